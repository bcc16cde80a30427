"""File formats of the drop-in command line: ctypes bindings of
``lib/libmf_io.so`` (MRC stacks, tilt-angle lists, .txt/.xf parameter files,
declared in ``include/markerfree/mf_io.h``) and a numpy MRC writer for
preparing inputs.

References: mrc/mrcheader.h:23-96, mrc/mrcstack.cpp:5-525,
main/Process.cpp:186-206 (ReadAngles), method/SaveParam.cpp:15-262.
"""
from __future__ import annotations

import ctypes as C
import struct
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libmf_io.so"
CLI_PATH = PKG / "bin" / "markerfree"

_lib = None


class MrcInfo(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("mode", C.c_int), ("next", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run "
                               "`python -m markerfree_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        f32p, f64p = C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.mfio_last_error.restype = C.c_char_p
        L.mfio_mrc_read_info.argtypes = [C.c_char_p, C.POINTER(MrcInfo)]
        L.mfio_mrc_read.argtypes = [C.c_char_p, f32p]
        L.mfio_mrc_read_slices.argtypes = [C.c_char_p, C.c_int, C.c_int,
                                           f32p]
        L.mfio_mrc_write.argtypes = [C.c_char_p, f32p, C.c_int, C.c_int,
                                     C.c_int, f32p]
        L.mfio_read_angles.argtypes = [C.c_char_p, f32p, C.c_int]
        L.mfio_param_path.argtypes = [C.c_char_p, C.c_int, C.c_char_p,
                                      C.c_size_t]
        L.mfio_save_txt.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int,
                                    f32p, f64p, f64p, f32p, C.c_float,
                                    C.c_int, C.c_float]
        L.mfio_save_xf.argtypes = [C.c_char_p, C.c_int, C.c_int, f32p, f64p,
                                   f64p, C.c_int]
        L.mfio_read_txt.argtypes = [C.c_char_p, C.c_int, f32p, f64p, f64p,
                                    f32p, f32p, C.POINTER(C.c_int)]
        L.mfio_fiducial_residuals.argtypes = [
            C.c_int, C.c_int, C.c_int, f32p, f64p, f64p, f32p, C.c_int, f32p,
            C.c_int, C.c_int, f64p, f64p, f64p]
        _lib = L
    return _lib


def _f32(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f64(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc):
    if rc != 0:
        raise OSError(lib().mfio_last_error().decode())


def mrc_info(path) -> MrcInfo:
    info = MrcInfo()
    _check(lib().mfio_mrc_read_info(str(path).encode(), C.byref(info)))
    return info


def read_mrc(path) -> np.ndarray:
    """The stack as float32 [nz][ny][nx]."""
    info = mrc_info(path)
    out = np.empty((info.nz, info.ny, info.nx), np.float32)
    _check(lib().mfio_mrc_read(str(path).encode(), _f32(out)))
    return out


def read_mrc_slices(path, z0, count) -> np.ndarray:
    """Images [z0, z0+count) as float32 [count][ny][nx]."""
    info = mrc_info(path)
    out = np.empty((max(count, 0), info.ny, info.nx), np.float32)
    _check(lib().mfio_mrc_read_slices(str(path).encode(), z0, count,
                                      _f32(out)))
    return out


def write_mrc_aligned(path, stack) -> np.ndarray:
    """The aligned-stack writer (float32, middle-slice statistics); returns
    (amin, amax, amean) as written to the header."""
    a = np.ascontiguousarray(stack, np.float32)
    nz, ny, nx = a.shape
    stats = np.empty(3, np.float32)
    _check(lib().mfio_mrc_write(str(path).encode(), _f32(a), nx, ny, nz,
                                _f32(stats)))
    return stats


def read_angles(path, cap=100000) -> np.ndarray:
    out = np.empty(cap, np.float32)
    n = lib().mfio_read_angles(str(path).encode(), _f32(out), cap)
    if n < 0:
        raise OSError(lib().mfio_last_error().decode())
    return out[:min(n, cap)].copy()


def param_path(outfile, mode) -> str:
    buf = C.create_string_buffer(4096)
    _check(lib().mfio_param_path(str(outfile).encode(), mode, buf, 4096))
    return buf.value.decode()


def save_txt(path, nx, ny, rotate, shift_x, shift_y, tilts, angle_offset,
             out_bin, elapsed):
    r = np.ascontiguousarray(rotate, np.float32)
    sx = np.ascontiguousarray(shift_x, np.float64)
    sy = np.ascontiguousarray(shift_y, np.float64)
    t = np.ascontiguousarray(tilts, np.float32)
    _check(lib().mfio_save_txt(str(path).encode(), nx, ny, len(r), _f32(r),
                               _f64(sx), _f64(sy), _f32(t), angle_offset,
                               out_bin, elapsed))


def save_xf(path, nx, rotate, shift_x, shift_y, inv):
    r = np.ascontiguousarray(rotate, np.float32)
    sx = np.ascontiguousarray(shift_x, np.float64)
    sy = np.ascontiguousarray(shift_y, np.float64)
    _check(lib().mfio_save_xf(str(path).encode(), nx, len(r), _f32(r),
                              _f64(sx), _f64(sy), int(inv)))


def read_txt(path, nz):
    """The .txt parameter file: (rotate, shift_x, shift_y, tilts,
    angle_offset, out_bin), rows in ascending tilt order."""
    rot = np.zeros(nz, np.float32)
    sx = np.zeros(nz, np.float64)
    sy = np.zeros(nz, np.float64)
    t = np.zeros(nz, np.float32)
    off = C.c_float()
    b = C.c_int()
    n = lib().mfio_read_txt(str(path).encode(), nz, _f32(rot), _f64(sx),
                            _f64(sy), _f32(t), C.byref(off), C.byref(b))
    if n < 0:
        raise OSError(f"cannot read {path}")
    if n != nz:
        raise ValueError(f"{path}: {n} rows, expected {nz}")
    return rot, sx, sy, t, float(off.value), int(b.value)


def fiducial_residuals(nx, ny, rotate, shift_x, shift_y, tilts, raw_xy,
                       min_views=2, fit_centre=True):
    """Reprojection residuals of marker tracks raw_xy [n][nz][2] (NaN where
    unseen) under an alignment (mfio_fiducial_residuals): returns a dict
    with mean, rms, max (px), observations, markers, centre offset, the
    per-section RMS and the markers' 3-D positions."""
    rot = np.ascontiguousarray(rotate, np.float32)
    nz = len(rot)
    sx = np.ascontiguousarray(shift_x, np.float64)
    sy = np.ascontiguousarray(shift_y, np.float64)
    t = np.ascontiguousarray(tilts, np.float32)
    xy = np.ascontiguousarray(raw_xy, np.float32)
    n = xy.shape[0]
    assert xy.shape == (n, nz, 2)
    per = np.zeros(nz, np.float64)
    xyz = np.zeros((n, 3), np.float64)
    summ = np.zeros(6, np.float64)
    st = lib().mfio_fiducial_residuals(nx, ny, nz, _f32(rot), _f64(sx),
                                       _f64(sy), _f32(t), n, _f32(xy),
                                       min_views, int(fit_centre), _f64(per),
                                       _f64(xyz), _f64(summ))
    if st:
        raise ValueError("bad fiducial input")
    return {"mean": summ[0], "rms": summ[1], "max": summ[2],
            "observations": int(summ[3]), "markers": int(summ[4]),
            "centre": summ[5], "per_section": per, "xyz": xyz}


# --------------------------------------------------------- numpy MRC writer

_MODES = {np.dtype(np.uint8): 0, np.dtype(np.int16): 1,
          np.dtype(np.float32): 2, np.dtype(np.uint16): 6}


def write_mrc(path, stack, next_bytes: bytes = b"") -> None:
    """Writes [nz][ny][nx] data as an MRC file (mode from the dtype: uint8 0,
    int16 1, float32 2, uint16 6) with an optional extended header, using
    the MRC2014 field offsets."""
    a = np.ascontiguousarray(stack)
    mode = _MODES[a.dtype]
    nz, ny, nx = a.shape
    h = bytearray(1024)
    struct.pack_into("<4i", h, 0, nx, ny, nz, mode)
    struct.pack_into("<3i", h, 28, nx, ny, nz)
    struct.pack_into("<6f", h, 40, nx, ny, nz, 90.0, 90.0, 90.0)
    struct.pack_into("<3i", h, 64, 1, 2, 3)
    f = a.astype(np.float64)
    struct.pack_into("<3f", h, 76, f.min(), f.max(), f.mean())
    struct.pack_into("<i", h, 92, len(next_bytes))
    h[208:212] = b"MAP "
    h[212:216] = bytes([0x44, 0x44, 0, 0])
    with open(path, "wb") as fh:
        fh.write(bytes(h))
        fh.write(next_bytes)
        fh.write(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes())


def read_header_fields(path) -> dict:
    """The header fields the writer sets, by MRC2014 offsets."""
    with open(path, "rb") as fh:
        h = fh.read(1024)
    nx, ny, nz, mode = struct.unpack_from("<4i", h, 0)
    return {
        "nx": nx, "ny": ny, "nz": nz, "mode": mode,
        "start": struct.unpack_from("<3i", h, 16),
        "mxyz": struct.unpack_from("<3i", h, 28),
        "cell": struct.unpack_from("<3f", h, 40),
        "angles": struct.unpack_from("<3f", h, 52),
        "map": struct.unpack_from("<3i", h, 64),
        "stats": struct.unpack_from("<3f", h, 76),
        "next": struct.unpack_from("<i", h, 92)[0],
        "creatid": struct.unpack_from("<h", h, 96)[0],
        "nlabl": struct.unpack_from("<i", h, 220)[0],
        "raw": h,
    }


def write_rawtlt(path, tilts) -> None:
    with open(path, "w") as fh:
        for t in tilts:
            fh.write(f"{float(t):.2f}\n")
