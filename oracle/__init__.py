"""CPU oracle for the Markerfree alignment path -- TEST INFRASTRUCTURE ONLY.

A C++ restatement of the reference algorithm (see ``oracle.h`` for the
function-by-function citations), used by ``tests/`` as the parity checker,
by ``__graft_entry__.smoke()`` as the checker of one small invocation, and by
``bench.py`` as the timed CPU baseline.  The product (markerfree_amd) never
imports or calls it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "lib" / "liborc.so"


def build():
    subprocess.run(["make", "-s", "-j8", "-C", str(HERE)], check=True)


if not LIB.exists():
    raise ImportError(f"{LIB} is not built; run `make -C oracle`")

_lib = C.CDLL(str(LIB))
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_ubyte)
_vp = C.c_void_p


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


def _f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def _u8(a):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(_u8p)


_sig("orc_find_proj_range", None, _fp, C.c_int, _u8p, C.c_int, C.c_int, _ip)
_sig("orc_reproject", None, _fp, C.c_int, C.c_int, C.c_int, C.c_int, _fp, _u8p,
     C.c_int, C.c_int, C.c_int, _fp, _fp)
_sig("orc_rp_inbounds_samples", C.c_longlong, C.c_int, C.c_int, C.c_float)
_sig("orc_fft_r2c_2d", None, _fp, C.c_int, C.c_int)
_sig("orc_fft_c2r_2d", None, _fp, C.c_int, C.c_int)
_sig("orc_fft_r2c_1d", None, _fp, C.c_int, C.c_int)
_sig("orc_fft_c2r_1d", None, _fp, C.c_int, C.c_int)
_sig("orc_num_threads", C.c_int)
_sig("orc_set_num_threads", None, C.c_int)


def num_threads() -> int:
    return int(_lib.orc_num_threads())


def set_num_threads(n: int):
    _lib.orc_set_num_threads(int(n))


def find_proj_range(tilts, skip, target, nproj):
    t = np.ascontiguousarray(tilts, np.float32)
    s = np.ascontiguousarray(skip, np.uint8)
    r = (C.c_int * 2)()
    _lib.orc_find_proj_range(_f(t), len(t), _u8(s), target, nproj, r)
    return int(r[0]), int(r[1])


def reproject(stack, tb, tilts, skip, target, nproj, acc64=True,
              want_slab=False):
    """One target of CalcReprojFBP::DoIt.  stack: [nz][by][bx] float32."""
    stack = np.ascontiguousarray(stack, np.float32)
    nz, by, bx = stack.shape
    t = np.ascontiguousarray(tilts, np.float32)
    s = np.ascontiguousarray(skip, np.uint8)
    out = np.empty((by, bx), np.float32)
    slab = np.empty((by, tb, 2 * bx), np.float32) if want_slab else None
    _lib.orc_reproject(_f(stack), bx, by, nz, tb, _f(t), _u8(s), target, nproj,
                       int(acc64), _f(out), _f(slab) if want_slab else None)
    return (out, slab) if want_slab else out


def rp_inbounds_samples(bx, tb, tilt):
    return int(_lib.orc_rp_inbounds_samples(bx, tb, float(tilt)))


def fft_r2c_2d(img):
    """cufftExecR2C of an ny x nx real image -> (ny, nx//2+1) complex64."""
    ny, nx = img.shape
    buf = np.zeros((ny, (nx // 2 + 1) * 2), np.float32)
    buf[:, :nx] = img
    _lib.orc_fft_r2c_2d(_f(buf), nx, ny)
    return buf.view(np.complex64)


def fft_c2r_2d(spec, nx):
    ny = spec.shape[0]
    buf = np.ascontiguousarray(spec, np.complex64).view(np.float32).copy()
    _lib.orc_fft_c2r_2d(_f(buf), nx, ny)
    return buf[:, :nx].copy()


# ------------------------------------------------------------ stage functions
_sig("orc_correct_proj", None, _fp, C.c_int, C.c_int, C.c_float, C.c_float,
     C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, _fp, _ip)
_sig("orc_binned_size", None, C.c_int, C.c_int, C.c_int, _ip)
_sig("orc_stretch_cc", None, _fp, _fp, C.c_int, C.c_int, C.c_float, C.c_float,
     C.c_float, C.c_float, _fp)
_sig("orc_pa_measure", None, _fp, _fp, C.c_int, C.c_int, _fp)
_sig("orc_tilt_axis", C.c_float, _fp, C.c_int, C.c_int, C.c_int, _fp, _fp, _fp,
     _fp, C.c_float, C.c_int, _fp, _ip)
_sig("orc_find_offset", C.c_float, _fp, C.c_int, C.c_int, C.c_int, _fp, _fp,
     _fp, _fp, _fp)


def binned_size(nx, ny, binning):
    r = (C.c_int * 2)()
    _lib.orc_binned_size(nx, ny, binning, r)
    return int(r[0]), int(r[1])


def correct_proj(raw, shift_x, shift_y, axis_deg, shift_only, random_fill,
                 binning, rweight):
    """CorrTomoStack::mCorrectProj for one raw image [ny][nx] -> [by][bx]."""
    raw = np.ascontiguousarray(raw, np.float32)
    ny, nx = raw.shape
    bx, by = binned_size(nx, ny, binning)
    out = np.empty((by, bx), np.float32)
    _lib.orc_correct_proj(_f(raw), nx, ny, shift_x, shift_y, axis_deg,
                          int(shift_only), int(random_fill), binning,
                          int(rweight), _f(out), None)
    return out


def stretch_cc(ref, img, ref_tilt, tilt, axis_deg, bfactor=400.0):
    ref = np.ascontiguousarray(ref, np.float32)
    img = np.ascontiguousarray(img, np.float32)
    by, bx = ref.shape
    out = np.zeros(2, np.float32)
    _lib.orc_stretch_cc(_f(ref), _f(img), bx, by, ref_tilt, tilt, axis_deg,
                        bfactor, _f(out))
    return float(out[0]), float(out[1])


def pa_measure(reproj, proj):
    reproj = np.ascontiguousarray(reproj, np.float32)
    proj = np.ascontiguousarray(proj, np.float32)
    by, bx = proj.shape
    out = np.zeros(2, np.float32)
    _lib.orc_pa_measure(_f(reproj), _f(proj), bx, by, _f(out))
    return float(out[0]), float(out[1])


def tilt_axis(stack, tilts, sx, sy, rot, arange, num=100):
    stack = np.ascontiguousarray(stack, np.float32)
    nz, ny, nx = stack.shape
    scores = np.zeros(num, np.float32)
    L = C.c_int()
    a = [np.ascontiguousarray(v, np.float32) for v in (tilts, sx, sy, rot)]
    axis = _lib.orc_tilt_axis(_f(stack), nx, ny, nz, *map(_f, a), arange, num,
                              _f(scores), C.byref(L))
    return float(axis), scores, int(L.value)


_sig("orc_common_lines", C.c_int, _fp, C.c_int, C.c_int, C.c_int, _fp, _fp,
     _fp, C.c_float, C.c_float, C.c_int, _fp)


def common_lines(stack, tilts, sx, sy, axis0, arange, num=100):
    """GRadon lines [nz][num][L] over the CLsetup grid around axis0."""
    stack = np.ascontiguousarray(stack, np.float32)
    nz, ny, nx = stack.shape
    a = [np.ascontiguousarray(v, np.float32) for v in (tilts, sx, sy)]
    L = _lib.orc_common_lines(_f(stack), nx, ny, nz, *map(_f, a), axis0,
                              arange, num, None)
    out = np.zeros((nz, num, L), np.float32)
    _lib.orc_common_lines(_f(stack), nx, ny, nz, *map(_f, a), axis0, arange,
                          num, _f(out))
    return out


_sig("orc_offset_correlate", C.c_float, _fp, _fp, C.c_int, C.c_int,
     C.c_float, C.c_float)


def offset_correlate(ref, img, ref_tilt, tilt):
    ref = np.ascontiguousarray(ref, np.float32)
    img = np.ascontiguousarray(img, np.float32)
    h, w = ref.shape
    return float(_lib.orc_offset_correlate(_f(ref), _f(img), w, h, ref_tilt,
                                           tilt))


_sig("orc_offset_moments", None, _fp, _fp, C.c_int, C.c_int, C.c_float,
     C.c_float, C.POINTER(C.c_double))
_sig("orc_offset_scan", C.c_float, _fp, C.c_int, C.c_int, C.c_int, _fp, _fp)
_sig("orc_peak2d", None, _fp, C.c_int, C.c_int, C.c_int, C.c_int, _fp)
_sig("orc_line_spectra", C.c_int, _fp, C.c_int, C.c_int, C.c_int, _fp, _fp,
     _fp, C.c_float, C.c_float, C.c_int, _fp)
_sig("orc_projalign_once_script", C.c_int, _fp, C.c_int)


def offset_moments(ref, img, ref_tilt, tilt):
    """The six masked moments of FindOffset::Correlate (float64[6])."""
    ref = np.ascontiguousarray(ref, np.float32)
    img = np.ascontiguousarray(img, np.float32)
    h, w = ref.shape
    m = np.zeros(6, np.float64)
    _lib.orc_offset_moments(_f(ref), _f(img), w, h, ref_tilt, tilt,
                            m.ctypes.data_as(C.POINTER(C.c_double)))
    return m


def offset_scan(corr, tilts):
    """SearchOffset over a bin-4 corrected stack [nz][h][w]: (best, ccs)."""
    corr = np.ascontiguousarray(corr, np.float32)
    nz, h, w = corr.shape
    t = np.ascontiguousarray(tilts, np.float32)
    ccs = np.zeros(31, np.float32)
    best = _lib.orc_offset_scan(_f(corr), w, h, nz, _f(t), _f(ccs))
    return float(best), ccs


def peak2d(cc_map, mode, w=None):
    """FindPeak (mode 0) / PAFindPeak (mode 1) on a map [h][pitch] whose
    first w columns are the map."""
    cc_map = np.ascontiguousarray(cc_map, np.float32)
    h, pitch = cc_map.shape
    w = pitch if w is None else w
    out = np.zeros(2, np.float32)
    _lib.orc_peak2d(_f(cc_map), w, h, pitch, mode, _f(out))
    return float(out[0]), float(out[1])


def line_spectra(stack, tilts, sx, sy, axis0, arange, num=100):
    """Radon + FFT1d line spectra [nz][num][L/2+1] complex64."""
    stack = np.ascontiguousarray(stack, np.float32)
    nz, ny, nx = stack.shape
    a = [np.ascontiguousarray(v, np.float32) for v in (tilts, sx, sy)]
    L = _lib.orc_line_spectra(_f(stack), nx, ny, nz, *map(_f, a), axis0,
                              arange, num, None)
    out = np.zeros((nz, num, L // 2 + 1), np.complex64)
    _lib.orc_line_spectra(_f(stack), nx, ny, nz, *map(_f, a), axis0, arange,
                          num, out.view(np.float32).ctypes.data_as(_fp))
    return out


def projalign_once_script(errs):
    """DoIt calls Process::ProjAlignOnce makes when its sweeps return errs."""
    e = np.ascontiguousarray(errs, np.float32)
    return int(_lib.orc_projalign_once_script(_f(e), len(e)))


def find_offset(stack, tilts, sx, sy, rot):
    stack = np.ascontiguousarray(stack, np.float32)
    nz, ny, nx = stack.shape
    ccs = np.zeros(31, np.float32)
    a = [np.ascontiguousarray(v, np.float32) for v in (tilts, sx, sy, rot)]
    best = _lib.orc_find_offset(_f(stack), nx, ny, nz, *map(_f, a), _f(ccs))
    return float(best), ccs


# ------------------------------------------------------------------ series
# Stage trace kinds (oracle.h ORC_EV_*).
EV_PREPROCESS, EV_COARSE, EV_AXIS, EV_OFFSET, EV_APPLY_OFFSET, EV_PA_SETUP, \
    EV_PA_ITER, EV_ONCE_BEGIN, EV_ONCE_END = range(1, 10)

class Options(C.Structure):
    _fields_ = [("offset", C.c_float), ("pitch_angle", C.c_float),
                ("zshift", C.c_float), ("thickness", C.c_int),
                ("align_z", C.c_int), ("out_bin", C.c_int), ("nproj", C.c_int),
                ("verbose", C.c_int)]


_sig("orc_series_create", _vp, C.c_int, C.c_int, C.c_int, _fp, _fp)
_sig("orc_series_destroy", None, _vp)
_sig("orc_preprocess", None, _vp)
_sig("orc_get_stack", None, _vp, _fp)
_sig("orc_get_tilts", None, _vp, _fp)
_sig("orc_set_tilts", None, _vp, _fp)
_sig("orc_get_params", None, _vp, _fp, _fp, _fp, _fp)
_sig("orc_set_params", None, _vp, _fp, _fp, _fp, C.c_float)
_sig("orc_coarse_setup", None, _vp)
_sig("orc_coarse_iter", C.c_float, _vp)
_sig("orc_series_tilt_axis", C.c_float, _vp, C.c_float, C.c_int)
_sig("orc_series_find_offset", C.c_float, _vp)
_sig("orc_apply_offset", None, _vp)
_sig("orc_projalign_setup", None, _vp, C.c_float, C.c_int, C.c_int)
_sig("orc_projalign_iter", C.c_float, _vp)
_sig("orc_projalign_iter_forced", C.c_float, _vp, _fp, _fp)
_sig("orc_series_trace", C.c_int, _vp, _ip, _fp, C.c_int)
_sig("orc_series_trace_clear", None, _vp)
_sig("orc_pa_z0", C.c_float, _vp)
_sig("orc_projalign_once", None, _vp)
_sig("orc_correct_final", None, _vp, C.c_int, _fp, _ip)
_sig("orc_run", None, _vp, C.POINTER(Options))
_sig("orc_get_corrected", None, _vp, _fp, _ip)


class Series:
    """orc_series: the reference pipeline state for one tilt series."""

    def __init__(self, stack, tilts_file_order):
        self.stack_raw = np.ascontiguousarray(stack, np.float32)
        self.nz, self.ny, self.nx = self.stack_raw.shape
        t = np.ascontiguousarray(tilts_file_order, np.float32)
        self.h = _lib.orc_series_create(self.nx, self.ny, self.nz,
                                        _f(self.stack_raw), _f(t))

    def __del__(self):
        if getattr(self, "h", None):
            _lib.orc_series_destroy(self.h)
            self.h = None

    def preprocess(self):
        _lib.orc_preprocess(self.h)

    def stack(self):
        out = np.empty((self.nz, self.ny, self.nx), np.float32)
        _lib.orc_get_stack(self.h, _f(out))
        return out

    def tilts(self):
        out = np.empty(self.nz, np.float32)
        _lib.orc_get_tilts(self.h, _f(out))
        return out

    def set_tilts(self, t):
        t = np.ascontiguousarray(t, np.float32)
        _lib.orc_set_tilts(self.h, _f(t))

    def params(self):
        sx, sy, rot = (np.empty(self.nz, np.float32) for _ in range(3))
        off = C.c_float()
        _lib.orc_get_params(self.h, _f(sx), _f(sy), _f(rot), C.byref(off))
        return sx, sy, rot, float(off.value)

    def set_params(self, sx, sy, rot, offset=0.0):
        a = [np.ascontiguousarray(v, np.float32) for v in (sx, sy, rot)]
        _lib.orc_set_params(self.h, *map(_f, a), offset)

    def coarse_setup(self):
        _lib.orc_coarse_setup(self.h)

    def coarse_iter(self):
        return float(_lib.orc_coarse_iter(self.h))

    def tilt_axis(self, arange, num=100):
        return float(_lib.orc_series_tilt_axis(self.h, arange, num))

    def find_offset(self):
        return float(_lib.orc_series_find_offset(self.h))

    def pa_z0(self):
        return float(_lib.orc_pa_z0(self.h))

    def apply_offset(self):
        _lib.orc_apply_offset(self.h)

    def projalign_setup(self, zshift, align_z, nproj):
        _lib.orc_projalign_setup(self.h, zshift, align_z, nproj)

    def projalign_iter(self):
        return float(_lib.orc_projalign_iter(self.h))

    def projalign_iter_forced(self, forced=None):
        """One sweep; 'forced' ([nz][2] raw px) replaces each measured shift
        before it is applied.  Returns (maxerr, own measurements [nz][2])."""
        own = np.zeros((self.nz, 2), np.float32)
        f = None
        if forced is not None:
            forced = np.ascontiguousarray(forced, np.float32).reshape(self.nz, 2)
            f = _f(forced)
        e = _lib.orc_projalign_iter_forced(self.h, f, _f(own))
        return float(e), own

    def projalign_once(self):
        _lib.orc_projalign_once(self.h)

    def trace(self):
        """Stage trace [(kind, value), ...] (ORC_EV_* kinds)."""
        n = _lib.orc_series_trace(self.h, None, None, 0)
        k = np.zeros(n, np.int32)
        v = np.zeros(n, np.float32)
        _lib.orc_series_trace(self.h, k.ctypes.data_as(_ip), _f(v), n)
        return list(zip(k.tolist(), v.tolist()))

    def trace_clear(self):
        _lib.orc_series_trace_clear(self.h)

    def correct_final(self, out_bin=1):
        bx, by = binned_size(self.nx, self.ny, out_bin)
        out = np.empty((self.nz, by, bx), np.float32)
        _lib.orc_correct_final(self.h, out_bin, _f(out), None)
        return out

    def run(self, align_z, nproj=10, offset=0.0, pitch_angle=0.0, zshift=0.0,
            out_bin=1, verbose=False):
        o = Options(offset, pitch_angle, zshift, 0, align_z, out_bin, nproj,
                    int(verbose))
        _lib.orc_run(self.h, C.byref(o))
