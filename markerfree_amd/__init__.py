"""MI355X-native Markerfree alignment: Python mirror of the C-ABI.

The compute path is ``lib/libmarkerfree_hip.so`` (hand-written HIP for
gfx950).  This module only binds it with ctypes; it never falls back to a CPU
implementation.  Importing it when the library has not been built raises
``ImportError`` with the build command.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libmarkerfree_hip.so"
SYNTH_PATH = PKG / "lib" / "libmf_synth.so"
CLI_PATH = PKG / "bin" / "markerfree"

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -m markerfree_amd.build` "
        "(or __graft_entry__.build()) first")

_lib = C.CDLL(str(LIB_PATH))

MF_OK, MF_EINVAL, MF_ENOMEM, MF_EDEVICE, MF_EIO, MF_EBUSY, MF_ESTATE = range(7)
_STATUS = {0: "MF_OK", 1: "MF_EINVAL", 2: "MF_ENOMEM", 3: "MF_EDEVICE",
           4: "MF_EIO", 5: "MF_EBUSY", 6: "MF_ESTATE"}

_vp = C.c_void_p
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_ubyte)


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("mf_version", C.c_char_p)
_sig("mf_device_count", C.c_int)
_sig("mf_ctx_create", C.c_int, C.c_int, C.c_size_t, C.POINTER(_vp))
_sig("mf_ctx_destroy", None, _vp)
_sig("mf_last_error", C.c_char_p, _vp)
_sig("mf_ctx_synchronize", C.c_int, _vp)
_sig("mf_dev_malloc", C.c_int, _vp, C.c_size_t, C.POINTER(_vp))
_sig("mf_dev_free", C.c_int, _vp, _vp)
_sig("mf_h2d", C.c_int, _vp, _vp, _vp, C.c_size_t)
_sig("mf_d2h", C.c_int, _vp, _vp, _vp, C.c_size_t)
_sig("mf_d2d", C.c_int, _vp, _vp, _vp, C.c_size_t)
_sig("mf_timing_enable", C.c_int, _vp, C.c_int)
_sig("mf_timing_reset", C.c_int, _vp)
_sig("mf_kernel_time", C.c_int, _vp, C.c_char_p, C.POINTER(C.c_double),
     C.POINTER(C.c_longlong))
_sig("mf_find_proj_range", C.c_int, _fp, C.c_int, _u8p, C.c_int, C.c_int, _ip)
_sig("mf_reproject", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int,
     _fp, _u8p, C.c_int, C.c_int, C.c_int, _vp, _vp)
_sig("mf_fft_r2c_2d", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float)
_sig("mf_fft_c2r_2d", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int)
_sig("mf_fft_r2c_1d", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int)
_sig("mf_fft_c2r_1d", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int)


class MFError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


def lib():
    """The loaded ctypes library (for symbol checks)."""
    return _lib


def version() -> str:
    return _lib.mf_version().decode()


def device_count() -> int:
    return int(_lib.mf_device_count())


def f32p(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def u8p(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(_u8p)


def find_proj_range(tilts, skip, target, nproj):
    t = np.ascontiguousarray(tilts, np.float32)
    s = np.ascontiguousarray(skip, np.uint8)
    r = (C.c_int * 2)()
    st = _lib.mf_find_proj_range(f32p(t), len(t), u8p(s), target, nproj, r)
    if st:
        raise MFError(st, "mf_find_proj_range")
    return int(r[0]), int(r[1])


class DeviceArray:
    """A device buffer owned by a Context (plumbing for tests/benchmarks)."""

    def __init__(self, ctx: "Context", shape, dtype=np.float32):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        ctx._check(_lib.mf_dev_malloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        self.ctx._check(_lib.mf_h2d(self.ctx.h, self.ptr, a.ctypes.data,
                                    self.nbytes))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(_lib.mf_d2h(self.ctx.h, out.ctypes.data, self.ptr,
                                    self.nbytes))
        return out

    def free(self):
        if self.ptr:
            self.ctx._check(_lib.mf_dev_free(self.ctx.h, self.ptr))
            self.ptr = None


class Context:
    """mf_ctx: one device, one stream, an HBM budget and kernel timers."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = _vp()
        st = _lib.mf_ctx_create(device, hbm_budget, C.byref(h))
        if st:
            raise MFError(st, "mf_ctx_create failed (no usable HIP device?)")
        self.h = h.value

    def _check(self, st):
        if st:
            raise MFError(st, _lib.mf_last_error(self.h).decode())

    def close(self):
        if self.h:
            _lib.mf_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def synchronize(self):
        self._check(_lib.mf_ctx_synchronize(self.h))

    def array(self, a: np.ndarray) -> DeviceArray:
        return DeviceArray(self, a.shape, a.dtype).upload(a)

    def empty(self, shape, dtype=np.float32) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    # ---- timing
    def timing(self, on: bool = True):
        self._check(_lib.mf_timing_enable(self.h, int(on)))

    def timing_reset(self):
        self._check(_lib.mf_timing_reset(self.h))

    def kernel_time(self, name: str):
        ms = C.c_double()
        n = C.c_longlong()
        self._check(_lib.mf_kernel_time(self.h, name.encode(), C.byref(ms),
                                        C.byref(n)))
        return ms.value, n.value

    # ---- kernel-level operations
    def reproject(self, d_stack: DeviceArray, bx, by, nz, tb, tilts, skip,
                  target, nproj, d_out: DeviceArray, mode=0, d_slab=None):
        t = np.ascontiguousarray(tilts, np.float32)
        s = np.ascontiguousarray(skip, np.uint8)
        self._check(_lib.mf_reproject(
            self.h, d_stack.ptr, bx, by, nz, tb, f32p(t), u8p(s), target,
            nproj, mode, d_out.ptr, d_slab.ptr if d_slab is not None else None))

    def fft_r2c_2d(self, d: DeviceArray, nx, ny, nimg=1, scale=1.0):
        self._check(_lib.mf_fft_r2c_2d(self.h, d.ptr, nx, ny, nimg, scale))

    def fft_c2r_2d(self, d: DeviceArray, nx, ny, nimg=1):
        self._check(_lib.mf_fft_c2r_2d(self.h, d.ptr, nx, ny, nimg))

    def fft_r2c_1d(self, d: DeviceArray, n, nlines, pitch):
        self._check(_lib.mf_fft_r2c_1d(self.h, d.ptr, n, nlines, pitch))

    def fft_c2r_1d(self, d: DeviceArray, n, nlines, pitch):
        self._check(_lib.mf_fft_c2r_1d(self.h, d.ptr, n, nlines, pitch))
