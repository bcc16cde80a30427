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
