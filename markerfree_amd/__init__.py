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
# MARKERFREE_HIP_LIB selects another build of the same library (e.g. a
# kernel variant under development); the default is the in-tree build.
LIB_PATH = Path(os.environ.get("MARKERFREE_HIP_LIB",
                               PKG / "lib" / "libmarkerfree_hip.so"))
SYNTH_PATH = PKG / "lib" / "libmf_synth.so"
# Launch shapes chosen by the autotuner (markerfree_amd.autotune) on MI355X.
DEFAULT_TUNING = PKG / "tuned" / "gfx950.json"
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
_sig("mf_ctx_memory", C.c_int, _vp, C.POINTER(C.c_size_t),
     C.POINTER(C.c_size_t), C.c_int)
_sig("mf_dev_free", C.c_int, _vp, _vp)
_sig("mf_h2d", C.c_int, _vp, _vp, _vp, C.c_size_t)
_sig("mf_d2h", C.c_int, _vp, _vp, _vp, C.c_size_t)
_sig("mf_d2d", C.c_int, _vp, _vp, _vp, C.c_size_t)
_sig("mf_timing_enable", C.c_int, _vp, C.c_int)
_sig("mf_timing_reset", C.c_int, _vp)
_sig("mf_timing_select", C.c_int, _vp, C.c_char_p)
_sig("mf_timing_every", C.c_int, _vp, C.c_int)
_sig("mf_kernel_count", C.c_int, _vp, C.c_char_p, C.POINTER(C.c_longlong))
_sig("mf_host_alloc", C.c_int, C.c_size_t, C.POINTER(_vp))
_sig("mf_host_free", C.c_int, _vp)
_sig("mf_tune_set", C.c_int, _vp, C.c_char_p, C.c_int)
_sig("mf_tune_get", C.c_int, _vp, C.c_char_p, C.POINTER(C.c_int))
_sig("mf_tune_load", C.c_int, _vp, C.c_char_p, C.POINTER(C.c_int))
_sig("mf_common_lines", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, _fp, _fp,
     _fp, C.c_float, C.c_float, C.c_int, _vp, C.POINTER(C.c_int))
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


def pinned_empty(shape, dtype=np.float32) -> np.ndarray:
    """A numpy array in page-locked host memory (mf_host_alloc), freed when
    the array and its views are collected; uploads from it and downloads into
    it run at the full PCIe rate."""
    import weakref
    dt = np.dtype(dtype)
    count = int(np.prod(shape))
    n = count * dt.itemsize
    p = _vp()
    st = _lib.mf_host_alloc(n, C.byref(p))
    if st:
        raise MFError(st, f"mf_host_alloc({n}) failed")
    buf = (C.c_char * max(n, 1)).from_address(p.value)
    weakref.finalize(buf, _lib.mf_host_free, p.value)
    return np.frombuffer(buf, dtype=dt, count=count).reshape(shape)


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

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device copy of 'other' (same size) on the context's
        stream."""
        assert other.nbytes == self.nbytes
        self.ctx._check(_lib.mf_d2d(self.ctx.h, self.ptr, other.ptr,
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

    def view(self, offset, shape) -> "DeviceView":
        """A part of this buffer, 'offset' elements in, of 'shape'; it does
        not own the memory and must not outlive the buffer."""
        n = int(np.prod(shape))
        if offset < 0 or (offset + n) * self.dtype.itemsize > self.nbytes:
            raise ValueError("view outside the buffer")
        return DeviceView(self, offset, shape)


class DeviceView(DeviceArray):
    """A non-owning window into a DeviceArray (see DeviceArray.view)."""

    def __init__(self, base: DeviceArray, offset, shape):
        self.ctx = base.ctx
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = base.dtype
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = base.ptr + int(offset) * self.dtype.itemsize

    def free(self):
        self.ptr = None


class Context:
    """mf_ctx: one device, one stream, an HBM budget and kernel timers."""

    def __init__(self, device: int = 0, hbm_budget: int = 0,
                 tuning: str | None = "default"):
        """'tuning': path of an autotuner JSON cache to apply, "default" for
        the committed cache of this GPU family (if present), None for the
        built-in launch-shape heuristics."""
        import weakref
        h = _vp()
        st = _lib.mf_ctx_create(device, hbm_budget, C.byref(h))
        if st:
            raise MFError(st, "mf_ctx_create failed (no usable HIP device?)")
        self.h = h.value
        # Series must be destroyed before their context (C-ABI rule).
        self._series = weakref.WeakSet()
        if tuning == "default":
            tuning = str(DEFAULT_TUNING) if DEFAULT_TUNING.exists() else None
        if tuning:
            self.tune_load(tuning)

    # ---- launch-shape tunables (markerfree_amd.autotune)
    def tune_set(self, key: str, value: int):
        self._check(_lib.mf_tune_set(self.h, key.encode(), int(value)))

    def tune_get(self, key: str):
        """Value of a tunable, or None when it is not set."""
        v = C.c_int()
        if _lib.mf_tune_get(self.h, key.encode(), C.byref(v)):
            return None
        return int(v.value)

    def tune_load(self, path) -> int:
        n = C.c_int()
        self._check(_lib.mf_tune_load(self.h, str(path).encode(), C.byref(n)))
        return int(n.value)

    def _check(self, st):
        if st:
            raise MFError(st, _lib.mf_last_error(self.h).decode())

    def close(self):
        if self.h:
            for s in list(self._series):
                s.close()
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

    def pinned_empty(self, shape, dtype=np.float32) -> np.ndarray:
        return pinned_empty(shape, dtype)

    # ---- timing
    def timing(self, on: bool = True, only: str | None = None, every: int = 1):
        """Kernel timing on or off; 'only' restricts it to one launch name,
        'every' times only every k-th launch of a name."""
        self._check(_lib.mf_timing_select(
            self.h, only.encode() if only else None))
        self._check(_lib.mf_timing_every(self.h, every))
        self._check(_lib.mf_timing_enable(self.h, int(on)))

    def kernel_count(self, name: str) -> int:
        """Launches of 'name' since timing_reset(), timed or not."""
        n = C.c_longlong()
        self._check(_lib.mf_kernel_count(self.h, name.encode(), C.byref(n)))
        return n.value

    def timing_reset(self):
        self._check(_lib.mf_timing_reset(self.h))

    def memory(self, reset_peak=False):
        """(bytes held now, most bytes held at once) by the context and its
        series; reset_peak restarts the peak from the current holding."""
        u, p = C.c_size_t(), C.c_size_t()
        self._check(_lib.mf_ctx_memory(self.h, C.byref(u), C.byref(p),
                                       int(reset_peak)))
        return u.value, p.value

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


# ------------------------------------------------------------ stage kernels
_dp = C.POINTER(C.c_double)
_sig("mf_correct_proj", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_float,
     C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, _vp)
_sig("mf_stretch_cc", C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, C.c_float,
     C.c_float, C.c_float, C.c_float, _fp)
_sig("mf_pa_measure", C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, _fp)
_sig("mf_series_upload", C.c_int, _vp, C.c_int, C.c_int, _fp)


class Options(C.Structure):
    """mf_options: the reference's -g / -p flags (main/opts.h:154-195)."""
    _fields_ = [("offset", C.c_float), ("pitch_angle", C.c_float),
                ("zshift", C.c_float), ("thickness", C.c_int),
                ("align_z", C.c_int), ("out_bin", C.c_int),
                ("nproj", C.c_int), ("verbose", C.c_int),
                ("reproj_mode", C.c_int), ("pa_pairs", C.c_int),
                ("graphs", C.c_int)]


_sig("mf_options_default", None, C.POINTER(Options))
_sig("mf_series_create", C.c_int, _vp, C.c_int, C.c_int, C.c_int, _fp, _fp,
     C.POINTER(_vp))
_sig("mf_series_destroy", None, _vp)
_sig("mf_series_last_error", C.c_char_p, _vp)
_sig("mf_series_set_verbose", C.c_int, _vp, C.c_int)
_sig("mf_series_set_graphs", C.c_int, _vp, C.c_int)
_sig("mf_series_graph_replays", C.c_int, _vp, C.POINTER(C.c_longlong))
_sig("mf_series_trim", C.c_int, _vp)
_sig("mf_preprocess", C.c_int, _vp)
_sig("mf_get_stack", C.c_int, _vp, _fp)
_sig("mf_get_tilts", C.c_int, _vp, _fp)
_sig("mf_set_tilts", C.c_int, _vp, _fp)
_sig("mf_get_params", C.c_int, _vp, _dp, _dp, _fp, _fp)
_sig("mf_set_params", C.c_int, _vp, _dp, _dp, _fp, C.c_float)
_sig("mf_coarse_setup", C.c_int, _vp)
_sig("mf_coarse_iter", C.c_int, _vp, _fp)
_sig("mf_tiltaxis", C.c_int, _vp, C.c_float, C.c_int, _fp)
_sig("mf_find_offset", C.c_int, _vp, _fp)
_sig("mf_apply_offset", C.c_int, _vp)
_sig("mf_projalign_setup", C.c_int, _vp, C.c_float, C.c_int, C.c_int, C.c_int)
_sig("mf_projalign_iter", C.c_int, _vp, _fp)
_sig("mf_projalign_once", C.c_int, _vp)
_sig("mf_correct", C.c_int, _vp, C.c_int, _fp, _ip, _ip)
_sig("mf_correct_async", C.c_int, _vp, C.c_int, _fp, _ip, _ip)
_sig("mf_series_wait", C.c_int, _vp)
_sig("mf_run", C.c_int, _vp, C.POINTER(Options))
_sig("mf_pa_measurements", C.c_int, _vp, _fp, _fp)
_sig("mf_axis_scores", C.c_int, _vp, _fp, C.c_int, _ip)
_sig("mf_offset_scores", C.c_int, _vp, _fp)
_sig("mf_series_trace", C.c_int, _vp, _ip, _fp, C.c_int, _ip)
_sig("mf_series_trace_clear", C.c_int, _vp)
_sig("mf_line_scores", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, _fp, _fp,
     _fp, C.c_float, C.c_float, C.c_int, _fp, _vp, _ip, _fp)
_sig("mf_ncc_moments", C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, C.c_float,
     C.c_float, _dp, _fp)
_sig("mf_offset_scan", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, _fp, _fp,
     _fp)
_sig("mf_peak2d", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _fp)
_sig("mf_peak2d_noalloc", C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int,
     C.c_int, _fp)

# Stage trace kinds (mf_api.h MF_EV_*).
EV_PREPROCESS, EV_COARSE, EV_AXIS, EV_OFFSET, EV_APPLY_OFFSET, EV_PA_SETUP, \
    EV_PA_ITER, EV_ONCE_BEGIN, EV_ONCE_END = range(1, 10)


def default_options(**kw) -> Options:
    o = Options()
    _lib.mf_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _ctx_correct_proj(self, d_raw, nx, ny, shift_x, shift_y, axis_deg,
                      shift_only, random_fill, binning, rweight, d_out):
    self._check(_lib.mf_correct_proj(self.h, d_raw.ptr, nx, ny, shift_x,
                                     shift_y, axis_deg, int(shift_only),
                                     int(random_fill), binning, int(rweight),
                                     d_out.ptr))


def _ctx_stretch_cc(self, d_ref, d_img, bx, by, ref_tilt, tilt, axis_deg,
                    bfactor=400.0):
    out = np.zeros(2, np.float32)
    self._check(_lib.mf_stretch_cc(self.h, d_ref.ptr, d_img.ptr, bx, by,
                                   ref_tilt, tilt, axis_deg, bfactor,
                                   f32p(out)))
    return float(out[0]), float(out[1])


def _ctx_pa_measure(self, d_reproj, d_proj, bx, by):
    out = np.zeros(2, np.float32)
    self._check(_lib.mf_pa_measure(self.h, d_reproj.ptr, d_proj.ptr, bx, by,
                                   f32p(out)))
    return float(out[0]), float(out[1])


def _ctx_common_lines(self, d_stack, nx, ny, nz, tilts, sx, sy, axis0,
                      arange, num=100):
    """GRadon lines [nz][num][L] (host array) of a resident stack."""
    a = [np.ascontiguousarray(v, np.float32) for v in (tilts, sx, sy)]
    L = C.c_int()
    self._check(_lib.mf_common_lines(self.h, d_stack.ptr, nx, ny, nz,
                                     *map(f32p, a), axis0, arange, num, None,
                                     C.byref(L)))
    out = self.empty((nz, num, L.value))
    try:
        self._check(_lib.mf_common_lines(self.h, d_stack.ptr, nx, ny, nz,
                                         *map(f32p, a), axis0, arange, num,
                                         out.ptr, C.byref(L)))
        return out.download()
    finally:
        out.free()


def _ctx_line_scores(self, d_stack, nx, ny, nz, tilts, sx, sy, axis0, arange,
                     num=100, spectra=False):
    """Common-line score of every axis candidate: (scores, best candidate
    before the wrap rules, L, spectra [nz][num][L/2+1] complex64 or None)."""
    a = [np.ascontiguousarray(v, np.float32) for v in (tilts, sx, sy)]
    scores = np.zeros(num, np.float32)
    L = C.c_int()
    best = C.c_float()
    d_spec = None
    if spectra:
        n = _line_len(self, d_stack, nx, ny, nz, a, axis0, arange, num)
        d_spec = self.empty((nz, num, n + 2))
    try:
        self._check(_lib.mf_line_scores(
            self.h, d_stack.ptr, nx, ny, nz, *map(f32p, a), axis0, arange,
            num, f32p(scores), d_spec.ptr if d_spec else None, C.byref(L),
            C.byref(best)))
        spec = None
        if d_spec is not None:
            spec = d_spec.download().view(np.complex64)
        return scores, float(best.value), int(L.value), spec
    finally:
        if d_spec is not None:
            d_spec.free()


def _line_len(ctx, d_stack, nx, ny, nz, a, axis0, arange, num):
    L = C.c_int()
    ctx._check(_lib.mf_common_lines(ctx.h, d_stack.ptr, nx, ny, nz,
                                    *map(f32p, a),
                                    axis0, arange, num, None, C.byref(L)))
    return int(L.value)


def _ctx_ncc_moments(self, d_ref, d_img, w, h, ref_tilt, tilt):
    """FindOffset::Correlate of one pair: (six moments, Q7 correlation)."""
    m = np.zeros(6, np.float64)
    cc = C.c_float()
    self._check(_lib.mf_ncc_moments(self.h, d_ref.ptr, d_img.ptr, w, h,
                                    ref_tilt, tilt, m.ctypes.data_as(_dp),
                                    C.byref(cc)))
    return m, float(cc.value)


def _ctx_offset_scan(self, d_corr, w, h, nz, tilts):
    """SearchOffset over a bin-4 corrected stack: (best, ccs[31])."""
    t = np.ascontiguousarray(tilts, np.float32)
    ccs = np.zeros(31, np.float32)
    best = C.c_float()
    self._check(_lib.mf_offset_scan(self.h, d_corr.ptr, w, h, nz, f32p(t),
                                    f32p(ccs), C.byref(best)))
    return float(best.value), ccs


def _ctx_peak2d(self, d_map, w, h, pitch, mode):
    out = np.zeros(2, np.float32)
    self._check(_lib.mf_peak2d(self.h, d_map.ptr, w, h, pitch, mode,
                               f32p(out)))
    return float(out[0]), float(out[1])


Context.line_scores = _ctx_line_scores
Context.ncc_moments = _ctx_ncc_moments
Context.offset_scan = _ctx_offset_scan
def _ctx_peak2d_noalloc(self, d_map, w, h, pitch, mode):
    """mf_peak2d_noalloc: the same search without any device allocation."""
    out = np.zeros(2, np.float32)
    self._check(_lib.mf_peak2d_noalloc(self.h, d_map.ptr, w, h, pitch, mode,
                                       f32p(out)))
    return float(out[0]), float(out[1])


Context.peak2d = _ctx_peak2d
Context.peak2d_noalloc = _ctx_peak2d_noalloc
Context.correct_proj = _ctx_correct_proj
Context.common_lines = _ctx_common_lines
Context.stretch_cc = _ctx_stretch_cc
Context.pa_measure = _ctx_pa_measure


class Series:
    """mf_series: one tilt series resident in HBM (Process counterpart)."""

    def __init__(self, ctx: Context, stack, tilts_file_order, shape=None):
        """stack [nz][ny][nx] in file order; or stack None and shape
        (nz, ny, nx), the images to follow through upload()."""
        if stack is None:
            self.nz, self.ny, self.nx = shape
            ptr = None
        else:
            stack = np.ascontiguousarray(stack, np.float32)
            self.nz, self.ny, self.nx = stack.shape
            ptr = f32p(stack)
        t = np.ascontiguousarray(tilts_file_order, np.float32)
        h = _vp()
        ctx._check(_lib.mf_series_create(ctx.h, self.nx, self.ny, self.nz,
                                         ptr, f32p(t), C.byref(h)))
        self.ctx = ctx
        self.h = h.value
        ctx._series.add(self)

    def _check(self, st):
        if st:
            raise MFError(st, _lib.mf_series_last_error(self.h).decode())

    def close(self):
        if self.h:
            _lib.mf_series_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def verbose(self, on=True):
        self._check(_lib.mf_series_set_verbose(self.h, int(on)))

    def upload(self, z0, images):
        """mf_series_upload: raw images [count][ny][nx] from z0 on (returns
        once the previous upload is done; 'images' must stay alive and
        unchanged until the next upload() call returns).  images None waits
        for the last upload."""
        if images is None:
            self._check(_lib.mf_series_upload(self.h, 0, 0, None))
            return
        assert images.dtype == np.float32 and images.flags.c_contiguous
        assert images.shape[1:] == (self.ny, self.nx)
        self._check(_lib.mf_series_upload(self.h, z0, images.shape[0],
                                          f32p(images)))

    def preprocess(self):
        self._check(_lib.mf_preprocess(self.h))

    def stack(self):
        out = np.empty((self.nz, self.ny, self.nx), np.float32)
        self._check(_lib.mf_get_stack(self.h, f32p(out)))
        return out

    def tilts(self):
        out = np.empty(self.nz, np.float32)
        self._check(_lib.mf_get_tilts(self.h, f32p(out)))
        return out

    def set_tilts(self, t):
        t = np.ascontiguousarray(t, np.float32)
        self._check(_lib.mf_set_tilts(self.h, f32p(t)))

    def params(self):
        sx = np.empty(self.nz, np.float64)
        sy = np.empty(self.nz, np.float64)
        rot = np.empty(self.nz, np.float32)
        off = C.c_float()
        self._check(_lib.mf_get_params(self.h, sx.ctypes.data_as(_dp),
                                       sy.ctypes.data_as(_dp), f32p(rot),
                                       C.byref(off)))
        return sx, sy, rot, float(off.value)

    def set_params(self, sx, sy, rot, offset=0.0):
        sx = np.ascontiguousarray(sx, np.float64)
        sy = np.ascontiguousarray(sy, np.float64)
        rot = np.ascontiguousarray(rot, np.float32)
        self._check(_lib.mf_set_params(self.h, sx.ctypes.data_as(_dp),
                                       sy.ctypes.data_as(_dp), f32p(rot),
                                       offset))

    def coarse_setup(self):
        self._check(_lib.mf_coarse_setup(self.h))

    def coarse_iter(self):
        e = C.c_float()
        self._check(_lib.mf_coarse_iter(self.h, C.byref(e)))
        return float(e.value)

    def tilt_axis(self, arange, num=100):
        a = C.c_float()
        self._check(_lib.mf_tiltaxis(self.h, arange, num, C.byref(a)))
        return float(a.value)

    def find_offset(self):
        b = C.c_float()
        self._check(_lib.mf_find_offset(self.h, C.byref(b)))
        return float(b.value)

    def apply_offset(self):
        self._check(_lib.mf_apply_offset(self.h))

    def projalign_setup(self, zshift, align_z, nproj, reproj_mode=8):
        self._check(_lib.mf_projalign_setup(self.h, zshift, align_z, nproj,
                                            reproj_mode))

    def projalign_iter(self):
        e = C.c_float()
        self._check(_lib.mf_projalign_iter(self.h, C.byref(e)))
        return float(e.value)

    def projalign_once(self):
        self._check(_lib.mf_projalign_once(self.h))

    def correct(self, out_bin=1, out=None, wait=True):
        """The aligned stack [nz][by][bx]; 'out', if given, receives it (for
        example a Context.pinned_empty array reused across series).  With
        wait=False the download into a page-locked 'out' is left running
        (mf_correct_async): 'out' holds the stack after wait()."""
        bx, by = C.c_int(), C.c_int()
        self._check(_lib.mf_correct(self.h, out_bin, None, C.byref(bx),
                                    C.byref(by)))
        shape = (self.nz, by.value, bx.value)
        if out is None:
            out = np.empty(shape, np.float32)
        if out.shape != shape or out.dtype != np.float32 or \
                not out.flags.c_contiguous:
            raise ValueError(f"out must be a contiguous float32 array of "
                             f"shape {shape}")
        fn = _lib.mf_correct if wait else _lib.mf_correct_async
        self._check(fn(self.h, out_bin, f32p(out), None, None))
        return out

    def wait(self):
        """Wait for the series' queued work (an unfinished download of
        correct(wait=False) included)."""
        self._check(_lib.mf_series_wait(self.h))

    def run(self, options: Options = None, **kw):
        o = options if options is not None else default_options(**kw)
        self._check(_lib.mf_run(self.h, C.byref(o)))

    def pa_measurements(self):
        """The last sweep's measured shifts [nz][2] (raw px) and lengths."""
        meas = np.zeros((self.nz, 2), np.float32)
        err = np.zeros(self.nz, np.float32)
        self._check(_lib.mf_pa_measurements(self.h, f32p(meas), f32p(err)))
        return meas, err

    def axis_scores(self):
        n = C.c_int()
        self._check(_lib.mf_axis_scores(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        self._check(_lib.mf_axis_scores(self.h, f32p(out), n.value, None))
        return out

    def offset_scores(self):
        out = np.zeros(31, np.float32)
        self._check(_lib.mf_offset_scores(self.h, f32p(out)))
        return out

    def trace(self):
        """Stage trace [(kind, value), ...] (EV_* kinds)."""
        n = C.c_int()
        self._check(_lib.mf_series_trace(self.h, None, None, 0, C.byref(n)))
        k = np.zeros(n.value, np.int32)
        v = np.zeros(n.value, np.float32)
        self._check(_lib.mf_series_trace(self.h, k.ctypes.data_as(_ip),
                                         f32p(v), n.value, None))
        return list(zip(k.tolist(), v.tolist()))

    def trace_clear(self):
        self._check(_lib.mf_series_trace_clear(self.h))

    def trim(self):
        """Release the stage workspaces; keep stack, tilts and parameters."""
        self._check(_lib.mf_series_trim(self.h))

    def set_graphs(self, on=True):
        """Replay projection-matching sweeps as a captured HIP graph."""
        self._check(_lib.mf_series_set_graphs(self.h, int(bool(on))))

    def graph_replays(self):
        n = C.c_longlong()
        self._check(_lib.mf_series_graph_replays(self.h, C.byref(n)))
        return n.value


# ------------------------------------------------ batched and served forms

_sig("mf_batch_run", C.c_int, C.POINTER(_vp), C.c_int, C.POINTER(Options),
     C.c_int)
_sig("mf_series_hbm_bytes", C.c_size_t, C.c_int, C.c_int, C.c_int,
     C.POINTER(Options))
_sig("mf_server_create", C.c_int, _vp, C.c_int, C.c_size_t, C.c_int,
     C.POINTER(_vp))
_sig("mf_server_destroy", None, _vp)
_sig("mf_submit", C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float),
     C.POINTER(C.c_float), C.POINTER(Options), C.POINTER(C.c_longlong))
_sig("mf_poll", C.c_int, _vp, C.c_longlong, C.POINTER(C.c_int))
_sig("mf_wait", C.c_int, _vp, C.c_longlong, C.POINTER(C.c_int))
_sig("mf_fetch", C.c_int, _vp, C.c_longlong, C.POINTER(C.c_double),
     C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_float),
     C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_int),
     C.POINTER(C.c_double), C.c_char_p, C.c_size_t)
_sig("mf_release", C.c_int, _vp, C.c_longlong)
_sig("mf_submit_borrowed", C.c_int, _vp, C.c_int, C.c_int, C.c_int, _fp, _fp,
     C.POINTER(Options), C.POINTER(C.c_longlong))
_sig("mf_server_clock", C.c_int, _vp, C.POINTER(C.c_double))
_sig("mf_server_stats", C.c_int, _vp, C.POINTER(C.c_size_t),
     C.POINTER(C.c_size_t), C.POINTER(C.c_int), C.POINTER(C.c_int),
     C.POINTER(C.c_longlong))

REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = range(4)


def series_hbm_bytes(nx, ny, nz, options: Options = None, **kw) -> int:
    o = options if options is not None else default_options(**kw)
    return int(_lib.mf_series_hbm_bytes(nx, ny, nz, C.byref(o)))


_BATCH_DONE = C.CFUNCTYPE(None, C.c_int, C.c_int, _fp, C.c_int, C.c_int,
                         C.c_int, _vp)
_sig("mf_batch_run_each", C.c_int, C.POINTER(_vp), C.c_int, C.POINTER(Options),
     C.c_int, C.c_int, _BATCH_DONE, _vp)


def batch_run(series, options: Options = None, concurrency=4, **kw):
    """mf_batch_run: mf_run on every series, 'concurrency' at a time."""
    o = options if options is not None else default_options(**kw)
    arr = (_vp * len(series))(*[s.h for s in series])
    st = _lib.mf_batch_run(arr, len(series), C.byref(o), concurrency)
    if st:
        msgs = [_lib.mf_series_last_error(s.h).decode() for s in series]
        raise MFError(st, "; ".join(m for m in msgs if m))


def batch_run_each(series, options, concurrency=4, correct=False,
                   on_done=None):
    """mf_batch_run_each: every series with its own Options, largest first;
    on_done(index, status, aligned) runs in the worker after each series,
    'aligned' a read-only view of its aligned stack (correct=True) valid only
    during the call, else None."""
    n = len(series)
    assert len(options) == n
    arr = (_vp * n)(*[s.h for s in series])
    opts = (Options * n)(*options)

    def _done(i, st, ptr, bx, by, nz, _user):
        if on_done is None:
            return
        a = None
        if ptr:
            a = np.ctypeslib.as_array(ptr, shape=(nz, by, bx))
            a.flags.writeable = False
        on_done(i, st, a)

    cb = _BATCH_DONE(_done)
    st = _lib.mf_batch_run_each(arr, n, opts, concurrency, int(bool(correct)),
                                cb, None)
    if st:
        msgs = [_lib.mf_series_last_error(s.h).decode() for s in series]
        raise MFError(st, "; ".join(m for m in msgs if m))


class Result:
    def __init__(self, shift_x, shift_y, rotate, offset, aligned, times):
        self.shift_x, self.shift_y, self.rotate = shift_x, shift_y, rotate
        self.offset, self.aligned = offset, aligned
        self.t_submit, self.t_start, self.t_done = times

    @property
    def latency(self):
        return self.t_done - self.t_submit


class Server:
    """mf_server: series submitted as they arrive, admitted in order against
    an HBM budget, run concurrently on 'workers' streams."""

    def __init__(self, ctx: Context, workers=2, hbm_budget=0, queue_cap=64):
        h = _vp()
        ctx._check(_lib.mf_server_create(ctx.h, workers, hbm_budget, queue_cap,
                                         C.byref(h)))
        self.ctx, self.h = ctx, h.value
        self._shapes = {}
        self._borrowed = {}
        ctx._series.add(self)

    def submit(self, stack, tilts, options: Options = None, borrow=False,
               **kw) -> int:
        """mf_submit (the server copies the stack), or with borrow=True
        mf_submit_borrowed (the server reads 'stack' in place; this object
        keeps a reference to it until the request is released)."""
        o = options if options is not None else default_options(**kw)
        if borrow:
            assert stack.dtype == np.float32 and stack.flags.c_contiguous
        stack = np.ascontiguousarray(stack, np.float32)
        t = np.ascontiguousarray(tilts, np.float32)
        nz, ny, nx = stack.shape
        rid = C.c_longlong()
        fn = _lib.mf_submit_borrowed if borrow else _lib.mf_submit
        st = fn(self.h, nx, ny, nz, f32p(stack), f32p(t), C.byref(o),
                C.byref(rid))
        if st:
            msg = "queue full" if st == MF_EBUSY else \
                _lib.mf_last_error(self.ctx.h).decode() or "submit failed"
            raise MFError(st, msg)
        self._shapes[rid.value] = nz
        if borrow:
            self._borrowed[rid.value] = stack
        return rid.value

    def poll(self, rid) -> int:
        s = C.c_int()
        self.ctx._check(_lib.mf_poll(self.h, rid, C.byref(s)))
        return s.value

    def wait(self, rid) -> int:
        s = C.c_int()
        self.ctx._check(_lib.mf_wait(self.h, rid, C.byref(s)))
        return s.value

    def fetch(self, rid, aligned=True) -> Result:
        nz = self._shapes[rid]
        sx, sy = np.empty(nz), np.empty(nz)
        rot = np.empty(nz, np.float32)
        off = C.c_float()
        bx, by = C.c_int(), C.c_int()
        times = np.empty(3)
        err = C.create_string_buffer(512)
        fp64 = C.POINTER(C.c_double)
        st = _lib.mf_fetch(self.h, rid, sx.ctypes.data_as(fp64),
                           sy.ctypes.data_as(fp64), f32p(rot), C.byref(off),
                           None, C.byref(bx), C.byref(by),
                           times.ctypes.data_as(fp64), err, 512)
        if st:
            raise MFError(st, err.value.decode() or "fetch failed")
        out = None
        if aligned:
            out = np.empty((nz, by.value, bx.value), np.float32)
            st = _lib.mf_fetch(self.h, rid, None, None, None, None, f32p(out),
                               None, None, None, err, 512)
            if st:
                raise MFError(st, err.value.decode())
        return Result(sx, sy, rot, off.value, out, tuple(times))

    def release(self, rid):
        self.ctx._check(_lib.mf_release(self.h, rid))
        self._shapes.pop(rid, None)
        self._borrowed.pop(rid, None)

    def clock(self) -> float:
        """Seconds since the server started (the clock of Result times)."""
        t = C.c_double()
        self.ctx._check(_lib.mf_server_clock(self.h, C.byref(t)))
        return t.value

    def stats(self):
        b, r = C.c_size_t(), C.c_size_t()
        q, n, c = C.c_int(), C.c_int(), C.c_longlong()
        self.ctx._check(_lib.mf_server_stats(self.h, C.byref(b), C.byref(r),
                                             C.byref(q), C.byref(n),
                                             C.byref(c)))
        return {"budget": b.value, "reserved": r.value, "queued": q.value,
                "running": n.value, "completed": c.value}

    def close(self):
        if self.h:
            _lib.mf_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
