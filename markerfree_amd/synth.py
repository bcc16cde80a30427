"""Seeded synthetic tilt series (phantoms of 3-D Gaussian blobs).

Binds ``lib/libmf_synth.so`` (host C++).  The presets follow the benchmark
workloads: ``small`` = 512x512x41 float32 series (axis 85.3 deg, shifts
U(-20,20), SNR 0.1), ``single`` = 1024x1024x61 int16-count series (axis
U(80,95), shifts U(-40,40), SNR 0.05), ``large`` = 5760x4092x61 Poisson series.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import SYNTH_PATH


class _Cfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("vol_x", C.c_int), ("vol_y", C.c_int), ("vol_z", C.c_int),
        ("nblobs", C.c_int),
        ("sigma_min", C.c_float), ("sigma_max", C.c_float),
        ("amp_min", C.c_float), ("amp_max", C.c_float),
        ("tilt_min", C.c_float), ("tilt_step", C.c_float),
        ("axis_min", C.c_float), ("axis_max", C.c_float),
        ("shift_max", C.c_float), ("mag_jitter", C.c_float),
        ("noise_kind", C.c_int), ("snr", C.c_float), ("dose", C.c_float),
        ("int16_counts", C.c_int),
        ("seed_phantom", C.c_uint64), ("seed_geom", C.c_uint64),
        ("seed_noise", C.c_uint64),
    ]


_lib = None


def _get():
    global _lib
    if _lib is None:
        if not SYNTH_PATH.exists():
            raise ImportError(f"{SYNTH_PATH} is not built")
        _lib = C.CDLL(str(SYNTH_PATH))
        _lib.mfs_generate.restype = C.c_int
        fp = C.POINTER(C.c_float)
        _lib.mfs_generate.argtypes = [C.POINTER(_Cfg), fp, fp, fp, fp, fp, fp]
        _lib.mfs_fiducials.restype = C.c_int
        _lib.mfs_fiducials.argtypes = [C.POINTER(_Cfg), C.c_int, fp]
    return _lib


@dataclass
class SeriesConfig:
    nx: int = 512
    ny: int = 512
    nz: int = 41
    vol: tuple = (512, 512, 128)
    nblobs: int = 400
    sigma: tuple = (2.0, 8.0)
    amp: tuple = (0.5, 1.0)
    tilt_min: float = -60.0
    tilt_step: float = 3.0
    axis: tuple = (85.3, 85.3)
    shift_max: float = 20.0
    mag_jitter: float = 0.0
    noise: str = "gauss"        # "none" | "gauss" | "poisson"
    snr: float = 0.1
    dose: float = 20.0
    int16: bool = False
    seeds: tuple = (1, 2, 3)


PRESETS = {
    # 512^2 x 41 float32, -60:3:60, axis 85.3, shifts U(-20,20), SNR 0.1.
    "small": SeriesConfig(),
    # 1024^2 x 61 int16 counts, -60:2:60, 3000 blobs in 1024x1024x300,
    # axis U(80,95), shifts U(-40,40), SNR 0.05.
    "single": SeriesConfig(nx=1024, ny=1024, nz=61, vol=(1024, 1024, 300),
                           nblobs=3000, tilt_step=2.0, axis=(80.0, 95.0),
                           shift_max=40.0, snr=0.05, int16=True),
    # 5760x4092 x 61, axis 86, shifts U(-80,80), 0.5 % magnification jitter,
    # Poisson at 20 e-/px.
    "large": SeriesConfig(nx=5760, ny=4092, nz=61, vol=(5760, 4092, 300),
                          nblobs=12000, tilt_step=2.0, axis=(86.0, 86.0),
                          shift_max=80.0, mag_jitter=0.005, noise="poisson"),
}


@dataclass
class Series:
    stack: np.ndarray          # [nz][ny][nx] float32 (as read from the MRC)
    tilts: np.ndarray          # [nz] float32, ascending
    shift_x: np.ndarray        # true shifts (px) applied per tilt
    shift_y: np.ndarray
    axis: float                # true tilt axis (deg, reference convention)
    count_scale: float = 1.0   # int16 quantisation factor (1 if float)
    config: SeriesConfig = field(default_factory=SeriesConfig)


def _cfg(cfg: SeriesConfig) -> _Cfg:
    c = _Cfg()
    c.nx, c.ny, c.nz = cfg.nx, cfg.ny, cfg.nz
    c.vol_x, c.vol_y, c.vol_z = cfg.vol
    c.nblobs = cfg.nblobs
    c.sigma_min, c.sigma_max = cfg.sigma
    c.amp_min, c.amp_max = cfg.amp
    c.tilt_min, c.tilt_step = cfg.tilt_min, cfg.tilt_step
    c.axis_min, c.axis_max = cfg.axis
    c.shift_max, c.mag_jitter = cfg.shift_max, cfg.mag_jitter
    c.noise_kind = {"none": 0, "gauss": 1, "poisson": 2}[cfg.noise]
    c.snr, c.dose = cfg.snr, cfg.dose
    c.int16_counts = int(cfg.int16)
    c.seed_phantom, c.seed_geom, c.seed_noise = cfg.seeds
    return c


def fiducials(cfg: SeriesConfig, count: int) -> np.ndarray:
    """Raw-image positions of the first 'count' blob centres in every
    section, [count][nz][2] (NaN outside the image): point-like markers with
    known 3-D positions for the fiducial residual evaluator."""
    out = np.empty((count, cfg.nz, 2), np.float32)
    n = _get().mfs_fiducials(C.byref(_cfg(cfg)), count,
                             out.ctypes.data_as(C.POINTER(C.c_float)))
    return out[:n]


def generate(cfg: SeriesConfig) -> Series:
    c = _cfg(cfg)
    stack = np.empty((cfg.nz, cfg.ny, cfg.nx), np.float32)
    tilts = np.empty(cfg.nz, np.float32)
    sx = np.empty(cfg.nz, np.float32)
    sy = np.empty(cfg.nz, np.float32)
    axis = C.c_float()
    scale = C.c_float()
    fp = C.POINTER(C.c_float)
    st = _get().mfs_generate(
        C.byref(c), stack.ctypes.data_as(fp), tilts.ctypes.data_as(fp),
        sx.ctypes.data_as(fp), sy.ctypes.data_as(fp), C.byref(axis),
        C.byref(scale))
    if st:
        raise ValueError("mfs_generate rejected the configuration")
    return Series(stack, tilts, sx, sy, float(axis.value), float(scale.value),
                  cfg)


def write_series(series: Series, directory, name: str = "series",
                 descending: bool = False):
    """Writes the series as the command line reads it: an MRC stack (mode 1
    int16 for count data, else mode 2 float32) and a .rawtlt list, in file
    order (reversed when 'descending').  Returns (stack_path, tilt_path)."""
    from pathlib import Path

    from . import fileio

    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    stack, tilts = series.stack, series.tilts
    if descending:
        stack, tilts = stack[::-1], tilts[::-1]
    data = stack.astype(np.int16) if series.config.int16 else stack
    mrc, tlt = d / f"{name}.mrc", d / f"{name}.rawtlt"
    fileio.write_mrc(mrc, np.ascontiguousarray(data))
    fileio.write_rawtlt(tlt, tilts)
    return mrc, tlt
