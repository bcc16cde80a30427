"""Shared test setup: repo root on sys.path, the ``gpu`` marker, fixtures."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library on device 0)")


@pytest.fixture(scope="session")
def mf():
    import markerfree_amd
    return markerfree_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def ctx(mf):
    if mf.device_count() < 1:
        raise RuntimeError("GPU test selected but no HIP device is visible")
    c = mf.Context(0)
    yield c
    c.close()


def rel_l2(a, b, mask=None):
    cplx = np.iscomplexobj(a) or np.iscomplexobj(b)
    a = np.asarray(a, np.complex128 if cplx else np.float64)
    b = np.asarray(b, np.complex128 if cplx else np.float64)
    if mask is not None:
        a, b = a[mask], b[mask]
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


# Workloads of the benchmark configurations (seeded; the generator is
# deterministic across machines, checked by the stack checksums the
# end-to-end fixtures record).
@pytest.fixture(scope="session")
def cfg1():
    """512^2 x 41 float32, -60:3:60, axis 85.3, shifts U(-20,20), SNR 0.1
    (run with AlignZ 100)."""
    from markerfree_amd import synth
    return synth.generate(synth.PRESETS["small"])


@pytest.fixture(scope="session")
def cfg2():
    """1024^2 x 61 int16 counts, -60:2:60, axis U(80,95), shifts U(-40,40),
    SNR 0.05 (run with AlignZ 300)."""
    from markerfree_amd import synth
    return synth.generate(synth.PRESETS["single"])


@pytest.fixture(scope="session")
def cfg1_pre(orc, cfg1):
    """The oracle's preprocessed cfg1 stack and ascending tilts."""
    O = orc.Series(cfg1.stack, cfg1.tilts)
    O.preprocess()
    return O.stack(), O.tilts()


def flush_print(*a):
    """Progress line that reaches the log at once (long GPU tests)."""
    print(*a, flush=True)
