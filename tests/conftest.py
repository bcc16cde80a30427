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
