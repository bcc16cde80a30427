"""The C-ABI library loads and exports every symbol its header declares."""
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared(header):
    text = Path(header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mf_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported(mf):
    names = declared(ROOT / "include" / "markerfree" / "mf_api.h")
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(mf.lib(), n)]
    assert not missing, missing


def test_version_and_errors_without_gpu(mf):
    assert mf.version().startswith("markerfree")
    # Host-only entry point: argument validation returns MF_EINVAL.
    import numpy as np
    t = np.zeros(3, np.float32)
    s = np.zeros(3, np.uint8)
    try:
        mf.find_proj_range(t, s, 5, 10)
    except mf.MFError as e:
        assert e.code == mf.MF_EINVAL
    else:
        raise AssertionError("out-of-range target accepted")


def test_empty_batches_need_no_gpu(mf):
    """An empty batch is a no-op in both batched entry points (no device is
    touched), and per-series options must match the series count."""
    mf.batch_run([], align_z=64, nproj=10)
    mf.batch_run_each([], [], concurrency=3, correct=True)
