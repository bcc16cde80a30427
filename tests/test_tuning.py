"""Launch-shape tunables: the registry, the JSON cache loader, and that the
results of the tuned kernels do not depend on the chosen shape (the
transform arithmetic is the same whatever the lines per workgroup and the
workgroup size)."""
import json

import numpy as np
import pytest

from test_stages import INVALID, blobs


@pytest.fixture()
def fresh(mf):
    c = mf.Context(0, tuning=None)
    yield c
    c.close()


@pytest.mark.gpu
def test_gpu_tunable_registry(fresh, mf):
    assert fresh.tune_get("cc_rowsf/64x64x1/lines") is None
    fresh.tune_set("cc_rowsf/64x64x1/lines", 2)
    assert fresh.tune_get("cc_rowsf/64x64x1/lines") == 2
    fresh.tune_set("cc_rowsf/64x64x1/lines", 0)  # removes the key
    assert fresh.tune_get("cc_rowsf/64x64x1/lines") is None
    with pytest.raises(mf.MFError):
        fresh.tune_set("x", -1)


@pytest.mark.gpu
def test_gpu_tuning_file(fresh, mf, tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"cc_cols/96x80x1/lines": 2,
                             "corr_rw/96x80x1/threads": 128}))
    assert fresh.tune_load(p) == 2
    assert fresh.tune_get("cc_cols/96x80x1/lines") == 2
    assert fresh.tune_get("corr_rw/96x80x1/threads") == 128
    bad = tmp_path / "bad.json"
    for text in ('["a", 1]', '{"a": -3}', '{"a": "x"}', '{"a": 1'):
        bad.write_text(text)
        with pytest.raises(mf.MFError):
            fresh.tune_load(bad)
    with pytest.raises(mf.MFError):
        fresh.tune_load(tmp_path / "missing.json")


SHAPES = [{}, {"lines": 1, "threads": 64}, {"lines": 2, "threads": 128},
          {"lines": 4, "threads": 256}, {"lines": 3, "threads": 192}]


def _apply(c, prefix, shape):
    for k in ("lines", "threads"):
        c.tune_set(prefix + k, shape.get(k, 0))


@pytest.mark.gpu
@pytest.mark.parametrize("h,w", [(120, 144), (1024, 1024), (2048, 256)])
def test_gpu_pa_measure_independent_of_launch_shape(fresh, h, w):
    """Launch shapes change nothing; at power-of-two heights the default
    column shape takes the in-place column transforms, the others the
    ping-pong ones."""
    rng = np.random.default_rng(3)
    ref = blobs(h, w, seed=5)
    img = (np.roll(ref, (-3, 4), axis=(0, 1)) +
           0.1 * rng.standard_normal(ref.shape)).astype(np.float32)
    ref[:, :5] = INVALID
    d_r, d_i = fresh.array(ref), fresh.array(img)
    base = fresh.pa_measure(d_r, d_i, w, h)
    for fam in ("cc_rowsf", "cc_cols", "cc_rowsi"):
        for shape in SHAPES:
            _apply(fresh, f"{fam}/{w}x{h}x1/", shape)
            assert fresh.pa_measure(d_r, d_i, w, h) == base, (fam, shape)
        _apply(fresh, f"{fam}/{w}x{h}x1/", {})
    d_r.free()
    d_i.free()


@pytest.mark.gpu
def test_gpu_correct_rweight_independent_of_launch_shape(fresh):
    rng = np.random.default_rng(4)
    ny, nx = 96, 128
    raw = (rng.random((ny, nx)) + 0.5).astype(np.float32)
    d_raw, d_out = fresh.array(raw), fresh.empty((ny, nx))
    fresh.correct_proj(d_raw, nx, ny, 2.5, -1.5, 84.0, False, True, 1, True,
                       d_out)
    base = d_out.download()
    for shape in SHAPES:
        _apply(fresh, f"corr_rw/{nx}x{ny}x1/", shape)
        fresh.correct_proj(d_raw, nx, ny, 2.5, -1.5, 84.0, False, True, 1,
                           True, d_out)
        np.testing.assert_array_equal(d_out.download(), base)
    d_raw.free()
    d_out.free()


@pytest.mark.gpu
def test_gpu_malformed_workgroup_size_is_an_error(fresh, mf):
    """A cached workgroup size below one wave or not a multiple of 64 is
    rejected with a clear error, not launched."""
    h, w = 64, 96
    ref = blobs(h, w, seed=6)
    d_r, d_i = fresh.array(ref), fresh.array(ref)
    for bad in (32, 96 + 1, 512):
        fresh.tune_set(f"cc_rowsf/{w}x{h}x1/threads", bad)
        with pytest.raises(mf.MFError, match="workgroup size"):
            fresh.pa_measure(d_r, d_i, w, h)
    fresh.tune_set(f"cc_rowsf/{w}x{h}x1/threads", 0)
    sx, sy = fresh.pa_measure(d_r, d_i, w, h)
    assert abs(sx) < 0.05 and abs(sy) < 0.05
    d_r.free()
    d_i.free()
