"""mf_peak2d_noalloc: the allocation-free peak search gives mf_peak2d's
results, on two calls in a row (the second on a smaller map, so that the
partials the first call left in the shared array must not leak into it), and
rejects maps beyond its maximum."""
import numpy as np
import pytest


def _map(h, w, pitch, peak, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    m = np.exp(-((xx - peak[0]) ** 2 + (yy - peak[1]) ** 2) / 14.0)
    m = m + 0.01 * rng.standard_normal((h, w))
    buf = np.zeros((h, pitch), np.float32)
    buf[:, :w] = m
    return buf


@pytest.mark.gpu
def test_gpu_peak2d_noalloc_matches_allocating_entry(ctx, mf, orc):
    # Large map first with its peak near the bottom, then a smaller one
    # whose peak lies elsewhere: a stale partial of the first call would
    # win the second call's argmax.
    cases = [(2048, 1536, 1538, (900.3, 1800.6), 1),
             (300, 260, 262, (100.2, 40.7), 2)]
    for mode in (0, 1):
        for h, w, pitch, peak, seed in cases:
            buf = _map(h, w, pitch, peak, seed + 10 * mode)
            d = ctx.array(buf)
            want = ctx.peak2d(d, w, h, pitch, mode)
            got = ctx.peak2d_noalloc(d, w, h, pitch, mode)
            d.free()
            assert got == want, (mode, h, got, want)
            assert got == orc.peak2d(buf, mode, w=w)


@pytest.mark.gpu
def test_gpu_peak2d_noalloc_rejects_oversized_map(ctx, mf):
    d = ctx.empty((8, 8))   # never read: the size check comes first
    with pytest.raises(mf.MFError, match="32768 rows"):
        ctx.peak2d_noalloc(d, 64, 32769, 64, 0)
    d.free()
