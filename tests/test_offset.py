"""Tilt-angle offset search (FindOffset.cu).

CPU tests pin the oracle's 31-step scan: a series whose tilt file is off by
a known angle comes back with the opposite offset; the scan curve moves with
the offset; each step's mean is the mean of the neighbour correlations
toward the zero tilt (quirk Q33), each correlation pinned in closed form by
test_oracle_pipeline.py (quirk Q7).  GPU tests compare the six masked
moments, the Q7 correlation and the whole scan with the oracle on the same
bin-4 corrected stacks.
"""
import numpy as np
import pytest


def _clean_series(nz=41, step=3.0):
    from markerfree_amd import synth
    cfg = synth.SeriesConfig(nx=512, ny=512, nz=nz, vol=(512, 512, 128),
                             nblobs=600, tilt_step=step,
                             tilt_min=-(nz - 1) / 2 * step, axis=(0.0, 0.0),
                             shift_max=0.0, noise="none")
    return synth.generate(cfg)


@pytest.fixture(scope="module")
def clean(orc):
    s = _clean_series()
    O = orc.Series(s.stack, s.tilts)
    O.preprocess()
    return O.stack(), O.tilts()


def _bin4(orc, stack):
    return np.stack([orc.correct_proj(im, 0.0, 0.0, 0.0, False, False, 4,
                                      False) for im in stack])


def test_oracle_offset_scan_recovers_known_offset(orc, clean):
    """Tilts read d degrees too high come back as an offset of -d (the
    scan adds its step to every tilt, FindOffset.cu:150-153), and the
    scan curve is the d = 0 curve moved by d steps."""
    stack, tilts = clean
    z = np.zeros(len(tilts), np.float32)
    b0, c0 = orc.find_offset(stack, tilts, z, z, z)
    assert b0 == 0.0
    for d in (3, -5, 7):
        b, c = orc.find_offset(stack, tilts + np.float32(d), z, z, z)
        assert b == -d
        lo, hi = max(0, d), min(31, 31 + d)
        np.testing.assert_allclose(c[lo - d:hi - d], c0[lo:hi], rtol=1e-6)


def test_oracle_offset_step_is_mean_of_neighbour_correlations(orc, clean):
    """One step of CalcAverangedCC (FindOffset.cu:148-174): every image but
    the zero tilt is correlated with its neighbour toward the zero tilt
    (quirk Q33; FindRefIndex's lower-|tilt| rule is not used here), and the
    step's value is their float mean."""
    stack, tilts = clean
    corr = _bin4(orc, stack)
    # Shift the tilts so that the zero tilt sits off the middle: the two
    # neighbour rules then pick different references for some images.
    t = (tilts + np.float32(4.0)).astype(np.float32)
    best, ccs = orc.offset_scan(corr, t)
    step = 15 - 2   # offset -2: tilts + 2, zero tilt where t - 2 ~ 0
    ang = (t + np.float32(-2.0)).astype(np.float32)
    z0 = int(np.argmin(np.abs(ang)))
    vals = []
    for k in range(len(ang)):
        if k == z0:
            continue
        r = k + 1 if k < z0 else k - 1
        vals.append(orc.offset_correlate(corr[r], corr[k], ang[r], ang[k]))
    mean = np.float32(0)
    for v in vals:
        mean = np.float32(mean + np.float32(v))
    mean = np.float32(mean / np.float32(len(vals)))
    assert ccs[step] == mean
    assert best == -4.0


@pytest.mark.gpu
def test_gpu_ncc_moments_match_oracle(ctx, orc, cfg1_pre):
    """The six masked moments and the Q7 correlation of neighbouring bin-4
    images of the 512^2 x 41 workload (FindOffset.cu:176-238): moments
    within 1e-9 relative (FP64 sums of the same FP32 terms in another
    order), correlation within 1e-6."""
    stack, tilts = cfg1_pre
    corr = _bin4(orc, stack[18:23])
    h, w = corr.shape[1:]
    d = ctx.array(corr)
    plane = h * w
    for k, r in ((0, 1), (4, 3), (2, 1)):
        mo = orc.offset_moments(corr[r], corr[k], float(tilts[18 + r]),
                                float(tilts[18 + k]))
        co = orc.offset_correlate(corr[r], corr[k], float(tilts[18 + r]),
                                  float(tilts[18 + k]))
        ref_v = d.view(r * plane, (h, w))
        img_v = d.view(k * plane, (h, w))
        mg, cg = ctx.ncc_moments(ref_v, img_v, w, h, float(tilts[18 + r]),
                                 float(tilts[18 + k]))
        assert mg[5] == mo[5]           # the same pixels are counted
        np.testing.assert_allclose(mg, mo, rtol=1e-9)
        assert abs(cg - co) <= 1e-6, (cg, co)
    d.free()


@pytest.mark.gpu
@pytest.mark.parametrize("shift", [0.0, 4.0])
def test_gpu_offset_scan_matches_oracle(ctx, orc, cfg1_pre, shift):
    """The whole 31-step scan (FindOffset.cu:129-174) on the bin-4 corrected
    512^2 x 41 workload: every step's mean correlation within 1e-5 of the
    oracle's, the same best step."""
    stack, tilts = cfg1_pre
    corr = _bin4(orc, stack)
    t = (tilts + np.float32(shift)).astype(np.float32)
    bo, co = orc.offset_scan(corr, t)
    nz, h, w = corr.shape
    d = ctx.array(corr)
    bg, cg = ctx.offset_scan(d, w, h, nz, t)
    d.free()
    assert float(np.abs(cg - co).max()) <= 1e-5 * float(np.abs(co).max())
    assert bg == bo
