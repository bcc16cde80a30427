"""Stack correction, stretched CC, projection-matching measurement.

CPU tests pin the oracle with closed forms (pixel-centre convention of the
nearest-neighbour correction, sum binning, the R-weighting filter as an FFT
product, shift recovery of the correlation); GPU tests compare the HIP
kernels with the oracle on the same inputs: correction bit-exact, correlated
shifts within 0.05 px.
"""
import numpy as np
import pytest

from conftest import rel_l2

INVALID = np.float32(-1e30)


def blobs(h, w, n=60, seed=0):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    img = np.zeros((h, w))
    for _ in range(n):
        cy, cx = rng.uniform(0, h), rng.uniform(0, w)
        s = rng.uniform(1.5, 5)
        img += rng.uniform(0.5, 1) * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2)
                                            / (2 * s * s))
    return (img + 0.1).astype(np.float32)


# ------------------------------------------------------------ oracle pins

def test_correct_integer_shift_convention(orc):
    """Shift-only correction maps out[y, i] = raw[y + sy + 1, i + sx + 1]:
    pixel centres at i + 0.5 and rounding by (int)(v + 0.5) (quirk Q11)."""
    raw = np.arange(40 * 32, dtype=np.float32).reshape(40, 32)
    out = orc.correct_proj(raw, 3.0, -2.0, 0.0, True, False, 1, False)
    assert out.shape == (40, 32)
    np.testing.assert_array_equal(out[5:30, 2:25], raw[4:29, 6:29])
    assert np.all(out[:, -3:] == INVALID)       # x + 4 >= 32 falls outside


def test_correct_rotation_half_turn(orc):
    raw = np.arange(24 * 16, dtype=np.float32).reshape(24, 16)
    # A quarter-pixel shift keeps q + 0.5 away from integers (a half turn
    # maps pixel centres exactly onto rounding boundaries otherwise).
    out = orc.correct_proj(raw, 0.25, 0.25, 180.0, False, False, 1, False)
    # q = -p + 0.25 + c  ->  index (int)(W + 0.25 - i) = W - i for i >= 1.
    np.testing.assert_array_equal(out[1:, 1:], raw[::-1, ::-1][:-1, :-1])


def test_correct_random_fill_is_nearby_raw_pixel(orc):
    raw = np.arange(64 * 64, dtype=np.float32).reshape(64, 64)
    out = orc.correct_proj(raw, 30.0, 0.0, 0.0, True, True, 1, False)
    # Columns whose source x = i + 31 >= 64 are filled from the mirrored
    # position (x -> 63 - x % 64) within the 51 x 51 window.
    for i in range(40, 64):
        for y in (5, 30, 60):
            v = int(out[y, i])
            sy_, sx_ = divmod(v, 64)
            x = i + 31
            xr = 63 - (x % 64)
            yr = y + 1 if y + 1 < 64 else 63 - ((y + 1) % 64)
            assert abs(sx_ - xr) <= 25 and abs(sy_ - yr) <= 25


def test_random_fill_keeps_the_window_bug_q10(orc):
    """mGRandom (CorrStack.cu:116-140) draws the x offset as next % 2601 - 25
    (quirk Q10: the window is 51 x 51 but x uses the whole LCG range) and the
    y offset as next / 51 - 25.  On an image wide enough to accept them, the
    fill pixels of out-of-image positions therefore lie up to ~2575 columns
    to the right of the mirrored source but never more than 25 rows away; a
    corrected window (next % 51) would keep both within 25.  The fill is a
    raw pixel, deterministic, and different positions draw different
    pixels."""
    w, h = 3200, 64
    raw = np.arange(w * h, dtype=np.float32).reshape(h, w)
    out = orc.correct_proj(raw, 2400.0, 0.0, 0.0, True, True, 1, False)
    again = orc.correct_proj(raw, 2400.0, 0.0, 0.0, True, True, 1, False)
    np.testing.assert_array_equal(out, again)
    dxs, dys = [], []
    for i in range(1000, w):
        for y in (3, 20, 40, 61):
            v = int(out[y, i])
            sy_, sx_ = divmod(v, w)
            # Nearest source pixel (int)(q + 0.5) of q = i + 0.5 + 2400
            # (quirk Q11), past the image's right edge, then mirrored.
            x = i + 2401
            xr = w - 1 - (x % w)
            yr = y + 1 if y + 1 < h else h - 1 - ((y + 1) % h)
            dxs.append(sx_ - xr)
            dys.append(sy_ - yr)
    dxs, dys = np.array(dxs), np.array(dys)
    assert np.abs(dys).max() <= 25
    assert dxs.min() >= -25 and dxs.max() > 1000
    assert len(np.unique(dxs)) > 100


def test_binning_sums(orc):
    raw = np.arange(20 * 16, dtype=np.float32).reshape(20, 16)
    out = orc.correct_proj(raw, -1.0, -1.0, 0.0, True, False, 2, False)
    # shift -1 cancels the +1 of the pixel-centre convention: identity.
    expect = raw.reshape(10, 2, 8, 2).sum(axis=(1, 3))
    np.testing.assert_array_equal(out, expect)
    # Any invalid pixel in a bin invalidates the bin sum.
    out2 = orc.correct_proj(raw, 0.0, -1.0, 0.0, True, False, 2, False)
    assert np.all(out2[:, -1] == INVALID)


def test_rweight_is_ramp_filter(orc):
    """R2C, x (1/N) ramp 2f(0.55+0.45cos 2pi f), C2R (unnormalised) equals
    irfft(rfft(x) * ramp) -- the filter of GRWeight (CorrStack.cu:431-446)."""
    rng = np.random.default_rng(3)
    raw = rng.standard_normal((12, 40)).astype(np.float32)
    out = orc.correct_proj(raw, -1.0, -1.0, 0.0, True, False, 1, True)
    n = 40
    f = np.arange(n // 2 + 1) / n
    ramp = 2 * f * (0.55 + 0.45 * np.cos(2 * np.pi * f))
    expect = np.fft.irfft(np.fft.rfft(raw.astype(np.float64), axis=1) * ramp,
                          n, axis=1)
    assert rel_l2(out, expect) < 1e-5


@pytest.mark.parametrize("dx,dy", [(5, -3), (-7, 4), (0, 0)])
def test_stretch_cc_recovers_shift(orc, dx, dy):
    ref = blobs(96, 128, seed=1)
    img = np.roll(np.roll(ref, dy, axis=0), dx, axis=1)
    sx, sy = orc.stretch_cc(ref, img, 0.0, 0.0, 0.0, 400.0)
    assert abs(sx - dx) < 0.3 and abs(sy - dy) < 0.3


def test_pa_measure_recovers_shift(orc):
    ref = blobs(128, 128, seed=2)
    img = np.roll(np.roll(ref, -4, axis=0), 6, axis=1)
    sx, sy = orc.pa_measure(ref, img)
    assert abs(sx - 6) < 0.3 and abs(sy + 4) < 0.3


def test_stretch_undoes_cosine_foreshortening(orc):
    """An image compressed along x by cos(t)/cos(t_ref) about the axis 0
    correlates best with its reference after stretching: zero shift."""
    ref = blobs(128, 128, seed=4)
    w = 128
    xs = (np.arange(w) + 0.5 - w / 2) / 0.8 + w / 2 - 0.5
    img = np.stack([np.interp(xs, np.arange(w), r) for r in ref]).astype(np.float32)
    ang = np.degrees(np.arccos(0.8))
    sx, sy = orc.stretch_cc(ref, img, 0.0, float(ang), 0.0, 400.0)
    assert abs(sx) < 0.5 and abs(sy) < 0.5


# --------------------------------------------------------------- GPU parity

@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    (3.5, -2.25, 0.0, True, True, 1, False),
    (-10.0, 7.0, 85.3, False, True, 1, False),
    (12.7, -3.3, 93.0, False, False, 4, False),
    (0.4, 0.6, 84.0, False, True, 1, True),
    (-5.5, 2.0, 86.0, False, True, 3, True),
])
def test_gpu_correct_matches_oracle(ctx, orc, args):
    raw = blobs(150, 182, seed=5)
    ref = orc.correct_proj(raw, *args)
    d_raw = ctx.array(raw)
    d_out = ctx.empty(ref.shape)
    ctx.correct_proj(d_raw, 182, 150, *args, d_out)
    got = d_out.download()
    if args[-1]:   # R-weighted: FP32 FFT against the FP64 oracle FFT
        assert rel_l2(got, ref) < 1e-5
    else:          # identical FP32 arithmetic: bit-exact
        np.testing.assert_array_equal(got, ref)
    d_raw.free()
    d_out.free()


@pytest.mark.gpu
@pytest.mark.parametrize("axis", [86.5, 93.0, 4.0])
def test_gpu_series_correction_bitexact(ctx, mf, orc, axis):
    """A series' final correction (bin 1, random fill): axes nearer a quarter
    turn gather from the transposed copy of the stack, the others from the
    stack itself; both equal the oracle's CorrectProj bit for bit."""
    nz = 5
    stack = np.stack([blobs(150, 182, seed=20 + z) for z in range(nz)])
    tilts = np.array([-8, -4, 0, 4, 8], np.float32)
    G = mf.Series(ctx, stack, tilts)
    G.preprocess()
    pre = G.stack()
    sx = np.array([3.25, -2.5, 0.0, 7.75, -11.0])
    sy = np.array([-1.5, 4.0, 2.25, -6.0, 0.5])
    G.set_params(sx, sy, np.full(nz, axis, np.float32))
    got = G.correct(1)
    for z in range(nz):
        ref = orc.correct_proj(pre[z], float(sx[z]), float(sy[z]), axis, False,
                               True, 1, False)
        np.testing.assert_array_equal(got[z], ref)
    G.close()


@pytest.mark.gpu
@pytest.mark.parametrize("tilts,axis", [((0.0, 0.0), 0.0), ((3.0, 6.0), 85.0),
                                        ((-30.0, -33.0), 93.0),
                                        ((50.0, 57.0), 80.0)])
def test_gpu_stretch_cc_matches_oracle(ctx, orc, tilts, axis):
    ref = blobs(128, 160, seed=6)
    img = np.roll(ref, (3, -5), axis=(0, 1)) + \
        0.05 * np.random.default_rng(7).standard_normal(ref.shape).astype(np.float32)
    img = img.astype(np.float32)
    o = orc.stretch_cc(ref, img, tilts[0], tilts[1], axis, 400.0)
    d_ref, d_img = ctx.array(ref), ctx.array(img)
    g = ctx.stretch_cc(d_ref, d_img, 160, 128, tilts[0], tilts[1], axis, 400.0)
    assert abs(g[0] - o[0]) <= 0.05 and abs(g[1] - o[1]) <= 0.05, (g, o)
    d_ref.free()
    d_img.free()


@pytest.mark.gpu
def test_gpu_pa_measure_matches_oracle(ctx, orc):
    rng = np.random.default_rng(8)
    ref = blobs(120, 144, seed=9)
    img = (np.roll(ref, (2, 5), axis=(0, 1)) +
           0.1 * rng.standard_normal(ref.shape)).astype(np.float32)
    ref2 = ref.copy()
    ref2[:, :6] = INVALID          # invalid reprojection columns
    o = orc.pa_measure(ref2, img)
    d_r, d_i = ctx.array(ref2), ctx.array(img)
    g = ctx.pa_measure(d_r, d_i, 144, 120)
    assert abs(g[0] - o[0]) <= 0.05 and abs(g[1] - o[1]) <= 0.05, (g, o)
    d_r.free()
    d_i.free()


# ------------------------------------------------------------ peak search

def _peak_maps():
    """Correlation-like maps: a smooth peak at a known sub-pixel place, a
    flat plateau (the first maximum in row-major order must win), a peak on
    the window's edge, and noise."""
    rng = np.random.default_rng(9)
    out = []
    for h, w, pitch in ((128, 96, 98), (100, 140, 142), (64, 64, 64)):
        yy, xx = np.mgrid[0:h, 0:w]
        m = np.exp(-((xx - w / 2 - 3.3) ** 2 + (yy - h / 2 + 5.7) ** 2) / 18.0)
        out.append((m, pitch))
        p = np.zeros((h, w))
        p[h // 2 - 2:h // 2 + 3, w // 2 - 4:w // 2 + 2] = 1.0
        out.append((p, pitch))
        e = rng.standard_normal((h, w)) * 0.01
        e[max(h // 20, 3) + 1, w - max(w // 20, 3) - 1] = 5.0
        e[(h - h * 8 // 20 * 2) // 2, (w - w * 8 // 20 * 2) // 2] = 4.0
        out.append((e, pitch))
    return out


def test_oracle_peak_recovers_parabola_vertex(orc):
    """A separable parabola sampled on the grid: the fit returns its vertex
    exactly in both windows (the three-point fit is exact for quadratics)."""
    h, w = 80, 96
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    vx, vy = w / 2 + 2.25, h / 2 - 3.75
    m = (100.0 - 0.5 * (xx - vx) ** 2 - 0.25 * (yy - vy) ** 2).astype(np.float32)
    for mode in (0, 1):
        sx, sy = orc.peak2d(m, mode)
        assert abs(sx - 2.25) < 1e-4 and abs(sy + 3.75) < 1e-4, (mode, sx, sy)


def test_oracle_peak_first_maximum_wins(orc):
    h, w = 64, 64
    m = np.zeros((h, w), np.float32)
    m[30, 40] = m[35, 20] = m[30, 45] = 2.0   # three equal maxima
    for mode in (0, 1):
        sx, sy = orc.peak2d(m, mode)
        # The fit is flat around an isolated spike: dx = dy = 0.
        assert (sx, sy) == (40 - 32, 30 - 32), (mode, sx, sy)


@pytest.mark.gpu
def test_gpu_peak2d_matches_oracle(ctx, orc):
    """mf_peak2d against the oracle's FindPeak / PAFindPeak: identical
    shifts (the same window, tie rule and fit arithmetic on the same
    floats)."""
    for m, pitch in _peak_maps():
        h, w = m.shape
        buf = np.zeros((h, pitch), np.float32)
        buf[:, :w] = m
        buf[:, w:] = 1e9   # pad columns are outside every window
        d = ctx.array(buf)
        for mode in (0, 1):
            want = orc.peak2d(buf, mode, w=w)
            got = ctx.peak2d(d, w, h, pitch, mode)
            assert got == want, (h, w, mode, got, want)
        d.free()
