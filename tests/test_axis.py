"""Common-line (Radon) sums of the tilt-axis search: the oracle pinned on
images whose lines are known in closed form, and the HIP lines bit-identical
to the oracle's, for candidate grids whose lines run along image rows, along
columns (read through the transposed stack) and across the image edges."""
import numpy as np
import pytest

INVALID = np.float32(-1e30)


def test_oracle_lines_of_constant_image(orc):
    """A constant image: every line that meets the image has the constant as
    its mean; lines that meet no pixel are invalid."""
    nz, ny, nx = 2, 260, 300
    stack = np.full((nz, ny, nx), 2.5, np.float32)
    tilts = np.array([0.0, 30.0], np.float32)
    z = np.zeros(nz, np.float32)
    lines = orc.common_lines(stack, tilts, z, z, 80.0, 180.0, num=12)
    valid = lines > -1e10
    assert valid.any() and np.all(lines[valid] == np.float32(2.5))
    # The cos(tilt) crop shortens lines, it never adds pixels.
    assert valid[1].sum() <= valid[0].sum()


def test_oracle_lines_skip_negative_pixels(orc):
    """Pixels below zero are left out of the mean (GRadon's v < 0 test)."""
    nz, ny, nx = 1, 256, 256
    stack = np.full((nz, ny, nx), 1.0, np.float32)
    stack[0, :, ::2] = -5.0
    z = np.zeros(nz, np.float32)
    lines = orc.common_lines(stack, np.zeros(1, np.float32), z, z, 90.0,
                             10.0, num=4)
    valid = lines > -1e10
    assert valid.any() and np.all(lines[valid] == np.float32(1.0))


def _stack(nz, ny, nx, seed, clean=()):
    """Positive pixels with 5% skipped (negative) ones, except in the images
    listed in 'clean', which hold none (the GPU sums those without tests)."""
    rng = np.random.default_rng(seed)
    s = (rng.random((nz, ny, nx)) * 3 + 0.1).astype(np.float32)
    skip = rng.random(s.shape) < 0.05
    skip[list(clean)] = False
    s[skip] = -1.0
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("axis0,arange", [
    (85.3, 180.0),   # all orientations: row- and column-like candidates
    (10.0, 30.0),    # column-like only (transposed reads)
    (92.0, 20.0),    # row-like only
])
def test_gpu_common_lines_bitexact(ctx, orc, axis0, arange):
    nz, ny, nx = 5, 150, 200
    # Images 1 and 3 hold no skipped pixels; image 4 only one, in its last
    # row (the last band of the per-image flags).
    stack = _stack(nz, ny, nx, seed=int(axis0), clean=(1, 3, 4))
    stack[4, ny - 1, nx // 2] = -1.0
    tilts = np.array([-12, -6, 0, 6, 12], np.float32)
    sx = np.array([3.5, -2.0, 1.25, 0.0, -7.0], np.float32)
    sy = np.array([-1.0, 4.0, -2.5, 6.0, 0.5], np.float32)
    ref = orc.common_lines(stack, tilts, sx, sy, axis0, arange, num=24)
    d = ctx.array(stack)
    got = ctx.common_lines(d, nx, ny, nz, tilts, sx, sy, axis0, arange,
                           num=24)
    d.free()
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)


# ------------------------------------------------ scores and wrap rules

def decision_agrees(score_g, score_o, kg, ko):
    """Two argmax decisions over candidate scores agree when they pick the
    same candidate, or when the oracle's own scores of the two picks differ
    by no more than the two curves differ anywhere (the implementations'
    rounding decides a tie).  Returns (agrees, is_tie)."""
    if kg == ko:
        return True, False
    diff = float(np.abs(score_g - score_o).max())
    return abs(float(score_o[ko]) - float(score_o[kg])) <= diff, True


def _grid(axis0, arange, num=100):
    step = np.float32(arange) / np.float32(num - 1)
    g = np.empty(num, np.float32)
    g[0] = np.float32(axis0) - step * np.float32(num) / np.float32(2)
    for i in range(1, num):
        g[i] = g[i - 1] + step
    return g


@pytest.mark.gpu
@pytest.mark.parametrize("state", ["coarse", "fine"])
def test_gpu_line_scores_match_oracle(ctx, orc, cfg1, cfg1_pre, state):
    """The common-line score curve (CalcScore of every candidate,
    Rotate.cu:489-561) and the line spectra (FFT1d, :430-451) of the
    512^2 x 41 SNR 0.1 workload against the oracle: scores within 1e-5 of
    the curve's peak, spectra within 1e-5 relative L2, the same best
    candidate or a tie within the curves' difference."""
    stack, tilts = cfg1_pre
    nz, ny, nx = stack.shape
    if state == "coarse":   # first search: zero shifts, axis 0, 180 deg
        sx = sy = np.zeros(nz, np.float32)
        axis0, arange = 0.0, 180.0
    else:                   # a fine search near the true axis
        z0 = int(np.argmin(np.abs(tilts)))
        sx = (cfg1.shift_x - cfg1.shift_x[z0]).astype(np.float32)
        sy = (cfg1.shift_y - cfg1.shift_y[z0]).astype(np.float32)
        axis0, arange = 85.0, 20.0
    rot = np.full(nz, axis0, np.float32)
    _, so, L = orc.tilt_axis(stack, tilts, sx, sy, rot, arange)
    spec_o = orc.line_spectra(stack, tilts, sx, sy, axis0, arange)
    d = ctx.array(stack)
    sg, best_g, Lg, spec_g = ctx.line_scores(d, nx, ny, nz, tilts, sx, sy,
                                             axis0, arange, spectra=True)
    d.free()
    assert Lg == L
    assert spec_g.shape == spec_o.shape
    from conftest import rel_l2
    assert rel_l2(spec_g, spec_o) <= 1e-5
    peak = float(np.abs(so).max())
    assert float(np.abs(sg - so).max()) <= 1e-5 * peak
    kg = int(np.argmin(np.abs(_grid(axis0, arange) - best_g)))
    ko = int(np.argmax(so))  # the first maximum, as findTiltAxis
    ok, tie = decision_agrees(sg, so, kg, ko)
    print(f"{state}: max |score diff| {np.abs(sg - so).max():.2e} "
          f"(peak {peak:.4f}), candidates {kg}/{ko}, tie {tie}")
    assert ok


def _wrap_series(axis):
    from markerfree_amd import synth
    cfg = synth.SeriesConfig(nx=256, ny=256, nz=21, vol=(256, 256, 64),
                             nblobs=300, tilt_step=6.0, axis=(axis, axis),
                             shift_max=0.0, noise="none")
    return synth.generate(cfg)


# (true axis, starting axis of every projection, range): the search result
# and the DoIt wrap (Rotate.cu:316-322) it takes.
WRAP_CASES = [
    (-50.0, 0.0, 180.0),    # best candidate below -45 with rotate[0] == 0: +180
    (140.0, 140.0, 20.0),   # best candidate above 135: -180
    (-50.0, -50.0, 20.0),   # below -45 but rotate[0] != 0: kept
]


def test_oracle_axis_wrap_rules(orc):
    """A series with its tilt axis at -50 deg searched from 0 over 180 deg
    comes back as about 130 (the +180 rule); one at 140 deg searched around
    140 comes back as about -40 (the -180 rule); the -50 deg series
    searched around -50 keeps its sign (the +180 rule needs rotate[0] ==
    0)."""
    want = [(120.0, 135.0), (-45.0, -35.0), (-55.0, -45.0)]
    for (axis, a0, arange), (lo, hi) in zip(WRAP_CASES, want):
        s = _wrap_series(axis)
        O = orc.Series(s.stack, s.tilts)
        O.preprocess()
        z = np.zeros(21, np.float32)
        O.set_params(z, z, np.full(21, a0, np.float32))
        got = O.tilt_axis(arange)
        assert lo <= got <= hi, (axis, a0, got)


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(WRAP_CASES)))
def test_gpu_axis_wrap_rules(ctx, mf, orc, case):
    axis, a0, arange = WRAP_CASES[case]
    s = _wrap_series(axis)
    G = mf.Series(ctx, s.stack, s.tilts)
    O = orc.Series(s.stack, s.tilts)
    G.preprocess()
    O.preprocess()
    z = np.zeros(21)
    rot = np.full(21, a0, np.float32)
    G.set_params(z, z, rot)
    O.set_params(z, z, rot)
    ag, ao = G.tilt_axis(arange), O.tilt_axis(arange)
    sg = G.axis_scores()
    G.close()
    if abs(ag - ao) <= 0.05:
        return
    # Otherwise only a tie of the two score curves may separate them.
    _, so, _ = orc.tilt_axis(O.stack(), O.tilts(), z.astype(np.float32),
                             z.astype(np.float32), rot, arange)
    g = _grid(a0, arange)
    unwrap = lambda a: a - 180.0 if a - a0 > 90 else (a + 180.0 if a0 - a > 90 else a)
    kg = int(np.argmin(np.abs(g - unwrap(ag))))
    ko = int(np.argmin(np.abs(g - unwrap(ao))))
    assert decision_agrees(sg, so, kg, ko)[0], (ag, ao)
