"""Pins for the oracle's series stages (no GPU), each against something the
mathematics or the generator fixes rather than the oracle's own formula:
preprocessing in closed form, the common-line axis search on a series with a
known axis that lies on the candidate grid, the offset correlation's quirk
Q7 on images whose moments are known, the rotation-centre least-squares fit
on shifts generated from a known centre height, and the coarse and
projection-matching sweeps against the generator's shifts."""
import numpy as np
import pytest


def _series(axis=60.0, shift_max=0.0, tilt_min=-60.0, seeds=(31, 32, 33)):
    from markerfree_amd import synth
    cfg = synth.SeriesConfig(nx=256, ny=256, nz=41, vol=(256, 256, 64),
                             nblobs=300, tilt_step=3.0, tilt_min=tilt_min,
                             axis=(axis, axis), shift_max=shift_max,
                             noise="none", seeds=seeds)
    return synth.generate(cfg)


@pytest.mark.parametrize("descending", [False, True])
def test_preprocess_closed_form(orc, descending):
    """SetPositive + MassNormalization (Preprocess.cu:103-177) on constant
    slices: every value is shifted by the stack minimum when that is
    negative, then slice z is scaled by ref/(mean_z + 1e-5), ref = the
    zero-tilt slice's mean capped at 1000.  A descending file is reversed
    into ascending order first, while the zero-tilt index is taken from
    the file order (quirk Q3).  Invalid pixels (-1e30) are left alone."""
    nz, n = 7, 24
    tilts = np.linspace(-30, 30, nz).astype(np.float32)
    levels = np.array([-3, 5, 900, 1500, 40, 7, 11], np.float32)
    stack = np.repeat(levels[:, None, None], n * n, 1).reshape(nz, n, n)
    stack[2, 0, 0] = -1e30
    if descending:
        tilts = tilts[::-1].copy()
    O = orc.Series(stack, tilts)
    O.preprocess()
    got = O.stack()
    shifted = (levels - levels.min()).astype(np.float32)
    order = shifted[::-1] if descending else shifted
    z0_file = int(np.argmin(np.abs(tilts)))
    ref = min(order[z0_file], np.float32(1000.0))
    want = order * (ref / (order + np.float32(1e-5)))
    np.testing.assert_allclose(got[:, 5, 5], want, rtol=1e-6)
    zi = nz - 1 - 2 if descending else 2
    assert got[zi, 0, 0] == np.float32(-1e30)
    np.testing.assert_array_equal(O.tilts(), np.sort(tilts))


def test_tilt_axis_recovers_grid_axis(orc):
    """CalcTIltAxis over 180 degrees around 0 (candidates -90.9 + k*180/99)
    finds the generator's 60-degree axis, which is candidate k = 83, from a
    noise-free series; the score curve peaks there."""
    s = _series(axis=60.0)
    O = orc.Series(s.stack, s.tilts)
    O.preprocess()
    z = np.zeros(41, np.float32)
    axis, scores, L = orc.tilt_axis(O.stack(), O.tilts(), z, z, z, 180.0)
    assert abs(axis - 60.0) < 1e-3
    assert int(np.argmax(scores)) == 83
    assert L == 154


def test_offset_correlate_quirk_q7(orc):
    """FindOffset::Correlate's "std" is sqrt(E[x^2] - E[x]) (quirk Q7), not
    the standard deviation.  For a {0,3} image with a fraction p of 3s
    correlated with itself, E[x] = 3p and E[x^2] = 9p, so
    cc = (9p - 9p^2) / (9p - 3p) = 1.5 (1 - p), where the textbook NCC gives
    1; against its complement 3 - x, cc = -1.5 sqrt(p (1 - p)).  The stretch
    samples bilinearly inside [0, w-1) x [0, h-1), so the last row and
    column are outside and skipped; a -1e30 pixel is skipped; an image
    constant along x is unchanged by the x-stretch between tilts."""
    n = 16
    yy, xx = np.mgrid[0:n, 0:n]
    a = np.where((xx + yy) % 2 == 0, 3.0, 0.0).astype(np.float32)
    inner = a[:n - 1, :n - 1]
    p = float((inner == 3).mean())                      # 113 / 225
    assert orc.offset_correlate(a, a, 0.0, 0.0) == pytest.approx(
        1.5 * (1 - p), abs=1e-6)
    assert orc.offset_correlate(a, 3 - a, 0.0, 0.0) == pytest.approx(
        -1.5 * np.sqrt(p * (1 - p)), abs=1e-6)
    rows = np.where(yy % 2 == 0, 3.0, 0.0).astype(np.float32)
    p_rows = float((rows[:n - 1, 0] == 3).mean())        # 8 / 15
    assert orc.offset_correlate(rows, rows, 10.0, 20.0) == pytest.approx(
        1.5 * (1 - p_rows), abs=1e-5)
    b = a.copy()
    b[3, 3] = -1e30                                      # a 3
    b[3, 4] = -1e30                                      # a 0
    valid = inner.copy()
    valid[3, 3] = valid[3, 4] = -1
    p_b = float((valid == 3).sum() / (valid >= 0).sum())  # 112 / 223
    assert orc.offset_correlate(b, b, 0.0, 0.0) == pytest.approx(
        1.5 * (1 - p_b), abs=1e-6)


def test_rotation_centre_fit_closed_form(orc):
    """FitRotCenterZ (ProjAlign.cpp:459-486): shifts equal to the shifts a
    rotation centre at height Z0 induces, R(axis) (-Z0 sin(theta), 0)
    (CalcZInducedShift, :553-561), fit back to exactly Z0."""
    s = _series(axis=60.0)
    O = orc.Series(s.stack, s.tilts)
    O.preprocess()
    th = np.deg2rad(O.tilts().astype(np.float64))
    a = np.deg2rad(60.0)
    z0 = 5.0
    u = -z0 * np.sin(th)
    sx = (np.cos(a) * u).astype(np.float32)
    sy = (np.sin(a) * u).astype(np.float32)
    O.set_params(sx, sy, np.full(41, 60.0, np.float32))
    O.projalign_setup(0.0, 64, 10)
    O.projalign_iter()
    assert O.pa_z0() == pytest.approx(z0, abs=1e-4)


def _axis_frame(dx, dy, axis):
    a = np.deg2rad(axis)
    return (np.cos(a) * dx + np.sin(a) * dy, -np.sin(a) * dx + np.cos(a) * dy)


def test_coarse_and_projection_matching_recover_shifts(orc):
    """Against the generator's shifts (relative to the zero tilt) on a
    noise-free series with a known axis: one coarse stretched-CC sweep
    (Transform::DoIt) removes most of the shift spread, and three
    projection-matching sweeps from zero shifts (ProjAlign::DoIt, quirk Q16)
    converge (last max update < 0.5 px) to the truth up to what projection
    matching leaves free: a rotation-centre height term a*sin(theta) across
    the axis, and the correction's half-pixel centre convention (a
    constant)."""
    s = _series(axis=60.0, shift_max=10.0)
    O = orc.Series(s.stack, s.tilts)
    O.preprocess()
    z0 = int(np.argmin(np.abs(s.tilts)))
    ex = (s.shift_x - s.shift_x[z0]).astype(np.float64)
    ey = (s.shift_y - s.shift_y[z0]).astype(np.float64)
    zeros = np.zeros(41, np.float32)
    rot = np.full(41, 60.0, np.float32)
    O.set_params(zeros, zeros, rot)
    O.coarse_setup()
    O.coarse_iter()
    sx, sy, _, _ = O.params()
    spread = np.sqrt(np.mean(ex ** 2 + ey ** 2))
    rms = np.sqrt(np.mean((sx - ex) ** 2 + (sy - ey) ** 2))
    assert rms < 0.3 * spread, (rms, spread)

    O.set_params(zeros, zeros, rot)
    O.projalign_setup(0.0, 64, 10)
    errs = [O.projalign_iter() for _ in range(3)]
    assert errs[-1] < 0.5 and errs[0] > 5.0, errs
    sx, sy, _, _ = O.params()
    u, v = _axis_frame(sx - ex, sy - ey, 60.0)
    th = np.deg2rad(s.tilts.astype(np.float64))
    A = np.stack([np.sin(th), np.ones_like(th)], 1)
    coef, *_ = np.linalg.lstsq(A, u, rcond=None)
    assert np.sqrt(np.mean((u - A @ coef) ** 2)) < 1.2
    assert np.sqrt(np.mean((v - v.mean()) ** 2)) < 0.6
