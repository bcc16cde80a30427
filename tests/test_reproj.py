"""Reprojection for projection matching (CalcReprojFBP, CalcReproj.cu).

CPU tests pin the oracle with properties the geometry fixes; GPU tests check
the HIP kernels against the FP64 oracle: identical invalid masks and
<= 1e-5 relative L2 on the valid reprojected pixels (the tolerance stated for
reprojected images).
"""
import numpy as np
import pytest

from conftest import rel_l2

INVALID = -1e30


def tilt_series(nz=21, step=3.0, start=None):
    start = -(nz // 2) * step if start is None else start
    return (start + step * np.arange(nz)).astype(np.float32)


# ---------------------------------------------------------------- oracle pins

def test_find_proj_range_hand_cases(orc):
    t = tilt_series(21, 3.0)              # -30 .. 30, zero at index 10
    skip = np.ones(21, np.uint8)
    skip[10] = 0
    # First target after the zero tilt: only the zero tilt is available.
    assert orc.find_proj_range(t, skip, 11, 10) == (10, 10)
    skip[11] = 0
    assert orc.find_proj_range(t, skip, 9, 10) == (10, 11)
    # Everything processed except the target: +-20.5 deg window at 3 deg steps
    # keeps 6 neighbours on each side; NPROJ=10 caps the count, trimming the
    # side away from zero tilt for negative targets and towards it otherwise.
    skip[:] = 0
    skip[4] = 1
    assert orc.find_proj_range(t, skip, 4, 10) == (0, 9)
    skip[:] = 0
    skip[16] = 1
    assert orc.find_proj_range(t, skip, 16, 10) == (11, 20)
    # Stretch limit cos(a_i)/cos(a_t) <= 1.2 at a 60 degree target.
    t2 = tilt_series(41, 3.0)             # -60 .. 60
    skip = np.zeros(41, np.uint8)
    skip[40] = 1
    lo, hi = orc.find_proj_range(t2, skip, 40, 10)
    assert (lo, hi) == (38, 39)           # a_i >= 53.13 deg: 54 and 57
    assert all(np.cos(np.deg2rad(t2[i])) / 0.5 <= 1.2 for i in range(lo, hi + 1))
    # Nothing qualifies: empty range.
    skip[:] = 1
    assert orc.find_proj_range(t2, skip, 20, 10) == (-1, -1)


def test_find_proj_range_matches_product(mf, orc):
    rng = np.random.default_rng(5)
    t = tilt_series(61, 2.0)
    for _ in range(200):
        skip = (rng.random(61) < 0.4).astype(np.uint8)
        target = int(rng.integers(0, 61))
        nproj = int(rng.integers(1, 14))
        assert mf.find_proj_range(t, skip, target, nproj) == \
            orc.find_proj_range(t, skip, target, nproj)


def test_zero_tilt_identity(orc):
    """One reference at 0 deg, target at 0 deg: RP(BP(p)) == p exactly,
    except the last detector column: its voxels have rho = bx - 0.5, above
    the reference's bound rho <= bx - 1 (CalcReproj.cu:17-24), so the column is
    never backprojected and reprojects as invalid."""
    bx, by, tb = 32, 4, 12
    rng = np.random.default_rng(0)
    stack = rng.random((3, by, bx), dtype=np.float32) + 0.5
    tilts = np.array([0.0, 0.0, 40.0], np.float32)
    skip = np.array([0, 1, 1], np.uint8)
    rp = orc.reproject(stack, tb, tilts, skip, target=1, nproj=10)
    np.testing.assert_array_equal(rp[:, :-1], stack[0][:, :-1])
    assert np.all(rp[:, -1] == np.float32(INVALID))


def test_linearity(orc):
    """BP (weighted sum / sum of weights) and RP (mean) are linear in the
    projection values: f(a*p + q) = a*f(p) + f(q) with unchanged masks."""
    rng = np.random.default_rng(1)
    bx, by, tb = 48, 5, 16
    tilts = tilt_series(11, 4.0)
    p = rng.random((11, by, bx), dtype=np.float32)
    q = rng.random((11, by, bx), dtype=np.float32)
    skip = np.zeros(11, np.uint8)
    skip[8] = 1
    rp = orc.reproject(p, tb, tilts, skip, 8, 6)
    rq = orc.reproject(q, tb, tilts, skip, 8, 6)
    r = orc.reproject(np.float32(3.0) * p + q, tb, tilts, skip, 8, 6)
    valid = rp > -1e10
    assert valid.sum() > 0.5 * rp.size
    np.testing.assert_array_equal(valid, r > -1e10)
    np.testing.assert_allclose(r[valid], 3.0 * rp[valid] + rq[valid],
                               rtol=1e-5, atol=1e-6)


def test_weights_closed_form(orc):
    """Constant rows p_i = v_i: every fully covered voxel is
    sum(v_i cos a_i w_i) / sum(w_i) with w_i = cos((a_t - a_i) * 0.01745)
    and the 0.01745 degree factor of the backprojector (quirk Q12), so the
    central reprojected pixels equal that value."""
    bx, by, tb = 96, 2, 16
    tilts = tilt_series(11, 4.0)           # -20 .. 20
    vals = np.linspace(1.0, 2.0, 11).astype(np.float32)
    stack = np.broadcast_to(vals[:, None, None], (11, by, bx)).copy()
    skip = np.zeros(11, np.uint8)
    target = 9
    skip[target] = 1
    lo, hi = orc.find_proj_range(tilts, skip, target, 10)
    a = tilts[lo:hi + 1].astype(np.float64)
    w = np.cos((tilts[target] - a) * 0.01745)
    c = np.cos(a * 0.01745)
    expect = np.sum(vals[lo:hi + 1] * c * w) / np.sum(w)
    rp = orc.reproject(stack, tb, tilts, skip, target, 10)
    centre = rp[:, bx // 2 - 8: bx // 2 + 8]
    np.testing.assert_allclose(centre, expect, rtol=1e-6)


def test_point_source_geometry(orc):
    """A point at (x0, z0) projects to rho_i = x0 cos a_i + z0 sin a_i + bx/2;
    backprojecting the neighbours and reprojecting at the target puts the
    peak at x0 cos a_t + z0 sin a_t + bx/2 (catches sign/axis mistakes)."""
    # A thin slab keeps the backprojected streaks from overlapping the target
    # ray far from the point (otherwise the response is a plateau); a flipped
    # sin term would move the peak by 2 z0 sin(30) = 6 px.
    bx, by, tb = 128, 1, 16
    tilts = tilt_series(13, 5.0)           # -30 .. 30
    x0, z0 = 17.0, 6.0
    stack = np.zeros((13, by, bx), np.float32)
    for i, a in enumerate(tilts):
        r = x0 * np.cos(np.deg2rad(a)) + z0 * np.sin(np.deg2rad(a)) + bx / 2
        stack[i, 0, int(r)] = 1.0
    skip = np.zeros(13, np.uint8)
    target = 12                            # +30 deg
    skip[target] = 1
    rp = orc.reproject(stack, tb, tilts, skip, target, 10)
    at = np.deg2rad(tilts[target])
    expect = x0 * np.cos(at) + z0 * np.sin(at) + bx / 2
    # Detector pixel u covers [u, u+1); spikes were placed at int(rho), so
    # the response centre sits near u + 1 on average.
    assert abs(int(np.argmax(rp[0])) + 1.0 - expect) <= 1.5


def test_invalid_band_matches_sample_count(orc):
    """Rays with fewer than 0.8 L in-slab samples are invalid (-1e30)."""
    bx, by, tb = 64, 2, 40
    tilts = tilt_series(21, 3.0)
    stack = np.ones((21, by, bx), np.float32)
    skip = np.zeros(21, np.uint8)
    target = 20                            # +30 deg
    skip[target] = 1
    rp = orc.reproject(stack, tb, tilts, skip, target, 10)
    valid = rp > -1e10
    assert 0 < valid.sum() < rp.size
    # Validity is a property of the geometry only: it must agree with the
    # closed-form count of in-slab samples of each ray (every in-slab voxel
    # is covered at this size), and be symmetric for the mirrored target.
    c = np.cos(np.deg2rad(np.float64(tilts[target])))
    L = int(np.float32(tb) / np.float32(c) + np.float32(0.5))
    assert valid[0].sum() == valid[1].sum()
    n_valid_rays = valid[0].sum()
    tilts_m = -tilts[::-1].copy()
    skip_m = skip[::-1].copy()
    rp_m = orc.reproject(stack, tb, tilts_m, skip_m, 20 - target, 10)
    assert (rp_m[0] > -1e10).sum() == n_valid_rays
    assert L >= tb


# --------------------------------------------------------------- GPU parity

def _series(bx, by, nz, step, seed):
    rng = np.random.default_rng(seed)
    tilts = tilt_series(nz, step)
    stack = (rng.standard_normal((nz, by, bx)) * 0.3 + 1.0).astype(np.float32)
    return stack, tilts


def _check_target(ctx, orc, stack, tilts, skip, target, nproj, tb, mode):
    nz, by, bx = stack.shape
    ref = orc.reproject(stack, tb, tilts, skip, target, nproj, acc64=True)
    d_stack = ctx.array(stack)
    d_out = ctx.empty((by, bx))
    ctx.reproject(d_stack, bx, by, nz, tb, tilts, skip, target, nproj, d_out,
                  mode=mode)
    got = d_out.download()
    d_stack.free()
    d_out.free()
    vr, vg = ref > -1e10, got > -1e10
    np.testing.assert_array_equal(vr, vg)
    assert rel_l2(got, ref, vr) <= 1e-5
    return got


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("bx,by,nz,step,tb", [
    (128, 16, 21, 3.0, 40),
    (256, 8, 41, 3.0, 100),   # small-series binned geometry, reduced rows
    (200, 6, 61, 2.0, 64),    # non-multiple-of-256 slab width
])
def test_gpu_reproject_parity(ctx, orc, mode, bx, by, nz, step, tb):
    stack, tilts = _series(bx, by, nz, step, seed=bx + nz)
    z0 = int(np.argmin(np.abs(tilts)))
    rng = np.random.default_rng(7)
    for target in [z0 + 1, z0 - 3, nz - 1, 0]:
        skip = (rng.random(nz) < 0.3).astype(np.uint8)
        skip[target] = 1
        # 14 neighbours exceed the table-driven kernel's limit (mode 1 then
        # runs the table-free form).
        for nproj in (1, 10, 14):
            _check_target(ctx, orc, stack, tilts, skip, target, nproj, tb, mode)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5, 6, 7])
def test_gpu_reproject_bitexact_fp32(ctx, orc, mode):
    """Against the oracle's FP32 mode (backprojection accumulated as
    fma(v*cos, w, acc) in reference order, ray sums in FP32 in increasing
    sample order) the HIP reprojection is bit-identical."""
    stack, tilts = _series(160, 12, 41, 3.0, seed=21)
    stack[3, 2, 40:44] = INVALID
    skip = np.zeros(41, np.uint8)
    for target in (27, 5):
        skip[target] = 1
        ref = orc.reproject(stack, 48, tilts, skip, target, 10, acc64=False)
        d_stack = ctx.array(stack)
        d_out = ctx.empty(ref.shape)
        ctx.reproject(d_stack, 160, 12, 41, 48, tilts, skip, target, 10, d_out,
                      mode=mode)
        np.testing.assert_array_equal(d_out.download(), ref)
        d_stack.free()
        d_out.free()


@pytest.mark.gpu
@pytest.mark.parametrize("bx,by,tb", [(512, 6, 150), (1024, 3, 300)])
def test_gpu_fast_paths_bitexact_wide(ctx, orc, bx, by, tb):
    """Production widths, where most slab blocks and ray samples take the
    kernels' test-free fast paths: the slab and every reprojection mode are
    bit-identical to the oracle's FP32 mode, for targets across the tilt
    range (full, partial and single-neighbour ranges)."""
    stack, tilts = _series(bx, by, 61, 2.0, seed=bx)
    nz = 61
    z0 = int(np.argmin(np.abs(tilts)))
    d_stack = ctx.array(stack)
    d_out = ctx.empty((by, bx))
    d_slab = ctx.empty((by, tb, 2 * bx))
    for k in (1, 6, 20, 30):
        for target in (z0 + k, z0 - k):
            skip = np.ones(nz, np.uint8)
            skip[max(0, z0 - k + 1): z0 + k] = 0
            skip[target] = 1
            ref, slab = orc.reproject(stack, tb, tilts, skip, target, 10,
                                      acc64=False, want_slab=True)
            ctx.reproject(d_stack, bx, by, nz, tb, tilts, skip, target, 10,
                          d_out, mode=0, d_slab=d_slab)
            np.testing.assert_array_equal(d_slab.download(), slab)
            np.testing.assert_array_equal(d_out.download(), ref)
            for mode in (1, 2, 3, 4, 5, 6, 7):
                ctx.reproject(d_stack, bx, by, nz, tb, tilts, skip, target, 10,
                              d_out, mode=mode)
                np.testing.assert_array_equal(d_out.download(), ref)
    for d in (d_stack, d_out, d_slab):
        d.free()


@pytest.mark.gpu
@pytest.mark.parametrize("bx,by,tb", [(1024, 6, 300), (512, 7, 150),
                                      (1100, 5, 90), (96, 1, 30), (2048, 3, 40)])
def test_gpu_one_division_mode(ctx, orc, bx, by, tb):
    """Mode 8 divides each ray's test-free interval once by the shared weight
    sum instead of every sample: at production and edge widths (four-row,
    two-row and one-row kernels, partial last groups), with skipped samples
    inside the test-free intervals, its reprojection agrees with the
    oracle's FP64 evaluation to 1e-5 (relative L2), has the same valid rays,
    and stays within 1e-5 of mode 1 ray by ray."""
    nz = 41
    stack, tilts = _series(bx, by, nz, 3.0, seed=5 * bx + by)
    stack[20, by // 2, bx // 3: bx // 3 + 9] = INVALID
    stack[22, by - 1, 7] = INVALID
    stack[24, 0, 0] = INVALID
    z0 = int(np.argmin(np.abs(tilts)))
    d_stack = ctx.array(stack)
    d_out = ctx.empty((by, bx))
    for k in (1, 6, 12):
        for target in (z0 + k, z0 - k):
            skip = np.zeros(nz, np.uint8)
            skip[target] = 1
            ref = orc.reproject(stack, tb, tilts, skip, target, 10, acc64=True)
            ctx.reproject(d_stack, bx, by, nz, tb, tilts, skip, target, 10,
                          d_out, mode=8)
            got = d_out.download()
            ctx.reproject(d_stack, bx, by, nz, tb, tilts, skip, target, 10,
                          d_out, mode=1)
            exact = d_out.download()
            vr, vg = ref > -1e10, got > -1e10
            np.testing.assert_array_equal(vr, vg)
            np.testing.assert_array_equal(exact > -1e10, vg)
            assert rel_l2(got, ref, vr) <= 1e-5
            err = np.abs(got[vg] - exact[vg])
            assert np.all(err <= 1e-5 * np.maximum(np.abs(exact[vg]), 1.0))
    d_stack.free()
    d_out.free()


@pytest.mark.gpu
@pytest.mark.parametrize("bx,by,tb,rows", [(1024, 6, 300, 4), (512, 7, 150, 4),
                                           (300, 5, 60, 2), (96, 1, 30, 4)])
def test_gpu_slab_rows_bitexact(ctx, orc, bx, by, tb, rows):
    """Modes 6 and 7 (materialised slab of 4 or 2 image rows per voxel): the
    interleaved slab, read back row by row, equals the oracle's slab, and the
    reprojection is bit-identical -- with row counts that leave a partial
    last group, skipped samples in one row of a group (the backprojection's
    fast path finds them from non-finite voxels), and a target with no
    neighbour (an invalid slab)."""
    stack, tilts = _series(bx, by, 41, 3.0, seed=3 * bx + by)
    stack[20, by // 2, bx // 3: bx // 3 + 9] = INVALID
    # Skipped samples at the detector edges, which the slab blocks with
    # projections partly in range reach.
    stack[19, by - 1, :5] = INVALID
    stack[21, 0, bx - 5:] = INVALID
    nz = 41
    mode = 6 if rows == 4 else 7
    groups = -(-by // rows)
    d_stack = ctx.array(stack)
    d_out = ctx.empty((by, bx))
    d_slab = ctx.empty((groups, tb, 2 * bx, rows))
    z0 = int(np.argmin(np.abs(tilts)))
    cases = []
    for k in (1, 4, 12):
        skip = np.ones(nz, np.uint8)
        skip[max(0, z0 - k + 1): z0 + k] = 0
        skip[z0 + k] = 1
        cases.append((skip, z0 + k))
    lonely = np.ones(nz, np.uint8)
    cases.append((lonely, 3))            # no neighbour: FindProjRange empty
    for skip, target in cases:
        ref, slab = orc.reproject(stack, tb, tilts, skip, target, 10,
                                  acc64=False, want_slab=True)
        ctx.reproject(d_stack, bx, by, nz, tb, tilts, skip, target, 10, d_out,
                      mode=mode, d_slab=d_slab)
        got = d_slab.download()               # [g][tb][2bx][R]
        flat = got.transpose(0, 3, 1, 2).reshape(groups * rows, tb, 2 * bx)
        np.testing.assert_array_equal(flat[:by], slab)
        np.testing.assert_array_equal(d_out.download(), ref)
    for d in (d_stack, d_out, d_slab):
        d.free()


@pytest.mark.gpu
@pytest.mark.parametrize("bx,by,tb", [(1100, 5, 90), (300, 7, 60), (96, 1, 30),
                                      (160, 2, 40), (250, 9, 50)])
def test_gpu_row_groups_bitexact(ctx, orc, bx, by, tb):
    """The multi-row forms (mode 1: four rows per unit where they fit in
    LDS, else two; modes 4 and 5: two rows with and without tables) against
    the oracle's FP32 mode and the one-row table form (mode 3): row counts
    that leave a partial last group, a single row, ray chunks that split an
    image row, and skipped samples in one row of a group but not the
    others."""
    stack, tilts = _series(bx, by, 41, 3.0, seed=bx + by)
    if by > 2:
        stack[20, 2, bx // 3: bx // 3 + 5] = INVALID
    stack[22, by - 1, 7] = INVALID
    # Position 0 is where the four-row form's out-of-range lookups read.
    stack[24, 0, 0] = INVALID
    skip = np.zeros(41, np.uint8)
    d_stack = ctx.array(stack)
    d_out = ctx.empty((by, bx))
    for target in (21, 33, 2):
        skip[:] = 0
        skip[target] = 1
        ref = orc.reproject(stack, tb, tilts, skip, target, 10, acc64=False)
        for mode in (1, 3, 4, 5):
            ctx.reproject(d_stack, bx, by, 41, tb, tilts, skip, target, 10,
                          d_out, mode=mode)
            np.testing.assert_array_equal(d_out.download(), ref)
    d_stack.free()
    d_out.free()


@pytest.mark.gpu
@pytest.mark.parametrize("bx", [1920, 2048])
def test_gpu_wide_rows_bitexact(ctx, orc, bx):
    """Widths where only the table-free two-row form fits (at 2048 columns
    and ten neighbours its slots take exactly the 160 KB of LDS): mode 1
    (which picks it), mode 5 and the one-row form against the oracle's FP32
    mode, with a skipped sample at position 0 of one row."""
    by, nz, tb = 3, 41, 40
    stack, tilts = _series(bx, by, nz, 3.0, seed=bx)
    stack[22, 1, 0] = INVALID
    d_stack = ctx.array(stack)
    d_out = ctx.empty((by, bx))
    skip = np.zeros(nz, np.uint8)
    for target in (20, 31):
        skip[:] = 0
        skip[target] = 1
        ref = orc.reproject(stack, tb, tilts, skip, target, 10, acc64=False)
        for mode in (1, 5, 3):
            ctx.reproject(d_stack, bx, by, nz, tb, tilts, skip, target, 10,
                          d_out, mode=mode)
            np.testing.assert_array_equal(d_out.download(), ref)
    d_stack.free()
    d_out.free()


@pytest.mark.gpu
def test_gpu_reproject_invalid_samples(ctx, orc):
    """Invalid (-1e30) corrected pixels are skipped exactly as the reference
    does (mDIntProj, CalcReproj.cu:17-24)."""
    stack, tilts = _series(96, 4, 21, 3.0, seed=3)
    stack[:, :, 10:14] = INVALID
    stack[4, 1, :] = INVALID
    skip = np.zeros(21, np.uint8)
    skip[15] = 1
    for mode in (0, 1, 2, 3, 4, 5, 6, 7, 8):
        _check_target(ctx, orc, stack, tilts, skip, 15, 10, 30, mode)


@pytest.mark.gpu
def test_gpu_reproject_empty_range(ctx, orc):
    stack, tilts = _series(64, 3, 11, 3.0, seed=4)
    skip = np.ones(11, np.uint8)
    d_stack = ctx.array(stack)
    d_out = ctx.empty((3, 64))
    for mode in (0, 1, 2, 3, 4, 5, 6, 7, 8):
        ctx.reproject(d_stack, 64, 3, 11, 20, tilts, skip, 5, 10, d_out,
                      mode=mode)
        assert np.all(d_out.download() == np.float32(INVALID))
    d_stack.free()
    d_out.free()


@pytest.mark.gpu
def test_gpu_slab_matches_oracle(ctx, orc):
    """The materialised slab itself (mGBackProj output) against the oracle."""
    stack, tilts = _series(64, 4, 21, 3.0, seed=9)
    skip = np.zeros(21, np.uint8)
    skip[18] = 1
    tb = 24
    ref, slab = orc.reproject(stack, tb, tilts, skip, 18, 10, want_slab=True)
    d_stack = ctx.array(stack)
    d_out = ctx.empty((4, 64))
    d_slab = ctx.empty((4, tb, 128))
    ctx.reproject(d_stack, 64, 4, 21, tb, tilts, skip, 18, 10, d_out, mode=0,
                  d_slab=d_slab)
    got = d_slab.download()
    vr = slab > -1e10
    np.testing.assert_array_equal(vr, got > -1e10)
    assert rel_l2(got, slab, vr) <= 1e-6
    for d in (d_stack, d_out, d_slab):
        d.free()
