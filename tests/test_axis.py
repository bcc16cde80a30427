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
