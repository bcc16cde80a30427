"""FFTs standing in for cuFFT (method/Cufft1&2.cu).

Pinned against the DFT definition through numpy.fft at every transform length
the alignment path meets (including the non-smooth lengths of the size table:
818 = 2*409, 1364 = 4*11*31, 646 = 2*17*19, common-line lengths 410, 922,
1946, 5228, ...), and against a direct O(n^2) sum on tiny lengths.
"""
import numpy as np
import pytest

from conftest import rel_l2

# (nx, ny) image sizes of the correlation stages at the benchmark shapes.
SIZES_2D = [(512, 512), (1024, 1024), (1152, 818), (1920, 1364), (910, 646),
            (818, 1152), (64, 48), (30, 22)]
# Real 1-D lengths: common-line lengths and R-weighting row lengths.
SIZES_1D = [410, 412, 922, 924, 1946, 1948, 3990, 3994, 5228, 5660, 544, 1024,
            1920, 1364, 16, 18, 82, 2 * 61, 2 * 67, 2 * 97]


def padded(img):
    ny, nx = img.shape
    buf = np.zeros((ny, (nx // 2 + 1) * 2), np.float32)
    buf[:, :nx] = img
    return buf


def test_oracle_matches_direct_dft(orc):
    rng = np.random.default_rng(0)
    for nx, ny in [(6, 4), (10, 7 * 2), (2 * 67, 4)]:
        img = rng.standard_normal((ny, nx)).astype(np.float32)
        jy, jx = np.arange(ny), np.arange(nx)
        fy = np.exp(-2j * np.pi * np.outer(jy, jy) / ny)
        fx = np.exp(-2j * np.pi * np.outer(jx, jx[: nx // 2 + 1]) / nx)
        direct = fy @ img.astype(np.float64) @ fx
        got = orc.fft_r2c_2d(img)
        assert rel_l2(got, direct) < 1e-6


@pytest.mark.parametrize("nx,ny", SIZES_2D[:6])
def test_oracle_fft2d_numpy(orc, nx, ny):
    rng = np.random.default_rng(nx + ny)
    img = rng.standard_normal((ny, nx)).astype(np.float32)
    ref = np.fft.rfft2(img.astype(np.float64))
    got = orc.fft_r2c_2d(img)
    assert rel_l2(got, ref) < 1e-6
    back = orc.fft_c2r_2d(got, nx)
    assert rel_l2(back, img * (nx * ny)) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny", SIZES_2D)
def test_gpu_fft2d(ctx, nx, ny):
    rng = np.random.default_rng(nx * 3 + ny)
    img = rng.standard_normal((2, ny, nx)).astype(np.float32)
    buf = np.stack([padded(im) for im in img])
    d = ctx.array(buf)
    ctx.fft_r2c_2d(d, nx, ny, nimg=2, scale=1.0)
    spec = d.download().view(np.complex64)
    for k in range(2):
        ref = np.fft.rfft2(img[k].astype(np.float64))
        assert rel_l2(spec[k], ref) < 2e-6, (nx, ny, k)
    ctx.fft_c2r_2d(d, nx, ny, nimg=2)
    back = d.download()[:, :, :nx]
    assert rel_l2(back, img * (nx * ny)) < 2e-6
    d.free()


@pytest.mark.gpu
def test_gpu_fft2d_scale(ctx):
    """bNorm: forward result multiplied by 1/(nx*ny) (Cufft1&2.cu:53-60)."""
    nx, ny = 64, 40
    img = np.random.default_rng(1).standard_normal((ny, nx)).astype(np.float32)
    d = ctx.array(padded(img))
    ctx.fft_r2c_2d(d, nx, ny, scale=1.0 / (nx * ny))
    ref = np.fft.rfft2(img.astype(np.float64)) / (nx * ny)
    assert rel_l2(d.download().view(np.complex64), ref) < 2e-6
    d.free()


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES_1D)
def test_gpu_fft1d(ctx, n):
    nl = 37
    pitch = (n // 2 + 1) * 2
    rng = np.random.default_rng(n)
    x = rng.standard_normal((nl, n)).astype(np.float32)
    buf = np.zeros((nl, pitch), np.float32)
    buf[:, :n] = x
    d = ctx.array(buf)
    ctx.fft_r2c_1d(d, n, nl, pitch)
    got = d.download().view(np.complex64)
    ref = np.fft.rfft(x.astype(np.float64), axis=1)
    assert rel_l2(got, ref) < 2e-6
    ctx.fft_c2r_1d(d, n, nl, pitch)
    back = d.download()[:, :n]
    assert rel_l2(back, x * n) < 2e-6
    d.free()
