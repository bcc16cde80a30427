"""The steps either side of the alignment path, against the oracle:

- ingest: a descending-tilt file (quirks Q2, Q3: the stack is reversed into
  ascending order, mass normalisation takes the zero-tilt index from the
  file order) and a stack streamed in chunks (mf_series_upload), through the
  GPU preprocessing, against the oracle's preprocessed stack;
- output: the final correction at OutBin 1, 2 and 3 (Process::CorrectStack,
  Process.cpp:329-372) against orc_correct_final;
- the .xf file (SaveParam::SaveXf, SaveParam.cpp:15-120) read the way an
  IMOD-style consumer reads it -- x_out = A x_in + d about the image centre
  -- maps every pixel of the GPU's aligned image back onto the raw pixel the
  correction took it from.
"""
import numpy as np
import pytest

from conftest import rel_l2


def _series(nz=23, tilt_min=-30.0, step=3.0, seed=31, nx=320, ny=256):
    from markerfree_amd import synth
    cfg = synth.SeriesConfig(nx=nx, ny=ny, nz=nz, vol=(nx, ny, 64),
                             nblobs=250, tilt_min=tilt_min, tilt_step=step,
                             axis=(84.0, 84.0), shift_max=8.0, snr=1.0,
                             seeds=(seed, seed + 1, seed + 2))
    return synth.generate(cfg)


@pytest.fixture(scope="module")
def asym():
    # -30 .. +36 deg: asymmetric, so that the zero-tilt index differs
    # between the file order and the ascending order when the file descends.
    return _series()


def test_oracle_descending_file_uses_file_order_zero_index(orc, asym):
    """Quirk Q3 (Process.cpp:228-235, Preprocess.cu:168-169): for a
    descending, asymmetric file the preprocessed stack differs from the
    ascending file's only by one global scale, the ratio of the box means of
    the two images the two zero-tilt indices pick."""
    s = asym
    A = orc.Series(s.stack, s.tilts)
    A.preprocess()
    D = orc.Series(s.stack[::-1].copy(), s.tilts[::-1].copy())
    D.preprocess()
    np.testing.assert_array_equal(A.tilts(), D.tilts())
    a, d = A.stack(), D.stack()
    ratio = d.astype(np.float64).sum() / a.astype(np.float64).sum()
    assert abs(ratio - 1.0) > 1e-3          # Q3 changes the scale
    np.testing.assert_allclose(d, a * ratio, rtol=2e-5, atol=1e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("descending", [False, True])
@pytest.mark.parametrize("streamed", [False, True])
def test_gpu_ingest_matches_oracle(ctx, mf, orc, asym, descending, streamed):
    """The raw stack in file order, whole or streamed in chunks of 5 images
    from page-locked memory, then GPU preprocessing: the preprocessed stack
    within 1e-6 relative of the oracle's, the same ascending tilts."""
    s = asym
    stack = s.stack[::-1].copy() if descending else s.stack
    tilts = s.tilts[::-1].copy() if descending else s.tilts
    O = orc.Series(stack, tilts)
    O.preprocess()
    if streamed:
        nz, ny, nx = stack.shape
        G = mf.Series(ctx, None, tilts, shape=(nz, ny, nx))
        bufs = [ctx.pinned_empty((5, ny, nx)) for _ in range(2)]
        for k, z0 in enumerate(range(0, nz, 5)):
            n = min(5, nz - z0)
            b = bufs[k % 2]
            b[:n] = stack[z0:z0 + n]
            G.upload(z0, b[:n])
        G.upload(0, None)
    else:
        G = mf.Series(ctx, stack, tilts)
    G.preprocess()
    assert rel_l2(G.stack(), O.stack()) <= 1e-6
    np.testing.assert_array_equal(G.tilts(), O.tilts())
    G.close()


@pytest.mark.gpu
@pytest.mark.parametrize("out_bin", [1, 2, 3])
def test_gpu_final_correction_matches_oracle(ctx, mf, orc, asym, out_bin):
    """Process::CorrectStack at OutBin 1, 2 and 3 from identical parameters:
    rotation, shift, random fill and the b x b sum binning (quirk Q6)
    bit-identical to the oracle."""
    s = asym
    nz = len(s.tilts)
    rng = np.random.default_rng(out_bin)
    sx = rng.uniform(-9, 9, nz)
    sy = rng.uniform(-9, 9, nz)
    rot = np.full(nz, 84.3, np.float32)
    G = mf.Series(ctx, s.stack, s.tilts)
    O = orc.Series(s.stack, s.tilts)
    G.preprocess()
    O.preprocess()
    G.set_params(sx, sy, rot)
    O.set_params(sx.astype(np.float32), sy.astype(np.float32), rot)
    got = G.correct(out_bin)
    want = O.correct_final(out_bin)
    G.close()
    bx, by = orc.binned_size(s.stack.shape[2], s.stack.shape[1], out_bin)
    assert got.shape == (nz, by, bx)
    np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("out_bin", [1, 2])
def test_gpu_async_correction_same_stack(ctx, mf, asym, out_bin):
    """mf_correct_async: the download into page-locked memory left running
    while another series preprocesses and corrects, then mf_series_wait,
    gives mf_correct's stack bit for bit; the same host buffer then takes
    the other series' stack; ordinary memory is filled before the call
    returns."""
    s = asym
    nz = len(s.tilts)
    rng = np.random.default_rng(10 + out_bin)
    G = [mf.Series(ctx, s.stack, s.tilts) for _ in range(2)]
    rot = np.full(nz, 84.3, np.float32)
    for k, g in enumerate(G):
        g.preprocess()
        g.set_params(rng.uniform(-9, 9, nz), rng.uniform(-9, 9, nz), rot + k)
    want = [g.correct(out_bin) for g in G]
    pin = mf.pinned_empty(want[0].shape)
    pin[:] = 0.0
    got = G[0].correct(out_bin, out=pin, wait=False)
    assert got is pin
    H = mf.Series(ctx, s.stack, s.tilts)   # other work meanwhile
    H.preprocess()
    other = H.correct(out_bin)
    G[0].wait()
    np.testing.assert_array_equal(pin, want[0])
    G[0].trim()
    G[1].correct(out_bin, out=pin, wait=False)
    G[1].wait()
    np.testing.assert_array_equal(pin, want[1])
    plain = np.zeros_like(want[1])
    G[1].correct(out_bin, out=plain, wait=False)
    np.testing.assert_array_equal(plain, want[1])
    assert other.shape == want[0].shape
    for g in G + [H]:
        g.close()


def _read_xf(path):
    rows = [list(map(float, line.split()))
            for line in path.read_text().splitlines()]
    return np.array(rows)


@pytest.mark.gpu
def test_gpu_xf_maps_aligned_pixels_back_to_their_source(ctx, mf, tmp_path):
    """The .xf written for a real GPU alignment, read as an IMOD-style
    consumer reads it (A = [[a11, a12], [a21, a22]], x_out = A x_in + d with
    coordinates about the image centre and pixel centres at i + 0.5), maps
    each aligned pixel back to within half a pixel (plus the 3-decimal print
    precision) of the raw pixel the GPU's correction took it from.  The
    correction is probed on an image whose pixels hold their own raw index;
    the reference's correction takes raw pixel (int)(q + 0.5) for the raw
    position q, i.e. raw pixel k sits at position k.  Positions below 0 are
    left out: there (int) truncates toward zero (quirk Q11), so the
    correction still takes pixel 0 for q in (-1, -0.5), which a consumer that
    rounds cannot reproduce."""
    from markerfree_amd import fileio
    s = _series(nz=9, tilt_min=-12.0, nx=256, ny=192, seed=41)
    nz, ny, nx = s.stack.shape
    G = mf.Series(ctx, s.stack, s.tilts)
    G.run(align_z=64, nproj=10)
    sx, sy, rot, _ = G.params()
    G.close()
    path = tmp_path / "series.xf"
    fileio.save_xf(path, nx, rot, sx, sy, 0)
    xf = _read_xf(path)
    assert xf.shape == (nz, 6)
    index = np.arange(nx * ny, dtype=np.float32).reshape(ny, nx)
    d_raw, d_out = ctx.array(index), ctx.empty((ny, nx))
    jj, ii = np.mgrid[0:ny, 0:nx]
    worst = 0.0
    for z in range(nz):
        ctx.correct_proj(d_raw, nx, ny, float(sx[z]), float(sy[z]),
                         float(rot[z]), False, False, 1, False, d_out)
        src = d_out.download()
        valid = src > -1e10
        k = src[valid].astype(np.int64)
        kr, kc = k // nx, k % nx
        a = xf[z, :4].reshape(2, 2)
        d = xf[z, 4:]
        ai = np.linalg.inv(a)
        po = np.stack([ii[valid] + 0.5 - nx / 2 - d[0],
                       jj[valid] + 0.5 - ny / 2 - d[1]])
        q = ai @ po
        qx, qy = q[0] + nx / 2, q[1] + ny / 2
        keep = (qx > 0.2) & (qy > 0.2)   # the Q11 band, with print margin
        err = max(np.abs(qx - kc)[keep].max(), np.abs(qy - kr)[keep].max())
        worst = max(worst, float(err))
        print(f"image {z}: {valid.mean():.2f} valid, {keep.mean():.3f} kept, "
              f"worst |q - k| {err:.3f}")
        assert valid.mean() > 0.5 and keep.mean() > 0.97
    d_raw.free()
    d_out.free()
    # 0.5 px of nearest-pixel rounding plus the %9.3f rounding of A and d
    # over coordinates of up to ~170 px from the centre.
    assert worst <= 0.5 + 0.1, worst
