"""Fiducial residual evaluator (the cal_res tool's evaluation,
tool/src/main/main.cpp:35-237, triangulation
tool/src/nbundle/bundle_core.cpp:149-216), an accuracy check independent of
parity: marker tracks with known 3-D positions come from the generator, the
alignment from the GPU run or from the oracle's run of the same series.

CPU tests pin the evaluator: zero residual under the true alignment, growth
with an axis error, a wrong shift found in its own section, the common
rotation-centre offset absorbed only when fitted, and the least-squares
positions against numpy's solver.
"""
import subprocess

import numpy as np
import pytest


def _clean(nz=21, axis=84.0, nblobs=60, seed=5):
    from markerfree_amd import synth
    cfg = synth.SeriesConfig(nx=256, ny=256, nz=nz, vol=(256, 256, 64),
                             nblobs=nblobs, tilt_step=6.0, axis=(axis, axis),
                             shift_max=8.0, noise="none",
                             seeds=(seed, seed + 1, seed + 2))
    return cfg, synth.generate(cfg), synth.fiducials(cfg, nblobs)


def _res(s, f, rot=None, sx=None, sy=None, tilts=None, **kw):
    from markerfree_amd import fileio
    nz = len(s.tilts)
    return fileio.fiducial_residuals(
        256, 256,
        np.full(nz, s.axis, np.float32) if rot is None else rot,
        s.shift_x.astype(np.float64) if sx is None else sx,
        s.shift_y.astype(np.float64) if sy is None else sy,
        s.tilts if tilts is None else tilts, f, **kw)


def test_true_alignment_has_no_residual():
    _, s, f = _clean()
    r = _res(s, f)
    assert r["markers"] == f.shape[0] and r["observations"] > 0.9 * f[..., 0].size
    assert r["rms"] < 1e-4 and r["max"] < 1e-4


def test_axis_error_grows_the_residual():
    _, s, f = _clean()
    nz = len(s.tilts)
    rms = [_res(s, f, rot=np.full(nz, s.axis + d, np.float32))["rms"]
           for d in (0.5, 1.0, 2.0)]
    assert 0.1 < rms[0] < rms[1] < rms[2]
    # Small rotations act linearly on the positions.
    assert 1.8 < rms[2] / rms[1] < 2.2 and 1.8 < rms[1] / rms[0] < 2.2


def test_wrong_shift_shows_in_its_section():
    _, s, f = _clean()
    sx = s.shift_x.astype(np.float64).copy()
    sx[4] += 1.5
    r = _res(s, f, sx=sx)
    per = r["per_section"]
    assert int(np.argmax(per)) == 4 and per[4] > 1.0
    assert np.median(per) < 0.2


def test_common_centre_offset_is_a_free_gauge():
    """Shifting every section by the same amount along the aligned x axis
    moves the rotation centre, not the markers: fitted, it leaves no
    residual and comes back as the centre offset; not fitted, it does."""
    _, s, f = _clean()
    th = np.deg2rad(s.axis)
    d = 3.0   # aligned-frame x offset -> raw shift R(axis) (-d, 0)
    sx = s.shift_x.astype(np.float64) - d * np.cos(th)
    sy = s.shift_y.astype(np.float64) - d * np.sin(th)
    fitted = _res(s, f, sx=sx, sy=sy)
    assert fitted["rms"] < 1e-4 and abs(abs(fitted["centre"]) - d) < 1e-3
    free = _res(s, f, sx=sx, sy=sy, fit_centre=False)
    assert free["rms"] > 0.5


def test_positions_match_a_direct_least_squares_solve():
    """Noisy tracks: each marker's (X, Z) (with the fitted common offset
    removed) and Y equal numpy's least-squares solution of the same
    equations, and the RMS equals the solution's residual."""
    _, s, f = _clean()
    rng = np.random.default_rng(2)
    f = (f + rng.normal(0, 0.3, f.shape)).astype(np.float32)
    r = _res(s, f, fit_centre=False)
    nz = len(s.tilts)
    th = np.deg2rad(s.axis)
    c, sn = np.cos(th), np.sin(th)
    beta = np.deg2rad(s.tilts.astype(np.float64))
    err2, n = 0.0, 0
    for i in range(f.shape[0]):
        ok = np.isfinite(f[i, :, 0])
        ux = f[i, ok, 0] - 128 - s.shift_x[ok]
        uy = f[i, ok, 1] - 128 - s.shift_y[ok]
        px, py = c * ux + sn * uy, -sn * ux + c * uy
        M = np.stack([np.cos(beta[ok]), np.sin(beta[ok])], 1)
        (X, Z), *_ = np.linalg.lstsq(M, px, rcond=None)
        Y = py.mean()
        np.testing.assert_allclose(r["xyz"][i], [X, Y, Z], atol=1e-3)
        err2 += ((px - M @ [X, Z]) ** 2 + (py - Y) ** 2).sum()
        n += ok.sum()
    assert abs(r["rms"] - np.sqrt(err2 / n)) < 1e-4


def test_txt_round_trip(tmp_path):
    from markerfree_amd import fileio
    rng = np.random.default_rng(4)
    nz = 11
    rot = rng.uniform(80, 95, nz).astype(np.float32)
    sx, sy = rng.uniform(-40, 40, nz), rng.uniform(-40, 40, nz)
    t = np.linspace(-50, 50, nz).astype(np.float32)
    p = tmp_path / "a.txt"
    fileio.save_txt(p, 512, 512, rot, sx, sy, t, -1.0, 2, 3.5)
    r, x, y, tt, off, b = fileio.read_txt(p, nz)
    np.testing.assert_allclose(r, rot, atol=5e-5)
    np.testing.assert_allclose(x, sx, atol=5e-4)
    np.testing.assert_allclose(y, sy, atol=5e-4)
    np.testing.assert_allclose(tt, t, atol=5e-3)
    assert off == -1.0 and b == 2


def test_cal_res_command_line(tmp_path):
    """cal_res on a stack, its .txt alignment and a marker file: the true
    alignment gives no residual, an axis 1 degree off does."""
    from markerfree_amd import fileio, synth
    from markerfree_amd import CLI_PATH
    cfg, s, f = _clean()
    mrc, tlt = synth.write_series(s, tmp_path, "s")
    nz = len(s.tilts)
    fid = tmp_path / "fids.txt"
    with open(fid, "w") as fh:
        for m in range(f.shape[0]):
            for z in range(nz):
                if np.isfinite(f[m, z, 0]):
                    fh.write(f"{m} {z} {f[m, z, 0]:.4f} {f[m, z, 1]:.4f}\n")
    tool = CLI_PATH.parent / "cal_res"
    out = {}
    for name, axis in (("true", s.axis), ("off", s.axis + 1.0)):
        txt = tmp_path / f"{name}.txt"
        fileio.save_txt(txt, 256, 256, np.full(nz, axis, np.float32),
                        s.shift_x.astype(np.float64),
                        s.shift_y.astype(np.float64), s.tilts, 0.0, 1, 0.0)
        p = subprocess.run([str(tool), "-i", str(mrc), "-a", str(tlt), "-x",
                            str(txt), "-f", str(fid)], capture_output=True,
                           text=True, check=True)
        line = [l for l in p.stdout.splitlines() if l.startswith("RMS")][0]
        out[name] = float(line.split(":")[1].split()[0])
    assert out["true"] < 1e-3 and out["off"] > 0.2, out


@pytest.mark.gpu
def test_gpu_alignment_residual_vs_oracle(ctx, mf, cfg1):
    """The 512^2 x 41 SNR 0.1 workload: residual of the generator's blob
    centres under the GPU's alignment and under the oracle's alignment of
    the same series (end-to-end fixture).  The two alignments differ by the
    method's own end-to-end sensitivity; their accuracy must be alike."""
    import json
    from pathlib import Path
    from markerfree_amd import fileio, synth
    path = Path(__file__).resolve().parent / "golden" / "e2e_small.json"
    if not path.exists():
        pytest.skip("e2e_small.json not generated")
    fx = json.loads(path.read_text())
    f = synth.fiducials(synth.PRESETS["small"], 400)
    G = mf.Series(ctx, cfg1.stack, cfg1.tilts)
    G.run(align_z=100, nproj=10)
    sx, sy, rot, off = G.params()
    G.close()
    nz = len(cfg1.tilts)
    nx = cfg1.stack.shape[2]

    def res(rot_, sx_, sy_, off_):
        return fileio.fiducial_residuals(
            nx, nx, np.full(nz, rot_, np.float32), np.asarray(sx_, np.float64),
            np.asarray(sy_, np.float64),
            (cfg1.tilts + np.float32(off_)).astype(np.float32), f,
            min_views=int(np.ceil(0.7 * nz)))

    g = res(float(rot[0]), sx, sy, off)
    o = res(fx["oracle"]["axis"], fx["oracle"]["sx"], fx["oracle"]["sy"],
            fx["oracle"]["offset"])
    t = res(cfg1.axis, cfg1.shift_x, cfg1.shift_y, 0.0)
    print(f"fiducial RMS: GPU {g['rms']:.3f} px, oracle {o['rms']:.3f} px, "
          f"truth {t['rms']:.4f} px over {g['markers']} markers")
    assert t["rms"] < 1e-3
    assert g["rms"] <= max(1.5 * o["rms"], o["rms"] + 0.5), (g["rms"], o["rms"])
