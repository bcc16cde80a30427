"""The drop-in command line end to end on the GPU: a synthetic series written
as an MRC stack and a .rawtlt file goes through `markerfree`, and its
outputs (aligned MRC, .txt or .xf) are checked against the library run on
the same data through the C-ABI (Process::DoIt, main/Process.cpp:34-92)."""
import subprocess

import numpy as np
import pytest


def _series():
    from markerfree_amd import synth
    cfg = synth.SeriesConfig(nx=256, ny=256, nz=41, vol=(256, 256, 64),
                             nblobs=300, tilt_step=3.0, axis=(85.3, 85.3),
                             shift_max=8.0, snr=1.0, int16=True,
                             seeds=(21, 22, 23))
    return synth.generate(cfg)


def _run_cli(stack, tlt, out, *extra):
    from markerfree_amd import fileio
    print(f"markerfree -i {stack} -o {out} {' '.join(extra)}", flush=True)
    r = subprocess.run([str(fileio.CLI_PATH), "-i", str(stack), "-o", str(out),
                        "-a", str(tlt), "-p", "10", *extra],
                       capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("descending", [False, True])
def test_cli_matches_library(ctx, mf, tmp_path, descending):
    from markerfree_amd import fileio, synth
    s = _series()
    stack, tlt = synth.write_series(s, tmp_path, "ts", descending=descending)
    out = tmp_path / "ts_ali.mrc"
    rc, log = _run_cli(stack, tlt, out, "-g", "0,0,0,0,64,1,0", "-s", "0")
    assert rc == 0, log
    for line in ("Set positivity ...", "Linear mass normalization: done.",
                 "Update head!", "Updating finished!"):
        assert line in log, line
    assert log.splitlines()[-1].startswith("Elapsed time: ")

    # The same series through the library (file order in, ascending out).
    raw = fileio.read_mrc(stack)
    G = mf.Series(ctx, raw, fileio.read_angles(tlt))
    G.run(align_z=64, nproj=10)
    sx, sy, rot, off = G.params()
    ali = G.correct(1)
    G.close()

    got = fileio.read_mrc(out)
    np.testing.assert_array_equal(got, ali)
    h = fileio.read_header_fields(out)
    assert (h["nx"], h["ny"], h["nz"], h["mode"]) == (256, 256, 41, 2)
    mid = ali[20][ali[20] > -1e10].astype(np.float64)
    np.testing.assert_allclose(h["stats"][:2], [mid.min(), mid.max()])

    txt = (tmp_path / "ts_ali.txt").read_text().splitlines()
    assert txt[0] == "# Markerfree Alignment"
    assert txt[1] == "# RawSize = 256 256 41"
    rows = np.array([[float(v) for v in l.split()] for l in txt[3:44]])
    np.testing.assert_array_equal(rows[:, 0], np.arange(41))
    np.testing.assert_allclose(rows[:, 1], rot, atol=1e-4)
    np.testing.assert_allclose(rows[:, 2], sx, atol=1e-3)
    np.testing.assert_allclose(rows[:, 3], sy, atol=1e-3)
    np.testing.assert_allclose(rows[:, 4], np.sort(s.tilts), atol=5e-3)
    assert txt[44] == f"angleoffset: {off:f}"
    assert txt[45] == "Output Bin: 1"
    assert txt[46].startswith("Elapsed time: ")

    if descending:
        # -s 1 writes the .xf in the file's (descending) order.
        rc, log = _run_cli(stack, tlt, tmp_path / "xf.mrc", "-g",
                           "0,0,0,0,64,1,0", "-s", "1")
        assert rc == 0, log
        xf = np.array([[float(v) for v in l.split()] for l in
                       (tmp_path / "xf.xf").read_text().splitlines()])
        assert xf.shape == (41, 6)
        i = 40 - np.arange(41)
        th = np.deg2rad(rot[i].astype(np.float64))
        c, sn = np.cos(th), np.sin(th)
        np.testing.assert_allclose(xf[:, 0], c, atol=6e-4)
        np.testing.assert_allclose(xf[:, 4], -(c * sx[i] + sn * sy[i]),
                                   atol=2e-3)


@pytest.mark.gpu
def test_cli_outbin_and_alignz_zero(ctx, tmp_path):
    """OutBin 2 sums 2x2 blocks into a (X/2)/2*2 wide stack; AlignZ 0 skips
    the alignment (the reference crashes there) and writes no parameter
    file."""
    from markerfree_amd import fileio, synth
    s = _series()
    stack, tlt = synth.write_series(s, tmp_path, "b")
    rc, log = _run_cli(stack, tlt, tmp_path / "b2.mrc", "-g",
                       "0,85.3,0,0,64,2,0", "-s", "0")
    assert rc == 0, log
    assert fileio.read_mrc(tmp_path / "b2.mrc").shape == (41, 128, 128)
    assert "New tilt axis" in log          # PA still refines a given axis
    rc, log = _run_cli(stack, tlt, tmp_path / "z.mrc", "-g",
                       "0,0,0,0,0,1,0", "-s", "0")
    assert rc == 0, log
    assert fileio.read_mrc(tmp_path / "z.mrc").shape == (41, 256, 256)
    assert not (tmp_path / "z.txt").exists()
