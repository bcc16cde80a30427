"""File formats and command-line behaviour of the drop-in boundary (no GPU):
MRC read/write, .rawtlt reading, .txt/.xf parameter files, option parsing,
messages and exit codes (main/opts.h, main/main.cpp, main/Process.cpp,
method/SaveParam.cpp, mrc/mrcstack.cpp)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def fio():
    from markerfree_amd import fileio
    return fileio


# ------------------------------------------------------------- .txt / .xf

def _parse_align_txt(path):
    rows, offset, out_bin, elapsed = [], None, None, None
    for line in open(path):
        if line.startswith("#"):
            continue
        if line.startswith("angleoffset:"):
            offset = float(line.split(":")[1])
        elif line.startswith("Output Bin:"):
            out_bin = int(line.split(":")[1])
        elif line.startswith("Elapsed time:"):
            elapsed = float(line.split(":")[1].split()[0])
        else:
            rows.append([float(v) for v in line.split()])
    return np.array(rows), offset, out_bin, elapsed


def test_txt_writer_reproduces_reference_example(fio, tmp_path):
    """Fed the values of the reference's own example output
    (example/TS_1_align.txt), the .txt writer reproduces the file byte for
    byte: header lines, column formats, trailer lines."""
    src = GOLDEN / "TS_1_align.txt"
    rows, offset, out_bin, elapsed = _parse_align_txt(src)
    assert rows.shape == (41, 5)
    out = tmp_path / "TS_1_align.txt"
    fio.save_txt(out, 910, 646, rows[:, 1], rows[:, 2], rows[:, 3],
                 rows[:, 4], offset, out_bin, elapsed)
    assert out.read_bytes() == src.read_bytes()


def test_xf_closed_form(fio, tmp_path):
    """SaveXf's translation terms reduce to (dx, dy) = -R(theta)^T s for the
    rotation R(theta) = [[c, -s], [s, c]] whatever the centre: the centre
    terms cancel.  Rows hold (c, s, -s, c, dx, dy) and follow the file order
    (reversed when the file's tilts descend)."""
    rng = np.random.default_rng(3)
    nz = 7
    rot = rng.uniform(80, 95, nz).astype(np.float32)
    sx = rng.uniform(-40, 40, nz)
    sy = rng.uniform(-40, 40, nz)
    for inv in (0, 1):
        path = tmp_path / f"a{inv}.xf"
        fio.save_xf(path, 1024, rot, sx, sy, inv)
        lines = path.read_text().splitlines()
        assert len(lines) == nz
        for k, line in enumerate(lines):
            assert line.endswith(" ") and len(line.split()) == 6
            i = nz - 1 - k if inv else k
            v = np.array([float(x) for x in line.split()])
            th = np.deg2rad(float(rot[i]))
            c, s = np.cos(th), np.sin(th)
            np.testing.assert_allclose(v[:4], [c, s, -s, c], atol=6e-4)
            dx = -(c * sx[i] + s * sy[i])
            dy = -(-s * sx[i] + c * sy[i])
            np.testing.assert_allclose(v[4:], [dx, dy], atol=2e-3)


@pytest.mark.parametrize("out,mode,want", [
    ("TS_1_align.mrc", 0, "TS_1_align.txt"),
    ("TS_1_align.mrc", 1, "TS_1_align.xf"),
    ("out.st", 0, "out.st.txt"),
    ("a.mrcs", 1, "a.xf"),                 # truncated at the first ".mrc"
    ("d.mrc/x.mrc", 0, "d.txt"),
    ("/tmp/run/al", 1, "/tmp/run/al.xf"),
])
def test_param_file_name(fio, out, mode, want):
    assert fio.param_path(out, mode) == want


# -------------------------------------------------------------------- MRC

def test_mrc_writer_header_and_middle_slice_stats(fio, tmp_path):
    """Header fields of InitializeHeader + UpdateHeader, read back by the
    MRC2014 offsets; statistics from the middle slice only, skipping values
    below -1e10; amax starts at FLT_MIN, so an all-negative slice reports
    FLT_MIN (quirk of the reference's UpdateHeader)."""
    nz, ny, nx = 5, 6, 8
    a = np.arange(nz * ny * nx, dtype=np.float32).reshape(nz, ny, nx)
    a[2, 0, :3] = -1e30
    path = tmp_path / "s.mrc"
    stats = fio.write_mrc_aligned(path, a)
    h = fio.read_header_fields(path)
    assert (h["nx"], h["ny"], h["nz"], h["mode"]) == (nx, ny, nz, 2)
    assert h["mxyz"] == (nx, ny, nz) and h["start"] == (0, 0, 0)
    assert h["cell"] == (1.0, 1.0, 1.0) and h["angles"] == (90.0, 90.0, 90.0)
    assert h["map"] == (1, 2, 3) and h["next"] == 0
    assert h["creatid"] == 1000 and h["nlabl"] == 0
    mid = a[2][a[2] > -1e10].astype(np.float64)
    np.testing.assert_allclose(h["stats"], [mid.min(), mid.max(), mid.mean()],
                               rtol=1e-6)
    np.testing.assert_array_equal(stats, np.array(h["stats"], np.float32))
    assert path.stat().st_size == 1024 + a.nbytes
    np.testing.assert_array_equal(fio.read_mrc(path), a)
    neg = -np.ones((3, 2, 2), np.float32)
    st = fio.write_mrc_aligned(tmp_path / "n.mrc", neg)
    assert st[1] == np.finfo(np.float32).tiny and st[0] == -1.0


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.float32, np.uint16])
def test_mrc_reader_modes_and_extended_header(fio, tmp_path, dtype):
    rng = np.random.default_rng(1)
    if dtype == np.float32:
        a = rng.standard_normal((3, 5, 7)).astype(dtype)
    else:
        info = np.iinfo(dtype)
        a = rng.integers(info.min, info.max, (3, 5, 7), endpoint=True).astype(dtype)
    path = tmp_path / "m.mrc"
    fio.write_mrc(path, a, next_bytes=bytes(range(100)))
    info = fio.mrc_info(path)
    assert (info.nx, info.ny, info.nz, info.next) == (7, 5, 3, 100)
    np.testing.assert_array_equal(fio.read_mrc(path), a.astype(np.float32))


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.float32, np.uint16])
def test_mrc_slice_reads_match_whole_read(fio, tmp_path, dtype):
    """Streaming ingest reads image ranges: every range equals the same
    images of the whole read, past an extended header; ranges outside the
    stack are refused."""
    rng = np.random.default_rng(2)
    if dtype == np.float32:
        a = rng.standard_normal((7, 6, 5)).astype(dtype)
    else:
        info = np.iinfo(dtype)
        a = rng.integers(info.min, info.max, (7, 6, 5), endpoint=True).astype(dtype)
    path = tmp_path / "s.mrc"
    fio.write_mrc(path, a, next_bytes=bytes(range(37)))
    whole = fio.read_mrc(path)
    for z0, n in ((0, 7), (0, 1), (3, 2), (6, 1), (2, 0)):
        np.testing.assert_array_equal(fio.read_mrc_slices(path, z0, n),
                                      whole[z0:z0 + n])
    for z0, n in ((5, 3), (-1, 1), (0, 8)):
        with pytest.raises(OSError, match="out of bounds"):
            fio.read_mrc_slices(path, z0, n)


def test_mrc_reader_rejects_unknown_mode_and_short_files(fio, tmp_path):
    path = tmp_path / "c.mrc"
    fio.write_mrc(path, np.zeros((2, 2, 2), np.float32))
    raw = bytearray(path.read_bytes())
    raw[12:16] = (4).to_bytes(4, "little")  # complex float
    path.write_bytes(bytes(raw))
    with pytest.raises(OSError, match="File type unknown"):
        fio.read_mrc(path)
    short = tmp_path / "short.mrc"
    fio.write_mrc(short, np.zeros((2, 4, 4), np.float32))
    short.write_bytes(short.read_bytes()[:-8])
    with pytest.raises(OSError, match="short MRC data"):
        fio.read_mrc(short)


def test_read_angles(fio, tmp_path):
    t = fio.read_angles(GOLDEN / "TS_1.rawtlt")
    assert len(t) == 41
    assert t[0] == np.float32(-40.01) and t[-1] == np.float32(39.98)
    p = tmp_path / "a.rawtlt"
    p.write_text("1.5 2.5\n3.5 x 4.5\n")   # parsing stops at the first failure
    np.testing.assert_array_equal(fio.read_angles(p), [1.5, 2.5, 3.5])
    with pytest.raises(OSError):
        fio.read_angles(tmp_path / "missing.rawtlt")


# ----------------------------------------------------- command-line errors

def _cli(*args):
    from markerfree_amd import fileio
    r = subprocess.run([str(fileio.CLI_PATH), *args], capture_output=True,
                       text=True, timeout=60)
    return r.returncode, r.stdout


def test_cli_help_and_bad_options():
    """GetOpts returns 0 after -h, which main treats as wrong input (exit
    -1 = 255); -g needs seven fields; -s must be 0 or 1; -p must be > 0."""
    rc, out = _cli("-h")
    assert rc == 255 and out.startswith("[-i INPUT FILENAME]")
    assert out.rstrip().endswith("***WRONG INPUT.")
    rc, out = _cli("-g", "0,0,0,0,300,1")
    assert rc == 255 and "***WRONG INPUT." in out
    rc, out = _cli("-s", "2")
    assert rc == 255 and out.startswith("Savemode must be 0 or 1.")
    rc, out = _cli("-p", "0")
    assert rc == 255 and out.startswith("nProj must be > 0.")
    rc, out = _cli("--bogus")
    assert rc == 255


def test_cli_missing_inputs_exit_zero(tmp_path):
    """An unreadable stack or angle file aborts with the reference's
    messages and exit code 0 (quirk Q27); options are echoed first."""
    from markerfree_amd import fileio
    rc, out = _cli("-i", str(tmp_path / "none.mrc"), "-o",
                   str(tmp_path / "o.mrc"), "-a", "x.rawtlt", "-g",
                   "0,0,0,0,300,1,0", "-p", "10", "-s", "0")
    assert rc == 0
    lines = out.splitlines()
    assert lines[:9] == ["pitch_angle = 0", "zshift = 0", "thickness = 0",
                         "offset = 0", "AlignZ = 300", "OutBin = 1",
                         "GPU = 0", "Savemode = 0", "nProj = 10"]
    assert lines[-2:] == [
        "[Abort] Input stack loading failed. Please check the file path.",
        "[Abort] DoIt() stopped due to input file loading failure."]
    stack = tmp_path / "s.mrc"
    fileio.write_mrc(stack, np.ones((3, 8, 8), np.float32))
    rc, out = _cli("-i", str(stack), "-o", str(tmp_path / "o.mrc"), "-a",
                   str(tmp_path / "missing.rawtlt"))
    assert rc == 0
    assert out.splitlines()[-2:] == [
        f"[Error] Cannot open angle file: {tmp_path / 'missing.rawtlt'}",
        "[Abort] DoIt() stopped due to angle file loading failure."]
