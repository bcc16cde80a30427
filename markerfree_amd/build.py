"""Build recipe for the HIP library, its CLI and the synthetic-series tool.

Compiles every source under ``markerfree_amd/csrc`` for gfx950 with hipcc and
links ``markerfree_amd/lib/libmarkerfree_hip.so`` plus the drop-in command
line ``markerfree_amd/bin/markerfree``.  Run ``python -m markerfree_amd.build``
or call :func:`build`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = ROOT / "_build" / "obj"
LIB = PKG / "lib" / "libmarkerfree_hip.so"
SYNTH = PKG / "lib" / "libmf_synth.so"
IO = PKG / "lib" / "libmf_io.so"
CLI = PKG / "bin" / "markerfree"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every FMA in the kernels is written explicitly so that
# index arithmetic matches the CPU oracle bit for bit.  Packed FP32 is turned
# off for all device code.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops"]
HOST = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-Wall",
        "-Wno-unused-result"]

# Sources of the shared library (the command line under csrc/cli and the
# synthetic generator are separate artefacts).
EXCLUDE = {"synth.cpp"}


def _sources():
    out = []
    for p in sorted(CSRC.iterdir()):
        if p.suffix in (".hip", ".cpp") and p.name not in EXCLUDE:
            out.append(p)
    return out


def _compile(src: Path) -> Path:
    obj = OBJ / (src.name + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        # Host-only translation units: the system C++ compiler with the HIP
        # runtime headers.
        cmd = ["g++", *COMMON, *HOST, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n"
                           f"{r.stdout}\n{r.stderr}")
    return obj


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(map(str, cmd))}\n{r.stdout}\n{r.stderr}")


def build(jobs: int | None = None, verbose: bool = False) -> None:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.parent.mkdir(parents=True, exist_ok=True)
    CLI.parent.mkdir(parents=True, exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    # Synthetic tilt-series generator: host C++ with OpenMP.
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
          f"-I{ROOT / 'include'}", str(CSRC / "synth.cpp"), "-o", str(SYNTH)])
    # File formats of the command line (host only), then the drop-in CLI
    # linked against both libraries.
    _run(["g++", *COMMON, "-Wall", "-shared", str(CSRC / "cli" / "io.cpp"),
          str(CSRC / "cli" / "residuals.cpp"), "-o", str(IO)])
    _run(["g++", *COMMON, "-Wall", str(CSRC / "cli" / "cli.cpp"), "-o",
          str(CLI), f"-L{LIB.parent}", "-lmarkerfree_hip", "-lmf_io",
          "-Wl,-rpath,$ORIGIN/../lib"])
    # The fiducial residual evaluator's command line (host only).
    _run(["g++", *COMMON, "-Wall", str(CSRC / "cli" / "cal_res.cpp"), "-o",
          str(CLI.parent / "cal_res"), f"-L{LIB.parent}", "-lmf_io",
          "-Wl,-rpath,$ORIGIN/../lib"])
    if verbose:
        print(f"built {LIB}, {SYNTH}, {IO}, {CLI}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
