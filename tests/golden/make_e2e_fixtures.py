"""Writes the end-to-end oracle fixtures e2e_<preset>.json.

For one benchmark workload (preset 'small' = 512^2 x 41 SNR 0.1, AlignZ 100;
'single' = 1024^2 x 61 int16 SNR 0.05, AlignZ 300) it runs the oracle's
Process::DoIt (orc_run) on the generated series and on K copies whose pixels
carry a 1e-6 relative Gaussian perturbation (seeds 1..K), and records:

- a checksum of the generated stack (so a test can prove it regenerated the
  same input);
- the oracle's final parameters (axis, offset, per-tilt shifts) and its stage
  trace;
- the oracle's own spread: for each result the largest difference between
  the unperturbed and a perturbed run.  A second implementation that differs
  from the oracle only by rounding cannot be expected to agree with it more
  closely than the oracle agrees with itself.

Usage: python tests/golden/make_e2e_fixtures.py small [K]
Run time (8 threads): about 1 min per oracle run for 'small', 13 min for
'single'.
"""
import json
import sys
import time
import zlib
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from markerfree_amd import synth  # noqa: E402

ALIGN_Z = {"small": 100, "single": 300}


def checksum(stack):
    return {"sum": float(stack.astype(np.float64).sum()),
            "crc32": zlib.crc32(np.ascontiguousarray(stack).tobytes())}


def run(stack, tilts, align_z):
    O = oracle.Series(stack, tilts)
    t0 = time.time()
    O.run(align_z=align_z, nproj=10)
    sx, sy, rot, off = O.params()
    return {"axis": float(rot[0]), "offset": off, "sx": sx.tolist(),
            "sy": sy.tolist(), "trace": O.trace(),
            "seconds": round(time.time() - t0, 1)}


def main():
    preset = sys.argv[1]
    k = int(sys.argv[2]) if len(sys.argv) > 2 else 4
    s = synth.generate(synth.PRESETS[preset])
    az = ALIGN_Z[preset]
    base = run(s.stack, s.tilts, az)
    print("base", base["axis"], base["offset"], base["seconds"], "s",
          flush=True)
    perts = []
    for seed in range(1, k + 1):
        noise = np.random.default_rng(seed).standard_normal(s.stack.shape)
        st = (s.stack * (1 + 1e-6 * noise)).astype(np.float32)
        r = run(st, s.tilts, az)
        perts.append(r)
        print("seed", seed, r["axis"], r["offset"], r["seconds"], "s",
              flush=True)
    bx, by = np.array(base["sx"]), np.array(base["sy"])

    def diffs(r):
        dx, dy = np.array(r["sx"]) - bx, np.array(r["sy"]) - by
        return {"axis": abs(r["axis"] - base["axis"]),
                "offset": abs(r["offset"] - base["offset"]),
                "shift_max": float(max(np.abs(dx).max(), np.abs(dy).max())),
                "shift_rms": float(np.sqrt(np.mean(dx ** 2 + dy ** 2)))}

    d = [diffs(r) for r in perts]
    band = {key: max(x[key] for x in d) for key in d[0]}
    out = {
        "preset": preset, "align_z": az, "nproj": 10,
        "stack": checksum(s.stack),
        "truth": {"axis": s.axis},
        "oracle": {key: base[key] for key in ("axis", "offset", "sx", "sy",
                                              "trace")},
        "perturbation": {"relative": 1e-6, "seeds": list(range(1, k + 1)),
                         "axis": [r["axis"] for r in perts],
                         "offset": [r["offset"] for r in perts],
                         "diffs": d},
        "band": band,
    }
    path = Path(__file__).with_name(f"e2e_{preset}.json")
    path.write_text(json.dumps(out, indent=1))
    print("band", band)


if __name__ == "__main__":
    main()
