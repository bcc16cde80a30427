"""Autotuner for the launch shapes of the line-transform kernels.

The correlation kernels (row transforms, column transforms with the spectral
product) and the fused correction + R-weighting kernel take their lines per
workgroup and workgroup size from tunables keyed by kernel family and shape
(``mf_tune_set``; see ``include/markerfree/mf_api.h``).  This tool times the
candidate shapes on the device with HIP events, keeps the fastest by
coordinate descent, and writes the result as a JSON object that
``Context(tuning=...)`` and ``mf_tune_load`` apply.  Keys it does not write
keep the library's built-in heuristic.

    python -m markerfree_amd.autotune --size 1024x1024 --nz 61 --out tuned.json
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import Context, Series

LINES = (1, 2, 4, 8)
THREADS = (64, 128, 256)


def _time(ctx, name, fn, reps):
    fn()
    ctx.synchronize()
    ctx.timing(True, only=name)
    ctx.timing_reset()
    for _ in range(reps):
        fn()
    ctx.synchronize()
    ms, n = ctx.kernel_time(name)
    ctx.timing(False)
    return ms / max(n, 1)


def _descend(ctx, keys, name, fn, reps, log):
    """Coordinate descent over (lines, threads) of each key prefix in turn;
    returns {key: value} of the chosen shapes."""
    best = {}
    base = _time(ctx, name, fn, reps)
    log(f"  {name}: heuristic {base * 1e3:.1f} us")
    for prefix in keys:
        for param, values in (("lines", LINES), ("threads", THREADS)):
            key = prefix + param
            choice, t_best = None, base
            for v in values:
                ctx.tune_set(key, v)
                try:
                    t = _time(ctx, name, fn, reps)
                except Exception as e:  # shape rejected (LDS, limits)
                    log(f"    {key}={v}: rejected ({e})")
                    continue
                log(f"    {key}={v}: {t * 1e3:.1f} us")
                if t < t_best * 0.97:
                    choice, t_best = v, t
            if choice is None:
                ctx.tune_set(key, 0)  # keep the heuristic
            else:
                ctx.tune_set(key, choice)
                best[key] = choice
                base = t_best
    log(f"  {name}: tuned {base * 1e3:.1f} us")
    return best


def tune(size=(1024, 1024), nz=61, reps=20, log=print):
    w, h = size
    out = {}
    rng = np.random.default_rng(0)
    with Context(0, tuning=None) as ctx:
        # Projection-matching measurement: one pair of w x h images.
        a = (rng.standard_normal((h, w)) + 1).astype(np.float32)
        b = np.roll(a, (3, -5), (0, 1)).copy()
        da, db = ctx.array(a), ctx.array(b)
        pre = f"/{w}x{h}x1/"
        out.update(_descend(ctx, ["cc_rowsf" + pre, "cc_cols" + pre,
                                  "cc_rowsi" + pre], "cc",
                            lambda: ctx.pa_measure(da, db, w, h), reps, log))
        # Correction + R-weighting of one projection at that size.
        raw = (rng.random((h, w)) + 0.5).astype(np.float32)
        draw, dout = ctx.array(raw), ctx.empty((h, w))
        out.update(_descend(
            ctx, [f"corr_rw/{w}x{h}x1/"], "correct_rweight",
            lambda: ctx.correct_proj(draw, w, h, 3.0, -2.0, 86.0, False, True,
                                     1, True, dout), reps, log))
        for d in (da, db, draw, dout):
            d.free()
        # Coarse alignment: the correlation of all adjacent pairs of an
        # nz-image series at this size, one sweep per timing.
        stack = (rng.random((nz, h, w)) + 0.5).astype(np.float32)
        tilts = (-60.0 + 120.0 / (nz - 1) * np.arange(nz)).astype(np.float32)
        s = Series(ctx, stack, tilts)
        s.preprocess()
        s.coarse_setup()
        pre = f"/{w}x{h}x{nz - 1}/"
        out.update(_descend(ctx, ["cc_rowsf" + pre, "cc_cols" + pre,
                                  "cc_rowsi" + pre], "cc",
                            lambda: s.coarse_iter(), max(2, reps // 10), log))
        s.close()
        # Projection-matching sweeps: the shapes of a step that measures and
        # re-corrects two targets at once, and of the whole-stack correction
        # that starts a sweep (the one-target shapes are tuned above).  One
        # sweep per timing; the other shapes of the sweep keep their values.
        s = Series(ctx, stack, tilts)
        s.preprocess()
        s.projalign_setup(0.0, 300, 10)
        sweeps = max(2, reps // 10)
        out.update(_descend(ctx, ["cc_rowsf" + f"/{w}x{h}x2/",
                                  "cc_cols" + f"/{w}x{h}x2/",
                                  "cc_rowsi" + f"/{w}x{h}x2/"], "cc",
                            lambda: s.projalign_iter(), sweeps, log))
        out.update(_descend(ctx, [f"corr_rw/{w}x{h}x2/", f"corr_rw/{w}x{h}x{nz}/"],
                            "correct_rweight", lambda: s.projalign_iter(),
                            sweeps, log))
        s.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--size", default="1024x1024")
    ap.add_argument("--nz", type=int, default=61)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--out", default="-")
    a = ap.parse_args(argv)
    w, h = (int(v) for v in a.size.split("x"))
    res = tune((w, h), a.nz, a.reps,
               log=lambda s: print(s, file=sys.stderr, flush=True))
    text = json.dumps(res, indent=1, sort_keys=True) + "\n"
    if a.out == "-":
        sys.stdout.write(text)
    else:
        with open(a.out, "w") as f:
            f.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
