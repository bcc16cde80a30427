"""Benchmark: tilt series aligned per hour on one MI355X.

Workload: the single-series configuration -- a 1024x1024x61 tilt series
(-60..60 deg in 2 deg steps, 3000-blob phantom, axis U(80,95), shifts
U(-40,40), SNR 0.05, stored as int16 counts), projection-matching thickness
300, NPROJ 10, aligned end to end by the plain per-series GPU form (the
reference's Process::DoIt sequence) with the output stack downloaded.  One
step = one series: preprocessing, 3 coarse sweeps + axis searches, offset
scan, projection matching (3 ProjAlignOnce rounds + 4 axis searches) and the
final corrected stack in host memory.  Raw stacks are resident in HBM before
the timed region starts (each step gets its own series).  Each series' stack
downloads over PCIe while the next series aligns (--serial: one series after
the other); latency is from a series' start to its stack in host memory.

Prints one JSON line with the series/hour, the backprojection kernel's
roofline and a CPU baseline (the oracle on a bounded sample, extrapolated).

Roofline of the dominant kernel, timed with HIP events around every 8th of
its launches in the timed steps.  Default (fused) mode: bprp_fused, which
backprojects and reprojects without a slab in HBM; its algorithmic bytes per
target are what it must move, the n neighbour images it reads and the
reprojection it writes, 4*(n+1)*bx*by.  That kernel is bound by VALU issue,
not HBM: 'valu' reports its issue rate from the committed counter summary
(SQ_INSTS_VALU over SQ_BUSY_CYCLES), and 'gvoxel_per_s' the backprojection
it delivers.  --mode materialized times bp_slab, whose algorithmic bytes are
the slab it writes plus the sinogram rows it reads, 4*(2bx*T*by) + 4*n*bx*by;
--mode slab4 / slab2 time bp_slab_rows, the same slab written four (two)
image rows per voxel.
'traffic' is the FETCH_SIZE + WRITE_SIZE counter reading per launch from the
committed summaries.

'value' counts series whose raw stacks are resident in HBM when the timed
region starts; 'pcie_inclusive' times as many series again from a
page-locked host stack to parameters and the aligned stack in host memory
(the served form's unit).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
TIME_EVERY = 8  # HIP events around every 8th launch of the timed kernel
CALIB_BYTES = 1 << 30


def profile_value(name, key):
    """A 'key: value' line of a committed profile summary
    (profiles/<name>), or None when absent."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def cpu_model():
    """The host CPU's model name with its family/model numbers (virtualised
    hosts may report only a brand name)."""
    info = {}
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if ":" in line:
                k, v = (t.strip() for t in line.split(":", 1))
                info.setdefault(k, v)
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    if "cpu family" in info and "model" in info:
        name += f" (family {info['cpu family']}, model {info['model']})"
    return name


def pa_sweep_bytes(mf, tilts, bx, by, tb, nproj, fused):
    """Algorithmic bytes of the dominant kernel over one projection-matching
    sweep, with n from FindProjRange under the sweep's skip pattern: per
    target 4*(n+1)*bx*by for the fused kernel (n neighbour images read, the
    reprojection written), 4*(2bx*tb*by) + 4*n*bx*by for bp_slab (the slab
    written, the sinogram rows read)."""
    nz = len(tilts)
    z0 = int(np.argmin(np.abs(tilts)))
    skip = np.ones(nz, np.uint8)
    skip[z0] = 0
    total, voxels, targets = 0, 0, 0
    order = []
    for i in range(1, nz):
        for t in (z0 + i, z0 - i):
            if 0 <= t < nz:
                order.append(t)
    for t in order:
        lo, hi = mf.find_proj_range(tilts, skip, t, nproj)
        n = hi - lo + 1 if lo >= 0 else 0
        if fused:
            total += 4 * (n + 1) * bx * by
        else:
            total += 4 * (2 * bx * tb * by) + 4 * n * bx * by
        voxels += 2 * bx * tb * by
        targets += 1
        skip[t] = 0
    return total, voxels, targets


def cpu_baseline(series, pa_targets, pa_sweeps, n_targets=20):
    """The oracle (OpenMP C++ restatement) timed per stage on the same
    1024^2 x 61 series, one instance of each stage, and extrapolated to one
    series with the GPU run's stage counts: preprocessing, 3 coarse sweeps,
    7 axis searches, the offset scan, per projection-matching sweep one
    corrected + R-weighted stack and per target a reprojection, a
    correlation and a re-correction, and the final correction."""
    import oracle
    stack = series.stack.astype(np.float32)
    nz, ny, nx = stack.shape
    tilts = series.tilts
    t = {}

    def timed(key, fn):
        t0 = time.perf_counter()
        r = fn()
        t[key] = time.perf_counter() - t0
        return r

    O = oracle.Series(stack, tilts)
    timed("preprocess", O.preprocess)
    pre = O.stack()
    zeros = np.zeros(nz, np.float32)
    O.set_params(zeros, zeros, zeros)
    O.coarse_setup()
    timed("coarse_sweep", O.coarse_iter)
    timed("axis_search", lambda: O.tilt_axis(180.0))
    timed("offset_scan", O.find_offset)
    axis = float(O.params()[2][0])
    timed("correct_image", lambda: oracle.correct_proj(
        pre[nz // 2], 0.0, 0.0, axis, False, True, 1, True))
    z0 = int(np.argmin(np.abs(tilts)))
    per = []
    for k in np.linspace(1, nz // 2, n_targets).astype(int):
        tg = min(z0 + k, nz - 1)
        skip = np.ones(nz, np.uint8)
        skip[max(0, z0 - k + 1): z0 + k] = 0
        skip[tg] = 1
        t0 = time.perf_counter()
        rp = oracle.reproject(pre, 300, tilts, skip, tg, 10, acc64=False)
        oracle.pa_measure(rp, pre[tg])
        per.append(time.perf_counter() - t0)
    t_target = float(np.mean(per)) + t["correct_image"]
    per_series = (t["preprocess"] + 3 * t["coarse_sweep"] +
                  7 * t["axis_search"] + t["offset_scan"] +
                  pa_sweeps * nz * t["correct_image"] + pa_targets * t_target +
                  nz * t["correct_image"])
    sampled = sum(t.values()) + sum(per)
    return {"value": round(3600.0 / per_series, 3), "unit": "series/hour",
            "cores": oracle.num_threads(), "cpu_model": cpu_model(),
            "kind": "port",
            "sample": (f"oracle on the same 1024x1024x61 T=300 series, one of "
                       f"each stage ({sampled:.1f} s): preprocess "
                       f"{t['preprocess']:.2f} s, coarse sweep "
                       f"{t['coarse_sweep']:.2f} s, axis search "
                       f"{t['axis_search']:.2f} s, offset scan "
                       f"{t['offset_scan']:.2f} s, corrected image "
                       f"{t['correct_image']:.3f} s, PA target "
                       f"{np.mean(per):.2f} s (mean of {len(per)}); "
                       f"extrapolated with the GPU run's {pa_sweeps:.0f} "
                       f"sweeps and {pa_targets} targets per series to "
                       f"{per_series:.0f} s/series")}


MIX_SIZES = [(1024, 1024), (2048, 2048), (5760, 4092)]


def mixture(synth, n, seed=7):
    """Configurations of the batched workload's mixture: sizes uniform from
    {1024^2, 2048^2, 5760x4092}, tilt counts U{35..61}, ranges +-U[50, 60]
    degrees, thickness {150, 300}, NPROJ 10, float32, seeded per series.
    Phantom and noise as the single-series workload (3000 blobs per 1024^2 of
    area, at most 12000; Gaussian noise at SNR 0.05; axis U(80, 95)); shifts
    U(-40, 40) px, U(-80, 80) px at 5760x4092."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx, ny = MIX_SIZES[int(rng.integers(0, len(MIX_SIZES)))]
        nz = int(rng.integers(35, 62))
        rmax = float(rng.uniform(50, 60))
        cfg = synth.SeriesConfig(
            nx=nx, ny=ny, nz=nz, vol=(nx, ny, 300),
            nblobs=min(3000 * nx * ny // (1 << 20), 12000),
            tilt_min=-rmax, tilt_step=2 * rmax / (nz - 1), axis=(80.0, 95.0),
            shift_max=80.0 if nx > 2048 else 40.0, snr=0.05,
            seeds=(100 + i, 200 + i, 300 + i))
        out.append((cfg, int(rng.choice([150, 300]))))
    return out


def generate_all(synth, specs, label):
    t0 = time.perf_counter()
    out = []
    for k, (cfg, tb) in enumerate(specs):
        out.append((synth.generate(cfg), tb))
        if (k + 1) % 8 == 0 or k + 1 == len(specs):
            print(f"{label}: generated {k + 1}/{len(specs)} series "
                  f"({time.perf_counter() - t0:.0f} s)", file=sys.stderr,
                  flush=True)
    return out


def _sample_sum(a):
    # A sampled checksum of an aligned stack (every 61st row, 67th column).
    return float(a[:, ::61, ::67].sum(dtype=np.float64))


def bench_batched(mf, synth, args):
    """The streaming batched form (mf_batch_run_each) on the mixture: every
    series aligned, corrected and its aligned stack handed to the caller in
    host memory, 'concurrency' series at a time, largest first; against the
    same series one at a time (the plain per-series form).  Raw stacks are
    resident in HBM before each timed pass."""
    if args.mix == "single":
        # The single-series workload (1024x1024x61, T 300), 'batch' series.
        one = synth.generate(synth.PRESETS["single"])
        mix = [(one, 300)] * args.batch
    else:
        mix = generate_all(synth, mixture(synth, args.batch), "batched")
    n = len(mix)
    opts = [mf.default_options(align_z=tb, nproj=10) for _, tb in mix]
    ctx = mf.Context(0)
    # Warm-up (tables, transform plans, page-locked output blocks): one
    # series of each size.
    warm, seen = [], set()
    for i, (s, _) in enumerate(mix):
        if s.stack.shape[1:] not in seen:
            seen.add(s.stack.shape[1:])
            warm.append(i)
    if args.passes == "plain":
        warm = []   # the trace of the pass alone (first kernel = its start)
    W = [mf.Series(ctx, mix[i][0].stack, mix[i][0].tilts) for i in warm]
    mf.batch_run_each(W, [opts[i] for i in warm],
                      concurrency=args.concurrency, correct=True)
    for g in W:
        g.close()

    def run(concurrency, label):
        pool = [mf.Series(ctx, s.stack, s.tilts) for s, _ in mix]
        ctx.synchronize()
        ctx.memory(reset_peak=True)
        sums, t_done = [None] * n, [None] * n
        count = [0]

        def on_done(i, st, a):
            sums[i] = _sample_sum(a)
            t_done[i] = time.perf_counter()
            count[0] += 1
            if count[0] % 8 == 0 or count[0] == n:
                print(f"{label}: {count[0]}/{n} series aligned "
                      f"({t_done[i] - t0:.1f} s)", file=sys.stderr,
                      flush=True)

        t0 = time.perf_counter()
        mf.batch_run_each(pool, opts, concurrency=concurrency, correct=True,
                          on_done=on_done)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        peak = ctx.memory()[1]
        params = [g.params() for g in pool]
        for g in pool:
            g.close()
        return dt, params, sums, peak

    if args.passes == "plain":
        # The plain pass alone (for a device busy-time trace).
        t_plain, _, _, peak = run(1, "plain")
        print(json.dumps({"plain_pass_seconds": round(t_plain, 3),
                          "series": n, "peak_hbm_gb": round(peak / 1e9, 1)}),
              flush=True)
        ctx.close()
        return
    t_plain, p_plain, s_plain, _ = run(1, "plain")
    t_batch, p_batch, s_batch, peak = run(args.concurrency, "batched")
    same = s_plain == s_batch and all(
        all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))
        for a, b in zip(p_plain, p_batch))
    ctx.close()
    counts = {f"{nx}x{ny}": sum(1 for s, _ in mix
                                if s.stack.shape[1:] == (ny, nx))
              for nx, ny in MIX_SIZES}
    out = {
        "metric": "tilt series aligned/hour (batched form)",
        "value": round(3600.0 * n / t_batch, 2), "unit": "series/hour",
        "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
        "data": ("synthetic (seeded 3000-blob phantom, int16 counts)"
                 if args.mix == "single" else
                 "synthetic mixture (seeded per series)"),
        "config": {"workload": (
                       f"{n} series 1024x1024x61, T 300, NPROJ 10 (the "
                       "single-series workload)" if args.mix == "single" else
                       f"{n} series, sizes {{1024^2, 2048^2, 5760x4092}}, "
                       "tilts U{35..61}, range +-U[50,60], T {150,300}, "
                       "NPROJ 10, float32"),
                   "concurrency": args.concurrency, "series_per_size": counts,
                   "sizes": [[s.stack.shape[2], s.stack.shape[1],
                              s.stack.shape[0], tb] for s, tb in mix]},
        "seconds": round(t_batch, 2),
        "plain_same_mix_seconds": round(t_plain, 2),
        "plain_same_mix_series_per_hour": round(3600.0 * n / t_plain, 2),
        "speedup_vs_plain": round(t_plain / t_batch, 3),
        "peak_hbm_gb": round(peak / 1e9, 1),
        "results_identical_to_plain": bool(same),
    }
    print(json.dumps(out), flush=True)


def bench_served(mf, synth, args):
    """The served form: series drawn (seeded) from a pool of base series of
    the batched mixture arrive as a Poisson stream at each rate.  Latency
    runs from the scheduled arrival to the parameters and the aligned stack
    in host memory; a submission the full queue refuses waits for the oldest
    outstanding request and is retried, its wait counted in its latency.
    The plain per-series form is the same server with one worker and one
    request at a time."""
    gen = generate_all(synth, mixture(synth, args.pool, seed=11), "served")
    # The base stacks sit in page-locked memory, read in place by the server
    # (mf_submit_borrowed), as a service reading files into its own staging
    # buffers would hand them over.
    base = []
    for s, tb in gen:
        pin = mf.pinned_empty(s.stack.shape)
        pin[:] = s.stack
        base.append((pin, s.tilts, tb))
    del gen
    opts = [mf.default_options(align_z=tb, nproj=10) for _, _, tb in base]
    ctx = mf.Context(0)
    rows = []
    with mf.Server(ctx, workers=1, queue_cap=2) as sv:
        for (st, tl, _), o in zip(base, opts):         # warm-up
            rid = sv.submit(st, tl, o, borrow=True)
            sv.wait(rid)
            sv.release(rid)
        plain = []
        for (st, tl, _), o in zip(base, opts):
            t = time.perf_counter()
            rid = sv.submit(st, tl, o, borrow=True)
            sv.wait(rid)
            plain.append(time.perf_counter() - t)
            sv.release(rid)
    mean_plain = float(np.mean(plain))
    print(json.dumps({"served_plain_s": [round(x, 3) for x in plain]}),
          file=sys.stderr, flush=True)
    with mf.Server(ctx, workers=args.concurrency,
                   queue_cap=args.queue) as sv:
        budget = sv.stats()["budget"]
        charges = {f"{st.shape[2]}x{st.shape[1]}x{st.shape[0]}":
                   round(mf.series_hbm_bytes(st.shape[2], st.shape[1],
                                             st.shape[0], o) / 1e9, 2)
                   for (st, _, _), o in zip(base, opts)}
        for lam in [float(r) for r in args.rates.split(",")]:
            rng = np.random.default_rng(int(lam * 100))
            draw = rng.integers(0, len(base), args.requests)
            sched = np.cumsum(rng.exponential(1.0 / lam, args.requests))
            t0 = time.perf_counter()
            c0 = sv.clock()
            ids, lat, done = [], [], []

            def collect(block):
                # Results in submission order, each released as soon as it
                # is read (the server holds a result until released).
                while len(done) < len(ids):
                    r = ids[len(done)]
                    if not block and sv.poll(r) not in (mf.REQ_DONE,
                                                        mf.REQ_FAILED):
                        return
                    sv.wait(r)
                    res = sv.fetch(r, aligned=False)
                    sv.release(r)
                    lat.append(res.t_done - (c0 + sched[len(done)]))
                    done.append(res.t_done)
                    if len(done) % 50 == 0:
                        print(f"rate {lam}: completed {len(done)}/"
                              f"{args.requests}", file=sys.stderr, flush=True)
                    if block:
                        return

            for k in range(args.requests):
                collect(False)
                dt = t0 + sched[k] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                st, tl, _ = base[draw[k]]
                while True:
                    try:
                        ids.append(sv.submit(st, tl, opts[draw[k]],
                                             borrow=True))
                        break
                    except mf.MFError as e:
                        if e.code != mf.MF_EBUSY:
                            raise
                        collect(True)   # the oldest outstanding request
                if (k + 1) % 50 == 0:
                    print(f"rate {lam}: submitted {k + 1}/{args.requests}",
                          file=sys.stderr, flush=True)
            while len(done) < len(ids):
                collect(True)
            lat = np.array(lat)
            span = max(done) - c0
            rows.append({"rate_per_s": lam, "requests": len(done),
                         "throughput_series_per_hour":
                             round(3600.0 * len(done) / span, 1),
                         "p50_ms": round(1000 * float(np.percentile(lat, 50)), 1),
                         "p99_ms": round(1000 * float(np.percentile(lat, 99)), 1)})
            print(json.dumps({"served_rate": rows[-1]}), file=sys.stderr,
                  flush=True)
    ctx.close()
    best = max(r["throughput_series_per_hour"] for r in rows)
    out = {
        "metric": "tilt series aligned/hour + p50/p99 latency (served form)",
        "value": best, "unit": "series/hour", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32",
        "data": f"synthetic: {len(base)} seeded base series of the batched "
                "mixture, requests drawn from them",
        "config": {"workload": "Poisson arrivals drawn from the mixture "
                               "{1024^2, 2048^2, 5760x4092}, tilts "
                               "U{35..61}, T {150,300}, NPROJ 10",
                   "workers": args.concurrency, "queue_cap": args.queue,
                   "base_sizes": [[st.shape[2], st.shape[1], st.shape[0], tb]
                                  for st, _, tb in base]},
        "admission": {"budget_gb": round(budget / 1e9, 1),
                      "charge_gb": charges},
        "plain_seconds_per_series": [round(x, 3) for x in plain],
        "plain_series_per_hour": round(3600.0 / mean_plain, 1),
        "plain_p50_ms": round(1000 * float(np.percentile(plain, 50)), 1),
        "plain_p99_ms": round(1000 * float(np.percentile(plain, 99)), 1),
        "rates": rows,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["materialized", "fused", "slab4",
                                       "slab2", "fused_exact"],
                    default="fused",
                    help="fused: the default fused form (mode 8); "
                         "fused_exact: the fused form bit-identical to the "
                         "reference's evaluation order (mode 1)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serial", action="store_true",
                    help="plain form without overlap: each series' download "
                         "finishes before the next series starts")
    ap.add_argument("--form", choices=["plain", "batched", "served"],
                    default="plain",
                    help="plain: the per-series form (the reported metric); "
                         "batched / served: the multi-series forms")
    ap.add_argument("--batch", type=int, default=64,
                    help="batched form: series in the mixture")
    ap.add_argument("--mix", choices=["mixture", "single"],
                    default="mixture",
                    help="batched form: the size mixture, or 'batch' copies "
                         "of the single-series workload")
    ap.add_argument("--concurrency", type=int, default=3)
    ap.add_argument("--passes", choices=["both", "plain"], default="both",
                    help="batched form: 'plain' runs the one-at-a-time pass "
                         "only")
    ap.add_argument("--requests", type=int, default=200,
                    help="served form: requests per rate")
    ap.add_argument("--pool", type=int, default=12,
                    help="served form: base series requests are drawn from")
    ap.add_argument("--queue", type=int, default=8,
                    help="served form: requests that may wait")
    ap.add_argument("--rates", default="0.5,1,2,4,8")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the alignment path runs one process "
              "on one MI355X (multi-GPU sharding is described in DESIGN.md, "
              "not measured)", file=sys.stderr)
        sys.exit(2)

    import markerfree_amd as mf
    from markerfree_amd import synth

    if args.form == "batched":
        return bench_batched(mf, synth, args)
    if args.form == "served":
        return bench_served(mf, synth, args)

    cfg = synth.PRESETS["single"]
    series = synth.generate(cfg)
    mode = {"materialized": 0, "fused": 8, "slab4": 6, "slab2": 7,
            "fused_exact": 1}[args.mode]
    fused = mode in (1, 8)
    align_z, nproj = 300, 10
    ctx = mf.Context(0)
    n = args.warmup + args.steps
    pool = [mf.Series(ctx, series.stack, series.tilts) for _ in range(n)]
    opts = mf.default_options(align_z=align_z, nproj=nproj, reproj_mode=mode)

    # The aligned stack of every step lands in one page-locked host buffer,
    # as a service reusing its output buffers would (DMA at the PCIe rate).
    aligned = mf.pinned_empty(series.stack.shape)

    def step(s):
        s.run(opts)
        out = s.correct(1, out=aligned)
        # The series' work buffers go back to the context (its block cache),
        # as in the batched and served forms, for the next series to reuse.
        s.trim()
        return out

    # Counter calibration: one device-to-device copy of a known size (1 GiB,
    # above the 256 MB last-level cache) outside the timed region; under
    # `rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` its dispatch gives the
    # counters' reading for known bytes.
    calib = [ctx.empty((CALIB_BYTES // 4,)) for _ in range(2)]
    calib[1].copy_from(calib[0])
    ctx.synchronize()
    for a in calib:
        a.free()

    for s in pool[:args.warmup]:
        step(s)
    ctx.synchronize()
    kname = {0: "bp_slab", 1: "bprp_fused", 8: "bprp_fused"}.get(
        mode, "bp_slab_rows")
    # HIP events around every TIME_EVERY-th backprojection launch only (each
    # timed launch costs two stream markers, a few microseconds of idle
    # device; sampling keeps that out of the series time).
    ctx.timing(True, only=kname, every=TIME_EVERY)
    ctx.timing_reset()
    lat = []

    def finish(s, ts):
        # Latency: from the start of the series' run to its aligned stack in
        # host memory.
        s.wait()
        lat.append(time.perf_counter() - ts)

    t0 = time.perf_counter()
    if args.serial:
        for s in pool[args.warmup:]:
            ts = time.perf_counter()
            step(s)
            lat.append(time.perf_counter() - ts)
    else:
        # Each series' aligned stack downloads into host memory on its own
        # stream while the next series aligns (mf_correct_async); its
        # workspaces go back to the block cache once the download is done,
        # before the next download starts into the same host buffer.
        prev = None
        for s in pool[args.warmup:]:
            ts = time.perf_counter()
            s.run(opts)
            if prev is not None:
                prev[1].join()
                prev[0].trim()
            s.correct(1, out=aligned, wait=False)
            th = threading.Thread(target=finish, args=(s, ts))
            th.start()
            prev = (s, th)
        if prev is not None:
            prev[1].join()
            prev[0].trim()
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    k_ms, k_timed = ctx.kernel_time(kname)
    k_n = ctx.kernel_count(kname)
    ctx.timing(False)

    s_last = pool[-1]
    tilts_final = s_last.tilts()
    bx = by = 1024
    sweep_bytes, sweep_vox, targets = pa_sweep_bytes(
        mf, tilts_final, bx, by, align_z, nproj, fused=fused)
    sweeps = k_n / max(args.steps, 1) / max(targets, 1)
    avg_ms = k_ms / max(k_timed, 1)
    bytes_per_launch = sweep_bytes / targets
    vox_per_launch = sweep_vox / targets
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    for s in pool:
        s.close()
    # As many series again, each from its raw stack in page-locked host
    # memory to the parameters and the aligned stack in host memory (the
    # upload inside the timing).
    raw = mf.pinned_empty(series.stack.shape)
    raw[:] = series.stack
    pcie = []
    tp0 = time.perf_counter()
    if args.serial:
        for _ in range(args.steps):
            tp = time.perf_counter()
            s = mf.Series(ctx, raw, series.tilts)
            step(s)
            s.params()
            pcie.append(time.perf_counter() - tp)
            s.close()
    else:
        # Pipelined: the next series' upload and the previous series'
        # download run over PCIe while a series aligns.  Latency is from the
        # start of a series' upload to its aligned stack in host memory.
        def new_series():
            tp = time.perf_counter()
            s = mf.Series(ctx, None, series.tilts, shape=series.stack.shape)
            s.upload(0, raw)
            return s, tp

        def finish_pcie(s, tp):
            s.wait()
            pcie.append(time.perf_counter() - tp)

        nxt = new_series() if args.steps > 0 else None
        prev = None
        for k in range(args.steps):
            s, tp = nxt
            if k + 1 < args.steps:
                nxt = new_series()
            s.run(opts)
            s.params()
            if prev is not None:
                prev[1].join()
                prev[0].close()
            s.correct(1, out=aligned, wait=False)
            th = threading.Thread(target=finish_pcie, args=(s, tp))
            th.start()
            prev = (s, th)
        if prev is not None:
            prev[1].join()
            prev[0].close()
    ctx.synchronize()
    pcie_elapsed = time.perf_counter() - tp0
    ctx.close()

    value = 3600.0 * args.steps / elapsed
    out = {
        "metric": "tilt series aligned/hour + p50/p99 latency; backproj "
                  "GVoxel/s vs HBM roofline",
        "value": round(value, 2),
        "unit": "series/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 2),
        "latency_ms": {"p50": round(1000 * float(np.percentile(lat, 50)), 2),
                       "p99": round(1000 * float(np.percentile(lat, 99)), 2)},
        "pcie_inclusive": {
            "series_per_hour": round(3600.0 * len(pcie) / pcie_elapsed, 2),
            "p50_ms": round(1000 * float(np.percentile(pcie, 50)), 2),
            "p99_ms": round(1000 * float(np.percentile(pcie, 99)), 2),
            "series": len(pcie),
            "unit": "raw stack in page-locked host memory to parameters and "
                    "aligned stack in host memory" + (
                        "" if args.serial else
                        "; uploads and downloads overlapped with the "
                        "alignment of other series, latency from upload "
                        "start")},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 3000-blob phantom, int16 counts)",
        "config": {"workload": "single series 1024x1024x61, plain per-series "
                               "GPU form" + (
                                   "" if args.serial else
                                   ", each download overlapped with the "
                                   "next series"),
                   "align_z": align_z, "nproj": nproj,
                   "reproj_mode": args.mode, "global_batch": 1},
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "avg_launch_ms": round(avg_ms, 4),
            "algorithmic_bytes_per_launch": int(bytes_per_launch),
            "gvoxel_per_s": round(vox_per_launch / (avg_ms * 1e-3) / 1e9, 1),
            "launches": int(k_n),
            "timed_launches": int(k_timed),
            "pa_sweeps_per_series": round(sweeps, 2),
        },
    }
    prof = f"traffic_{kname}_series_1024x1024x61_T300.md"
    tr = profile_value(prof, "traffic_bytes_per_launch")
    out["roofline"]["traffic"] = int(tr) if tr is not None else None
    if mode == 8:
        # The committed counter summary is of the default mode's kernel.
        cnt = "counters_bprp_free_series_1024x1024x61_T300.md"
        frac = profile_value(cnt, "valu_issue_frac")
        clock = profile_value(cnt, "clock_ghz")
        out["roofline"]["valu"] = {
            "bound": "valu issue",
            "frac": frac,
            "peak": "1024 SIMDs x clock / 4 wave-instructions/s",
            "clock_ghz": clock,
            "source": "profiles/" + cnt,
        }
        out["roofline"]["note"] = (
            "the fused kernel reads n neighbour images and writes one; its "
            "time goes to VALU issue ('valu'), so its HBM frac is low by "
            "construction")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(series,
                                           int(k_n / max(args.steps, 1)),
                                           sweeps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
