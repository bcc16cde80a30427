"""Batched and served forms: several series of different shapes run
concurrently on one GPU, and every result is identical to the result of the
same series run alone (deterministic kernels, no device state shared between
series).  Admission and queue limits are exercised from the host."""
import numpy as np
import pytest


def _mix():
    from markerfree_amd import synth
    shapes = [(256, 256, 41, 3.0, 85.3), (320, 224, 35, 3.5, 80.0),
              (224, 288, 29, 4.0, 92.0)]
    out = []
    for i, (nx, ny, nz, step, axis) in enumerate(shapes):
        cfg = synth.SeriesConfig(nx=nx, ny=ny, nz=nz, vol=(nx, ny, 64),
                                 nblobs=250, tilt_step=step,
                                 tilt_min=-step * (nz // 2),
                                 axis=(axis, axis), shift_max=8.0, snr=0.5,
                                 seeds=(40 + i, 50 + i, 60 + i))
        out.append(synth.generate(cfg))
    return out


OPTS = dict(align_z=64, nproj=10)


def _alone(mf, ctx, s):
    G = mf.Series(ctx, s.stack, s.tilts)
    G.run(**OPTS)
    res = G.params(), G.correct(1)
    G.close()
    return res


def _same(a, b):
    (sx, sy, rot, off), ali = a
    (sx2, sy2, rot2, off2), ali2 = b
    np.testing.assert_array_equal(sx, sx2)
    np.testing.assert_array_equal(sy, sy2)
    np.testing.assert_array_equal(rot, rot2)
    assert off == off2
    np.testing.assert_array_equal(ali, ali2)


@pytest.mark.gpu
def test_gpu_batch_equals_alone(ctx, mf):
    mix = _mix()
    alone = [_alone(mf, ctx, s) for s in mix]
    series = [mf.Series(ctx, s.stack, s.tilts) for s in mix] * 1
    mf.batch_run(series, concurrency=3, **OPTS)
    for G, ref in zip(series, alone):
        _same((G.params(), G.correct(1)), ref)
        G.close()


@pytest.mark.gpu
def test_gpu_server_mix_equals_alone(ctx, mf):
    """Requests submitted back to back overlap on the device (three
    workers); each fetched result equals the series run alone."""
    mix = _mix()
    alone = [_alone(mf, ctx, s) for s in mix]
    with mf.Server(ctx, workers=3) as sv:
        ids = [sv.submit(s.stack, s.tilts, **OPTS) for s in mix * 2]
        for rid in ids:
            assert sv.wait(rid) == mf.REQ_DONE
        for k, rid in enumerate(ids):
            r = sv.fetch(rid)
            _same(((r.shift_x, r.shift_y, r.rotate, r.offset), r.aligned),
                  alone[k % 3])
            assert r.t_submit <= r.t_start <= r.t_done
            sv.release(rid)
        st = sv.stats()
        assert st["completed"] == 6 and st["reserved"] == 0


@pytest.mark.gpu
def test_gpu_server_admission_and_limits(ctx, mf):
    """With a budget that holds one series at a time the requests run one
    after another (never two at once) and still give the same results; a
    series larger than the budget is refused at submission, and a full
    queue answers MF_EBUSY."""
    mix = _mix()
    s0 = mix[0]
    need = mf.series_hbm_bytes(256, 256, 41, **OPTS)
    ref = _alone(mf, ctx, s0)
    with mf.Server(ctx, workers=3, hbm_budget=int(need * 1.5),
                   queue_cap=2) as sv:
        ids = [sv.submit(s0.stack, s0.tilts, **OPTS) for _ in range(2)]
        with pytest.raises(mf.MFError) as e:
            for _ in range(4):      # the queue holds at most 2 waiting
                ids.append(sv.submit(s0.stack, s0.tilts, **OPTS))
        assert e.value.code == mf.MF_EBUSY
        big = np.zeros((41, 1024, 1024), np.float32)
        with pytest.raises(mf.MFError) as e:
            sv.submit(big, s0.tilts, **OPTS)
        assert e.value.code == mf.MF_ENOMEM
        for rid in ids:
            assert sv.wait(rid) == mf.REQ_DONE
        rs = [sv.fetch(rid) for rid in ids]
        for r in rs:
            _same(((r.shift_x, r.shift_y, r.rotate, r.offset), r.aligned), ref)
        spans = sorted((r.t_start, r.t_done) for r in rs)
        for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
            assert b0 >= a1 - 1e-6        # serialised by the budget


@pytest.mark.gpu
def test_gpu_pinned_host_buffers_same_results(ctx, mf):
    """A stack uploaded from page-locked memory and an aligned stack
    downloaded into it (mf_host_alloc) equal the pageable paths."""
    s = _mix()[0]
    ref = _alone(mf, ctx, s)
    pin = mf.pinned_empty(s.stack.shape)
    pin[:] = s.stack
    G = mf.Series(ctx, pin, s.tilts)
    G.run(**OPTS)
    out = mf.pinned_empty(ref[1].shape)
    got = G.correct(1, out=out)
    assert got is out
    _same(ref, (G.params(), got))
    with pytest.raises(ValueError):
        G.correct(1, out=np.empty((1, 2, 3), np.float32))
    G.close()


@pytest.mark.gpu
def test_gpu_streamed_upload_equals_whole_stack(ctx, mf):
    """Streaming ingest: a series created without images and filled chunk by
    chunk (page-locked and ordinary host buffers, uneven chunks) gives the
    same results as one created from the whole stack."""
    s = _mix()[1]
    ref = _alone(mf, ctx, s)
    nz, ny, nx = s.stack.shape
    for pinned, chunk in ((True, 8), (False, 5)):
        G = mf.Series(ctx, None, s.tilts, shape=(nz, ny, nx))
        bufs = [mf.pinned_empty((chunk, ny, nx)) if pinned
                else np.empty((chunk, ny, nx), np.float32) for _ in range(2)]
        for k, z0 in enumerate(range(0, nz, chunk)):
            n = min(chunk, nz - z0)
            b = bufs[k % 2][:n]
            b[...] = s.stack[z0:z0 + n]  # this buffer's last upload is done
            G.upload(z0, b)
        G.upload(0, None)
        with pytest.raises(mf.MFError, match="out of bounds"):
            G.upload(nz - 1, bufs[0][:2])
        G.run(**OPTS)
        _same((G.params(), G.correct(1)), ref)
        with pytest.raises(mf.MFError, match="preprocessing"):
            G.upload(0, bufs[0][:1])
        G.close()



@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 6, 0])
def test_gpu_admission_charge_bounds_device_use(ctx, mf, mode):
    """The served form admits against mf_series_hbm_bytes: the most device
    memory the context holds while one series is created, aligned, corrected
    and destroyed (the context's own lazily made tables included) stays
    within that bound, for each reprojection form and two shapes."""
    for s, opts in ((_mix()[0], dict(OPTS)),
                    (_mix()[1], dict(align_z=96, nproj=10, out_bin=2))):
        nz, ny, nx = s.stack.shape
        opts["reproj_mode"] = mode
        bound = mf.series_hbm_bytes(nx, ny, nz, **opts)
        ctx.synchronize()
        u0, _ = ctx.memory(reset_peak=True)
        G = mf.Series(ctx, s.stack, s.tilts)
        G.run(**opts)
        G.correct(opts.get("out_bin", 1))
        _, peak = ctx.memory()
        G.close()
        used = peak - u0
        print(f"{nx}x{ny}x{nz} mode {mode}: peak {used / 2**20:.1f} MiB, "
              f"bound {bound / 2**20:.1f} MiB")
        assert 0 < used <= bound, (used, bound)
        assert ctx.memory()[0] <= u0 + (1 << 20)   # released again


@pytest.mark.gpu
def test_gpu_trim_keeps_stack_and_parameters(ctx, mf):
    """mf_batch_run trims every series after its run: the series then holds
    little more than its preprocessed stack, its correction and parameters
    equal an untrimmed run's, stage calls ask for their setup again, and a
    new setup works."""
    s = _mix()[0]
    ref = _alone(mf, ctx, s)
    nz, ny, nx = s.stack.shape
    ctx.synchronize()
    u0, _ = ctx.memory()
    G = mf.Series(ctx, s.stack, s.tilts)
    mf.batch_run([G], concurrency=1, **OPTS)
    held = ctx.memory()[0] - u0
    assert held <= 4 * nx * ny * nz + (1 << 20), held
    _same((G.params(), G.correct(1)), ref)
    with pytest.raises(mf.MFError) as e:
        G.projalign_iter()
    assert e.value.code == mf.MF_ESTATE
    G.projalign_setup(0.0, OPTS["align_z"], OPTS["nproj"])
    assert np.isfinite(G.projalign_iter())
    G.close()


@pytest.mark.gpu
def test_gpu_batch_each_streams_corrected_stacks(ctx, mf):
    """The streaming batched form: per-series options (thickness and output
    binning differ), each series corrected into a worker's page-locked block
    and handed to the callback; every stack and parameter set equals the
    series run alone, one call per series, largest series started first."""
    mix = _mix()
    opts = [mf.default_options(align_z=64, nproj=10, out_bin=1),
            mf.default_options(align_z=96, nproj=10, out_bin=2),
            mf.default_options(align_z=64, nproj=10, out_bin=1)]
    alone = []
    for s, o in zip(mix, opts):
        G = mf.Series(ctx, s.stack, s.tilts)
        G.run(o)
        alone.append((G.params(), G.correct(o.out_bin)))
        G.close()
    for conc in (1, 3):
        series = [mf.Series(ctx, s.stack, s.tilts) for s in mix]
        got, calls = {}, []

        def on_done(i, st, aligned):
            calls.append(i)
            assert st == mf.MF_OK
            got[i] = aligned.copy()

        mf.batch_run_each(series, opts, concurrency=conc, correct=True,
                          on_done=on_done)
        assert sorted(calls) == [0, 1, 2]
        if conc == 1:
            sizes = [np.prod(mix[i].stack.shape) for i in calls]
            assert sizes == sorted(sizes, reverse=True)
        for i, G in enumerate(series):
            _same((G.params(), got[i]), alone[i])
            G.close()


@pytest.mark.gpu
def test_gpu_borrowed_submission_equals_alone(ctx, mf):
    """mf_submit_borrowed: the server reads the caller's stack in place (one
    page-locked, one ordinary) instead of copying it; results equal the
    series run alone, and the caller's stacks are left unchanged."""
    mix = _mix()
    alone = [_alone(mf, ctx, s) for s in mix[:2]]
    pin = mf.pinned_empty(mix[0].stack.shape)
    pin[:] = mix[0].stack
    plain = mix[1].stack.copy()
    with mf.Server(ctx, workers=2) as sv:
        ids = [sv.submit(pin, mix[0].tilts, borrow=True, **OPTS),
               sv.submit(plain, mix[1].tilts, borrow=True, **OPTS)]
        for k, rid in enumerate(ids):
            assert sv.wait(rid) == mf.REQ_DONE
            r = sv.fetch(rid)
            _same(((r.shift_x, r.shift_y, r.rotate, r.offset), r.aligned),
                  alone[k])
            sv.release(rid)
    np.testing.assert_array_equal(pin, mix[0].stack)
    np.testing.assert_array_equal(plain, mix[1].stack)


@pytest.mark.gpu
def test_gpu_failed_allocations_do_not_poison_later_calls(ctx, mf):
    """A device allocation larger than the device is refused with
    MF_ENOMEM, and the HIP error it leaves on the thread is cleared: the
    next series run on the same thread still succeeds with the same
    results."""
    s = _mix()[0]
    ref = _alone(mf, ctx, s)
    with pytest.raises(mf.MFError) as e:
        ctx.empty((1 << 40,))              # 4 TiB of device memory
    assert e.value.code == mf.MF_ENOMEM
    _same(_alone(mf, ctx, s), ref)


@pytest.mark.gpu
def test_gpu_reused_device_blocks_give_identical_results(mf):
    """A released series' device blocks stay in its context for reuse, still
    holding that series' data.  A series run after another of the same shape
    (so on reused blocks) gives bit-identical parameters and aligned stack
    to the same series in a fresh context (fresh allocations): no stage
    depends on the prior contents of its buffers."""
    from markerfree_amd import synth
    mix = _mix()
    x = mix[0]
    cfg = synth.SeriesConfig(nx=256, ny=256, nz=41, vol=(256, 256, 64),
                             nblobs=250, tilt_step=3.0, tilt_min=-60.0,
                             axis=(80.0, 80.0), shift_max=8.0, snr=0.3,
                             seeds=(7, 8, 9))
    y = synth.generate(cfg)
    with mf.Context() as fresh:
        ref = _alone(mf, fresh, x)
    with mf.Context() as c:
        _alone(mf, c, y)        # its blocks go to the context's cache
        for _ in range(2):      # and are taken again by x, twice
            _same(_alone(mf, c, x), ref)
