"""Projection-matching sweeps replayed as a captured HIP graph
(mf_options.graphs, mf_series_set_graphs): within one setup the second sweep
is captured and later sweeps replay it, with the device-resident parameters
uploaded before each replay.  The results must be bit-identical to launching
every kernel, over several replays with other work on the device in between
(the axis searches of mf_run, another series' eager sweeps), and a new setup
must drop the graph (new angles and thickness).
"""
import numpy as np
import pytest

from conftest import flush_print


def _same_params(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(np.asarray(x), np.asarray(y))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [8, 1, 6, 0])
def test_gpu_graph_run_matches_eager(ctx, mf, cfg1, mode):
    """mf_run on the 512^2 x 41 series with and without graphs (fused,
    4-row materialised and one-row materialised reprojection): identical
    parameters, stage trace and last-sweep measurements; the graph run
    replayed at least three sweeps, the eager run none."""
    flush_print(f"graph vs eager mf_run, reproj_mode {mode}")
    out = {}
    for graphs in (0, 1):
        G = mf.Series(ctx, cfg1.stack, cfg1.tilts)
        G.run(align_z=100, nproj=10, reproj_mode=mode, graphs=graphs)
        out[graphs] = (G.params(), G.trace(), G.pa_measurements(),
                       G.graph_replays())
        G.close()
    assert out[0][3] == 0
    assert out[1][3] >= 3, out[1][3]
    _same_params(out[0][0], out[1][0])
    _same_params(out[0][1], out[1][1])
    _same_params(out[0][2], out[1][2])


@pytest.mark.gpu
def test_gpu_graph_stage_calls_replay_and_setup_drops_graph(ctx, mf, cfg1):
    """Stage by stage: six sweeps with an axis search and another series'
    eager sweep between replays, then a new setup with another thickness
    and three more sweeps.  Every sweep's maximum error and the final
    parameters equal the eager series'; replays happen only from the second
    sweep of each setup on, and kernel timing suspends them."""
    flush_print("graph stage calls vs eager")
    E = mf.Series(ctx, cfg1.stack, cfg1.tilts)
    G = mf.Series(ctx, cfg1.stack, cfg1.tilts)
    X = mf.Series(ctx, cfg1.stack, cfg1.tilts)   # interleaved eager work
    G.set_graphs(True)
    for s in (E, G, X):
        s.preprocess()
        s.set_params(np.zeros(len(cfg1.tilts)), np.zeros(len(cfg1.tilts)),
                     np.full(len(cfg1.tilts), cfg1.axis, np.float32))
        s.projalign_setup(0.0, 100, 10)
    X.projalign_setup(0.0, 120, 10)
    replays = []
    for i in range(6):
        assert E.projalign_iter() == G.projalign_iter(), i
        replays.append(G.graph_replays())
        X.projalign_iter()
        if i == 2:
            assert E.tilt_axis(10.0, 40) == G.tilt_axis(10.0, 40)
    # Sweep 1 eager, sweep 2 captured and launched, 3..6 replayed.
    assert replays == [0, 1, 2, 3, 4, 5], replays
    _same_params(E.params(), G.params())
    ctx.timing(True)
    assert E.projalign_iter() == G.projalign_iter()
    ctx.timing(False)
    assert G.graph_replays() == 5
    for s in (E, G):
        s.projalign_setup(0.0, 80, 10)
    for i in range(3):
        assert E.projalign_iter() == G.projalign_iter(), ("setup 2", i)
    assert G.graph_replays() == 7
    _same_params(E.params(), G.params())
    _same_params(E.pa_measurements(), G.pa_measurements())
    for s in (E, G, X):
        s.close()
