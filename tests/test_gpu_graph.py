"""The persistent sweep launch replayed from a captured HIP graph (dqmc_batch_use_graph) gives
the chains a direct launch gives, bit for bit: fields, random state, sign and G after every
replay, with the program's own eager work (state readbacks, a measured sweep outside the graph)
and synchronizations between the replays.  The captured graph holds the sweep kernel only."""
import numpy as np
import pytest

import dqmc_amd as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return D.Context(0)


def batch(ctx, graph):
    ps = [D.make_params(6, 24, 2.0, pace=6, checkerboard=True, delay=16, U=4.0, mu=-0.2, seed=s) for s in range(8)]
    ps += [D.make_params(4, 20, 2.0, pace=5, model=D.ATTRACTIVE, U=3.0, mu=0.1, seed=30 + s) for s in range(4)]
    b = D.Batch(ctx, ps)
    b.use_graph(graph)
    b.fields_random()
    b.init()
    return b


def test_graph_replay_equals_direct_launch(ctx):
    print("[gpu-graph] replay vs direct", flush=True)
    g, d = batch(ctx, True), batch(ctx, False)
    for it in range(6):
        g.sweep(1)
        d.sweep(1)
        g.synchronize(); d.synchronize()
        for c in range(g.n):
            ig, idd = g.info(c), d.info(c)
            assert ig.rng == idd.rng and ig.sign == idd.sign and ig.pos == idd.pos, (it, c)
            assert np.array_equal(g.fields(c), d.fields(c)), (it, c)
            for spin in (1, -1):
                assert np.array_equal(g.green(c, spin), d.green(c, spin)), (it, c, spin)
        if it == 2:
            # eager work on both between the replays: a measured two-sweep launch (a new capture
            # for the graph batch, since n_sweeps and flags change) and readbacks
            g.sweep(2, measure=True); d.sweep(2, measure=True)
            sg, cg = g.measure_accum(0)
            sd, cd = d.measure_accum(0)
            assert np.array_equal(sg, sd) and np.array_equal(cg, cd)
    assert g.last_sweep_ms() > 0
