"""Spin-pair launches (DQMC_LAUNCH_PAIR): each repulsive chain on two co-resident workgroups,
one per spin, exchanging the k x k block of the walk per delayed-update block; and quad
launches (DQMC_LAUNCH_QUAD) that add a helper workgroup per spin computing half of the output
columns of every dense wrap and slice product.

A pair launch computes the same numbers as the one-workgroup launch: every spin's next
G0[S,S] is formed from the block's factors in both, each 16x16 block of an update is one
wavefront's product whichever wavefronts run it, the walk gets identical inputs in both
workgroups, and the helpers split the GEMMs at 64-column tile boundaries (each element summed in
the same order).  So the tests demand bit-for-bit equality with DQMC_LAUNCH_SINGLE (fields, accept
bits, random state, sign, G of every recorded slot, measurement sums, wrap error), and the
one-workgroup launch is itself pinned to the oracle by the parity tests.  Shapes: dense and
checkerboard, the headline 12x12 chains, the 16x16 precision-workload chains, an uneven
batch of four sizes, a chain count that is not a multiple of 8 (idle padding workgroups),
and a failing chain beside healthy ones.  The reference runs the two spins one after the
other in one process (src/dqmc_walker.cpp:212-216).
"""
import numpy as np
import pytest

import _oracle as O
import dqmc_amd as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return D.Context(0)


def run(ctx, ps, mode, sweeps, measure_sweeps=2, record=True, expk=None):
    b = D.Batch(ctx, ps, record=record)
    b.set_launch_mode(mode)
    b.fields_random()
    if expk is not None:
        for c, (e, ie) in expk.items():
            b.set_expk(c, e, ie)
    b.init()
    if sweeps:
        b.sweep(sweeps)
    if measure_sweeps:
        b.sweep(measure_sweeps, measure=True, record=record)
    b.synchronize()
    return b


def assert_same(a, b, n, record=True):
    for c in range(n):
        ia, ib = a.info(c), b.info(c)
        assert ia.rng == ib.rng and ia.sign == ib.sign and ia.pos == ib.pos, c
        assert ia.status == ib.status and ia.wrap_error == ib.wrap_error, (c, ia.wrap_error, ib.wrap_error)
        assert np.array_equal(a.fields(c), b.fields(c)), c
        for spin in (1, -1):
            assert np.array_equal(a.green(c, spin), b.green(c, spin), equal_nan=True), (c, spin)
        sa, ca = a.measure_accum(c)
        sb, cb = b.measure_accum(c)
        assert np.array_equal(sa, sb) and np.array_equal(ca, cb), c
        if record:
            assert np.array_equal(a.accepts(c), b.accepts(c)), c
            assert np.array_equal(a.slice_signs(c), b.slice_signs(c)), c
            for slot in (0, 1, a.L[c] - 1):
                for spin in (1, -1):
                    assert np.array_equal(a.slice_green(c, slot, spin), b.slice_green(c, slot, spin), equal_nan=True), (c, slot, spin)


CASES = {
    "6x6-dense": [D.make_params(6, 40, 4.0, pace=8, U=4.0, mu=-0.3, delay=8, seed=s) for s in range(8)],
    "12x12-checkerboard": [D.make_params(12, 60, 6.0, pace=10, checkerboard=True, delay=32, U=4.0, mu=0.0, seed=s)
                           for s in range(8)],
    "uneven-13": [D.make_params(side, 24 + 8 * (c % 3), 3.0, pace=5 + c % 3, U=5.0, mu=-0.2 * (c % 4), delay=(8, 16, 32)[c % 3],
                                checkerboard=(side % 2 == 0 and c % 2 == 1), seed=100 + c)
                  for c, side in enumerate([4, 6, 8, 10, 4, 6, 8, 10, 5, 7, 9, 3, 6])],
}


DENSE = {
    "6x6-dense": CASES["6x6-dense"],   # N = 36: the helpers get no columns
    "12x12-dense": [D.make_params(12, 40, 4.0, pace=8, U=4.0, mu=-0.1 * s, delay=32, seed=s) for s in range(5)],
    "uneven-dense": [D.make_params(side, 24 + 8 * (c % 3), 3.0, pace=6, U=5.0, mu=-0.2 * (c % 4), delay=(8, 16, 32)[c % 3],
                                   seed=200 + c) for c, side in enumerate([4, 8, 10, 11, 12, 16, 9, 14, 13])],
}


@pytest.mark.skip(reason="quad launches (a helper workgroup per spin for half of every dense product) did not "
                         "finish on the device in two runs (the second after making the helpers' flag polls "
                         "relaxed); the cause is not found, so dqmc_batch_set_launch_mode refuses DQMC_LAUNCH_QUAD")
@pytest.mark.parametrize("name", sorted(DENSE))
def test_quad_equals_single_bit_for_bit(ctx, name):
    ps = DENSE[name]
    print(f"[gpu-pair] {name}: {len(ps)} chains, quad vs single", flush=True)
    single = run(ctx, ps, D.LAUNCH_SINGLE, 3)
    quad = run(ctx, ps, D.LAUNCH_QUAD, 3)
    assert quad.last_launch_mode() == D.LAUNCH_QUAD
    assert_same(single, quad, len(ps))
    single.close(); quad.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_pair_equals_single_bit_for_bit(ctx, name):
    ps = CASES[name]
    print(f"[gpu-pair] {name}: {len(ps)} chains, pair vs single", flush=True)
    single = run(ctx, ps, D.LAUNCH_SINGLE, 3)
    pair = run(ctx, ps, D.LAUNCH_PAIR, 3)
    assert single.last_launch_mode() == D.LAUNCH_SINGLE and pair.last_launch_mode() == D.LAUNCH_PAIR
    assert_same(single, pair, len(ps))
    print(f"[gpu-pair] {name}: identical; wrap error max {max(pair.info(c).wrap_error for c in range(len(ps))):.2e}",
          flush=True)
    single.close(); pair.close()


def test_pair_16x16_precision_workload_chains(ctx):
    """Eight chains of the 16x16 workload (U 6, beta 10, L 200, pace 8, dense, delay 32, four mu)
    over one sweep each way: identical in both launch forms, and AUTO picks the pair form."""
    print("[gpu-pair] 16x16 precision-workload chains", flush=True)
    mus = [0.0, -0.2, -0.4, -0.6]
    ps = [D.make_params(16, 200, 10.0, pace=8, U=6.0, mu=mus[c % 4], delay=32, seed=c) for c in range(8)]
    single = run(ctx, ps, D.LAUNCH_SINGLE, 1, measure_sweeps=1)
    pair = run(ctx, ps, D.LAUNCH_PAIR, 1, measure_sweeps=1)
    auto = run(ctx, ps, D.LAUNCH_AUTO, 1, measure_sweeps=1)
    assert pair.last_launch_mode() == D.LAUNCH_PAIR and auto.last_launch_mode() == D.LAUNCH_PAIR
    assert_same(single, pair, len(ps))
    assert_same(single, auto, len(ps))
    single.close(); pair.close(); auto.close()


def test_pair_chain_matches_the_oracle(ctx):
    """One 6x6 chain in a pair launch against the oracle over a forward and a backward sweep."""
    print("[gpu-pair] 6x6 pair launch vs oracle", flush=True)
    p = D.make_params(6, 40, 4.0, pace=8, U=4.0, mu=-0.3, delay=8, seed=5)
    b = D.Batch(ctx, [p], record=True)
    b.set_launch_mode(D.LAUNCH_PAIR)
    b.fields_random()
    orc = O.Chain(6, 40, 4.0, pace=8, U=4.0, mu=-0.3)
    orc.seed(5); orc.fields_random()
    orc.set_expK(b.expk(0)[0], b.expk(0)[1])
    b.init(); orc.init()
    for up in (True, False):
        b.sweep(1, record=True)
        orc.sweep_forward() if up else orc.sweep_backward()
        assert b.last_launch_mode() == D.LAUNCH_PAIR
        assert np.array_equal(b.fields(0).astype(np.float64), orc.fields())
        assert np.array_equal(b.accepts(0), orc.accepts())
        assert b.info(0).sign == orc.sign and b.info(0).rng == orc.rng_state
        for spin in (1, -1):
            assert np.max(np.abs(b.green(0, spin) - orc.green(spin))) < 1e-10
    b.close()


def test_pair_join_falls_back_on_a_busy_device(ctx):
    """A pair launch while another stream's 250-chain launch holds 250 of the 256 CUs: the chains
    whose partner workgroup cannot get a CU within 100 us run alone, the others pair once CUs free
    up, and every chain equals the one-workgroup launch bit for bit, with no synchronization
    failure (dqmc_chain_info.paired tells which way each went)."""
    print("[gpu-pair] pair launch beside a 250-chain launch on another stream", flush=True)
    busy = D.Batch(ctx, [D.make_params(12, 60, 6.0, pace=10, checkerboard=True, delay=32, U=4.0, seed=s)
                         for s in range(250)])
    busy.set_launch_mode(D.LAUNCH_SINGLE)
    busy.fields_random(); busy.init(); busy.synchronize()
    ps = [D.make_params(8, 40, 4.0, pace=8, U=4.0, mu=-0.2, delay=16, seed=s) for s in range(16)]
    ref = run(ctx, ps, D.LAUNCH_SINGLE, 2, measure_sweeps=0, record=False)
    b = D.Batch(ctx, ps)
    b.set_launch_mode(D.LAUNCH_PAIR)
    b.fields_random(); b.init(); b.synchronize()
    busy.sweep(6)   # asynchronous, on the busy batch's stream
    b.sweep(2)
    b.synchronize(); busy.synchronize()
    paired = [b.info(c).paired for c in range(16)]
    print(f"[gpu-pair] {sum(paired)} of 16 chains paired, {16 - sum(paired)} ran alone", flush=True)
    assert b.last_launch_mode() == D.LAUNCH_PAIR
    assert all(b.info(c).status == 0 for c in range(16))
    assert all(busy.info(c).status == 0 for c in range(250))
    assert_same(ref, b, 16, record=False)
    ref.close(); b.close(); busy.close()


def test_pair_launch_with_attractive_chains(ctx):
    """Attractive chains in a pair launch run alone (the second workgroup leaves); the batch equals
    the one-workgroup launch bit for bit and AUTO takes the pair form."""
    print("[gpu-pair] mixed repulsive / attractive batch", flush=True)
    ps = [D.make_params(6, 24, 3.0, pace=6, model=(D.ATTRACTIVE if c % 3 == 0 else D.REPULSIVE), U=4.0, mu=-0.3,
                        delay=8, seed=300 + c) for c in range(9)]
    single = run(ctx, ps, D.LAUNCH_SINGLE, 2)
    auto = run(ctx, ps, D.LAUNCH_AUTO, 2)
    assert auto.last_launch_mode() == D.LAUNCH_PAIR
    assert [auto.info(c).paired for c in range(9)] == [0 if c % 3 == 0 else 1 for c in range(9)]
    assert_same(single, auto, 9)
    single.close(); auto.close()


def test_failing_chain_in_a_pair_launch(ctx):
    """An overflowing exp(-dtau K) fails its chain at init (DQMC_STATUS_NONFINITE): neither of its
    workgroups sweeps or waits, and the healthy chains are the chains the one-workgroup launch gives."""
    print("[gpu-pair] failing chain beside healthy ones", flush=True)
    side = 4
    ps = [D.make_params(side, 20, 2.0, pace=5, U=4.0, seed=s) for s in (3, 11, 12)]
    huge = np.full((side * side, side * side), 1e200, order="F")
    single = run(ctx, ps, D.LAUNCH_SINGLE, 2, expk={0: (huge, huge)})
    pair = run(ctx, ps, D.LAUNCH_PAIR, 2, expk={0: (huge, huge)})
    assert pair.info(0).status & D.STATUS_FATAL and pair.info(0).pos == 0
    assert not pair.info(0).status & D.STATUS_PAIR_SYNC
    for c in (1, 2):
        assert pair.info(c).status == 0
    assert_same(single, pair, 3)
    single.close(); pair.close()


def test_launch_mode_rules(ctx):
    """PAIR needs a repulsive chain and the pair grid resident at once; the time-displaced sweep
    falls back to one workgroup per chain."""
    print("[gpu-pair] launch-mode rules", flush=True)
    att = D.Batch(ctx, [D.make_params(4, 16, 2.0, pace=4, model=D.ATTRACTIVE, U=4.0, seed=1)])
    with pytest.raises(D.DqmcError):
        att.set_launch_mode(D.LAUNCH_PAIR)
    att.fields_random(); att.init(); att.sweep(1); att.synchronize()
    assert att.last_launch_mode() == D.LAUNCH_SINGLE
    att.close()
    cb = D.Batch(ctx, [D.make_params(4, 8, 1.0, pace=4, checkerboard=True, U=4.0, seed=s) for s in range(4)])
    with pytest.raises(D.DqmcError):
        cb.set_launch_mode(D.LAUNCH_QUAD)   # refused (reserved)
    cb.set_launch_mode(D.LAUNCH_PAIR)
    cb.close()
    mid = D.Batch(ctx, [D.make_params(4, 8, 1.0, pace=4, U=4.0, seed=s) for s in range(72)])
    with pytest.raises(D.DqmcError):
        mid.set_launch_mode(D.LAUNCH_QUAD)   # refused (reserved)
    mid.set_launch_mode(D.LAUNCH_PAIR)
    mid.close()
    many = D.Batch(ctx, [D.make_params(4, 8, 1.0, pace=4, U=4.0, seed=s) for s in range(256)])
    with pytest.raises(D.DqmcError):
        many.set_launch_mode(D.LAUNCH_PAIR)
    many.fields_random(); many.init(); many.sweep(1); many.synchronize()
    assert many.last_launch_mode() == D.LAUNCH_SINGLE   # AUTO: 512 workgroups do not fit
    many.close()
    ps = [D.make_params(6, 24, 3.0, pace=6, U=4.0, mu=-0.2, delay=8, seed=s) for s in range(3)]
    out = {}
    for mode in (D.LAUNCH_SINGLE, D.LAUNCH_PAIR):
        b = D.Batch(ctx, ps)
        b.set_launch_mode(mode)
        b.fields_random(); b.init()
        b.sweep(2)
        b.sweep(2, measure=True, dynamic=True)
        b.synchronize()
        assert b.last_launch_mode() == D.LAUNCH_SINGLE   # the time-displaced sweeps ran unpaired
        out[mode] = b
    assert_same(out[D.LAUNCH_SINGLE], out[D.LAUNCH_PAIR], 3, record=False)
    for c in range(3):
        assert np.array_equal(out[D.LAUNCH_SINGLE].dynamic_accum(c)["gk"], out[D.LAUNCH_PAIR].dynamic_accum(c)["gk"])
    for b in out.values():
        b.close()


def test_autotune_launch(ctx):
    """dqmc_autotune_launch times both mappings where pairs apply (8 chains of 12x12: pairs are
    faster, the CUs being idle) and only the one-workgroup form where they do not (256 chains)."""
    print("[gpu-pair] launch autotuner", flush=True)
    p = D.make_params(12, 60, 6.0, pace=10, checkerboard=True, delay=32, U=4.0, seed=1)
    best, ms = ctx.autotune_launch(p, 8, sweeps=1)
    print(f"[gpu-pair] 8 x 12x12: {ms}", flush=True)
    assert set(ms) == {D.LAUNCH_SINGLE, D.LAUNCH_PAIR} and best == D.LAUNCH_PAIR
    assert ms[D.LAUNCH_PAIR] < ms[D.LAUNCH_SINGLE]
    q = D.make_params(4, 8, 1.0, pace=4, U=4.0, seed=1)
    best, ms = ctx.autotune_launch(q, 256, sweeps=1)
    assert set(ms) == {D.LAUNCH_SINGLE} and best == D.LAUNCH_SINGLE
