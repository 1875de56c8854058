"""GPU parity of the DQMC sweep engine (HIP, through the C-ABI) against the
oracle on identical fields, random streams and exp(-dtau K).

Bar: fields and accept/reject bits bit-identical; slice signs identical;
max|dG| <= 1e-10 on every slice whose G was freshly recomputed from the UDV
stacks (stabilization slices); on wrapped slices the two FP64 implementations
can differ by the algorithm's own wrap error (max|G_fresh - G_wrapped| of the
oracle, the reference's WrapError(), 1e-12..1e-7 depending on dtau and U),
so there the bar is max(1e-9, 10 x oracle wrap error); observables to 1e-10
relative.
"""
import numpy as np
import pytest

import _oracle as O
import dqmc_amd as D

pytestmark = pytest.mark.gpu

TOL_G = 1e-10


@pytest.fixture(scope="module")
def ctx():
    return D.Context(0)


def pair(ctx, side, L, beta, pace, model=0, cb=False, U=4.0, mu=0.0, seed=42, delay=16, share_expk=True):
    orc = O.Chain(side, L, beta, pace=pace, model=model, checkerboard=cb, U=U, mu=mu)
    orc.seed(seed)
    orc.fields_random()
    p = D.make_params(side, L, beta, pace=pace, model=model, checkerboard=cb, delay=delay, U=U, mu=mu, seed=seed)
    b = D.Batch(ctx, [p], record=True)
    b.fields_random()
    assert np.array_equal(b.fields(0).astype(np.float64), orc.fields())
    assert b.info(0).rng == orc.rng_state
    if share_expk and not cb:
        b.set_expk(0, orc.expK(0), orc.expK(1))
    orc.init()
    b.init()
    return orc, b


def fresh_slots(L, pace, up):
    """Slots whose stored G comes straight from the stacks (no wraps since)."""
    if up:
        return {t - 1 for t in range(1, L + 1) if t % pace == 0 or t == L}
    return {t - 1 for t in range(1, L) if t % pace == 0} | {L - 1}


def compare_sweep(orc, b, label, pace, up):
    n, L = orc.N, orc.L
    assert np.array_equal(b.fields(0).astype(np.float64), orc.fields()), label
    assert np.array_equal(b.accepts(0), orc.accepts()), label
    sg = b.slice_signs(0)
    fresh = fresh_slots(L, pace, up)
    worst_fresh = worst_wrapped = 0.0
    for slot in range(L):
        assert sg[slot] == orc.slice_sign(slot), (label, slot)
        for spin in (1, -1):
            e = np.max(np.abs(b.slice_green(0, slot, spin) - orc.slice_green(slot, spin)))
            if slot in fresh:
                worst_fresh = max(worst_fresh, e)
            else:
                worst_wrapped = max(worst_wrapped, e)
    tol_wrapped = max(1e-9, 10.0 * orc.wrap_error)
    print(f"[gpu-parity] {label}: max|dG| fresh {worst_fresh:.2e}, wrapped {worst_wrapped:.2e} "
          f"(oracle wrap error {orc.wrap_error:.2e})", flush=True)
    assert worst_fresh < TOL_G, (label, worst_fresh)
    assert worst_wrapped < tol_wrapped, (label, worst_wrapped, tol_wrapped)
    return max(worst_fresh, worst_wrapped)


CASES = [
    # side, L, beta, pace, model, cb, U, mu
    (4, 80, 4.0, 10, 0, False, 4.0, 0.0),     # configuration 1 (4x4, beta 4, 80 slices, dense)
    (4, 40, 2.0, 10, 0, True, 4.0, 0.0),
    (6, 30, 3.0, 7, 0, True, 4.0, -0.3),      # L % pace != 0
    (4, 40, 4.0, 10, 1, False, 4.0, -0.5),    # attractive, dense
    (4, 40, 4.0, 10, 1, True, 4.0, -0.5),     # attractive, checkerboard (mu-sign convention of the break-up)
    (2, 16, 1.0, 4, 0, False, 6.0, 0.2),      # side 2 (double-counted bonds)
]


@pytest.mark.parametrize("case", CASES, ids=[f"s{c[0]}L{c[1]}m{c[4]}cb{int(c[5])}" for c in CASES])
def test_sweeps_match_oracle(ctx, case):
    side, L, beta, pace, model, cb, U, mu = case
    print(f"[gpu-parity] case {case}", flush=True)
    orc, b = pair(ctx, side, L, beta, pace, model, cb, U, mu)
    for spin in (1, -1):   # the initial G is recomputed from the UDV stack: the fresh-G bar
        assert np.max(np.abs(b.green(0, spin) - orc.green(spin))) < TOL_G
    assert b.info(0).sign == orc.sign
    for it in range(3):
        orc.sweep_forward(); b.sweep(1, record=True)
        compare_sweep(orc, b, f"up{it}", pace, True)
        orc.sweep_backward(); b.sweep(1, record=True)
        compare_sweep(orc, b, f"down{it}", pace, False)
    info = b.info(0)
    assert info.sign == orc.sign
    assert info.rng == orc.rng_state
    # the wrap error is a max of amplified rounding noise: same order, not the same value
    assert info.wrap_error < max(1e-9, 10.0 * orc.wrap_error)


@pytest.mark.parametrize("delay", [1, 3, 16, 32])
def test_delay_depth_does_not_change_the_chain(ctx, delay):
    print(f"[gpu-parity] delay {delay}", flush=True)
    orc, b = pair(ctx, 6, 20, 2.0, 5, 0, True, 5.0, 0.0, seed=7, delay=delay)
    for it in range(2):
        orc.sweep_forward(); b.sweep(1, record=True); compare_sweep(orc, b, f"k{delay}up{it}", 5, True)
        orc.sweep_backward(); b.sweep(1, record=True); compare_sweep(orc, b, f"k{delay}down{it}", 5, False)


def test_measurements_match_oracle(ctx):
    print("[gpu-parity] measurements", flush=True)
    orc, b = pair(ctx, 4, 40, 3.0, 10, 0, False, 4.0, -0.2, seed=5)
    for _ in range(4):
        orc.sweep_forward(); orc.measure()
        orc.sweep_backward(); orc.measure()
    b.sweep(8, measure=True)
    s_o, c_o = orc.measure_accum()
    s_g, c_g = b.measure_accum(0)
    assert np.array_equal(c_o, c_g)
    assert np.allclose(s_g, s_o, rtol=1e-10, atol=1e-10), (s_g, s_o)


def test_host_expk_matches_oracle(ctx):
    for side, beta, L, mu in [(4, 4.0, 80, 0.0), (6, 3.0, 30, -0.4), (2, 1.0, 10, 0.3), (12, 6.0, 60, 0.0)]:
        orc = O.Chain(side, L, beta, U=4.0, mu=mu)
        b = D.Batch(ctx, [D.make_params(side, L, beta, U=4.0, mu=mu)])
        e, ie = b.expk(0)
        assert np.max(np.abs(e - orc.expK(0))) < 1e-14
        assert np.max(np.abs(ie - orc.expK(1))) < 1e-14


def test_uneven_batch_equals_chains_alone(ctx):
    print("[gpu-parity] uneven batch", flush=True)
    ps = [D.make_params(4, 40, 4.0, pace=10, seed=1),
          D.make_params(6, 24, 2.0, pace=6, checkerboard=True, seed=2, U=3.0),
          D.make_params(4, 30, 3.0, pace=8, model=D.ATTRACTIVE, mu=-0.3, seed=3),
          D.make_params(8, 20, 2.0, pace=5, checkerboard=True, seed=4, mu=0.1)]
    mix = D.Batch(ctx, ps, record=True)
    mix.fields_random(); mix.init(); mix.sweep(3, measure=True)
    for c, p in enumerate(ps):
        alone = D.Batch(ctx, [p], record=True)
        alone.fields_random(); alone.init(); alone.sweep(3, measure=True)
        assert np.array_equal(alone.fields(0), mix.fields(c))
        for spin in (1, -1):
            assert np.array_equal(alone.green(0, spin), mix.green(c, spin))
        assert np.array_equal(alone.measure_accum(0)[0], mix.measure_accum(c)[0])


def test_example_fields_stress_green(ctx):
    # the reference's beta = 8 field snapshot (example/output/fields.out), attractive 8x8
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "example_fields_L160_N64.npy"))
    orc = O.Chain(8, 160, 8.0, pace=10, model=1, U=4.0, mu=0.0)
    orc.set_fields(f.astype(np.float64))
    orc.init()
    b = D.Batch(ctx, [D.make_params(8, 160, 8.0, pace=10, model=D.ATTRACTIVE, U=4.0)])
    b.set_fields(0, f)
    b.set_expk(0, orc.expK(0), orc.expK(1))
    b.init()
    assert np.max(np.abs(b.green(0, 1) - orc.green(1))) < TOL_G
