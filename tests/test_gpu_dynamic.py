"""GPU parity of the time-displaced path (DqmcWalker::sweep_for_dynamic_greens,
src/dqmc_walker.cpp:424-524, and MeasureHandler::dynamic_measure,
src/measure/measure_methods.cpp:264-461) against the oracle, through the C-ABI.

Bar: G(t,t), G(t,0), G(0,t) within 1e-10 on slots recomputed from the stacks and within
max(1e-9, 10 x oracle wrap error) on wrapped slots (the two FP64 implementations differ by
rounding that the wraps amplify, as in the equal-time sweep); fields, accept bits and slice
signs of the updating sweeps around it identical; the dynamic sums (G(k,t) over every k-star,
D(t), DSS(t), rho_s) within N x the wrapped-slot bar, sign and count exact.
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


def pair(ctx, side, L, beta, pace, model=0, cb=False, U=4.0, mu=0.0, seed=42):
    orc = O.Chain(side, L, beta, pace=pace, model=model, checkerboard=cb, U=U, mu=mu)
    orc.seed(seed)
    orc.fields_random()
    p = D.make_params(side, L, beta, pace=pace, model=model, checkerboard=cb, U=U, mu=mu, seed=seed)
    b = D.Batch(ctx, [p], record=True)
    b.fields_random()
    if not cb:
        b.set_expk(0, orc.expK(0), orc.expK(1))
    orc.init()
    b.init()
    return orc, b


def fresh_slots(L, pace):
    return {t - 1 for t in range(1, L + 1) if t % pace == 0 or t == L}


def compare_dynamic_slots(orc, b, label, pace):
    L = orc.L
    fresh = fresh_slots(L, pace)
    tol_wrapped = max(1e-9, 10.0 * orc.wrap_error)
    worst = {True: 0.0, False: 0.0}
    for slot in range(L):
        for spin in (1, -1):
            mats = [(b.slice_green(0, slot, spin), orc.slice_green(slot, spin)),
                    (b.dynamic_green(0, slot, spin, 0), orc.dynamic_green(slot, spin, 0)),
                    (b.dynamic_green(0, slot, spin, 1), orc.dynamic_green(slot, spin, 1))]
            for g, o in mats:
                e = np.max(np.abs(g - o))
                worst[slot in fresh] = max(worst[slot in fresh], e)
    print(f"[gpu-dynamic] {label}: max|dG| fresh {worst[True]:.2e}, wrapped {worst[False]:.2e} "
          f"(oracle wrap error {orc.wrap_error:.2e})", flush=True)
    assert worst[True] < TOL_G, (label, worst[True])
    assert worst[False] < tol_wrapped, (label, worst[False], tol_wrapped)
    return tol_wrapped


def compare_eq_sweep(orc, b, label):
    assert np.array_equal(b.fields(0).astype(np.float64), orc.fields()), label
    assert np.array_equal(b.accepts(0), orc.accepts()), label
    sg = b.slice_signs(0)
    for slot in range(orc.L):
        assert sg[slot] == orc.slice_sign(slot), (label, slot)


def compare_dynamic_sums(orc, b, tol, label):
    a = orc.dynamic_accum()
    g = b.dynamic_accum(0)
    assert g["count"] == a["count"] and g["sign"] == a["sign"], label
    n = orc.N
    for key in ("gk", "dos", "dss"):
        err = np.max(np.abs(g[key] - a[key]))
        assert err < n * tol, (label, key, err)
    assert abs(g["rho"] - a["rho"]) < n * tol, (label, g["rho"], a["rho"])


CASES = [
    # side, L, beta, pace, model, cb, U, mu
    (4, 40, 3.0, 10, 0, False, 4.0, -0.2),    # dense repulsive
    (4, 40, 2.0, 10, 0, True, 4.0, 0.0),      # checkerboard
    (6, 30, 3.0, 7, 1, False, 4.0, -0.3),     # attractive, L % pace != 0
    (4, 24, 2.0, 5, 1, True, 4.0, -0.5),      # attractive checkerboard
    (2, 16, 1.0, 4, 0, False, 6.0, 0.2),      # side 2
    (4, 10, 1.0, 10, 0, False, 4.0, 0.1),     # one stack level: the recompute at t = beta only
    (3, 1, 0.2, 1, 0, False, 2.0, 0.1),       # a single slice
    (5, 12, 1.2, 5, 0, False, 4.0, 0.0),      # odd side (site maps, the stiffness k-star of an odd lattice)
]


@pytest.mark.parametrize("case", CASES, ids=[f"s{c[0]}L{c[1]}p{c[3]}m{c[4]}cb{int(c[5])}" for c in CASES])
def test_dynamic_sweeps_match_oracle(ctx, case):
    side, L, beta, pace, model, cb, U, mu = case
    print(f"[gpu-dynamic] case {case}", flush=True)
    orc, b = pair(ctx, side, L, beta, pace, model, cb, U, mu)
    # one updating forth-and-back, then the reference's measuring schedule twice:
    # dynamic sweep + dynamic measurement, updating backward sweep + equal-time measurement
    orc.sweep_forward(); b.sweep(1, record=True)
    orc.sweep_backward(); b.sweep(1, record=True)
    for it in range(2):
        orc.sweep_dynamic(); orc.dynamic_measure()
        b.sweep(1, measure=True, record=True, dynamic=True)
        assert np.array_equal(b.fields(0).astype(np.float64), orc.fields())
        tol = compare_dynamic_slots(orc, b, f"dyn{it}", pace)
        compare_dynamic_sums(orc, b, tol, f"dyn{it}")
        orc.sweep_backward(); orc.measure()
        b.sweep(1, measure=True, record=True, dynamic=True)
        compare_eq_sweep(orc, b, f"down{it}")
    s_o, c_o = orc.measure_accum()
    s_g, c_g = b.measure_accum(0)
    assert np.array_equal(c_o, c_g)          # equal-time counts: backward sweeps only
    assert np.allclose(s_g, s_o, rtol=1e-9, atol=1e-9), (s_g, s_o)
    info = b.info(0)
    assert info.sign == orc.sign and info.rng == orc.rng_state
    assert info.wrap_error < max(1e-9, 10.0 * orc.wrap_error)


def test_batched_dynamic_sweeps(ctx):
    # several forth-and-back iterations in one launch equal the same work launched per direction
    ps = [D.make_params(4, 20, 2.0, pace=5, seed=3, U=4.0),
          D.make_params(6, 24, 2.0, pace=6, checkerboard=True, seed=4, U=3.0),
          D.make_params(4, 16, 2.0, pace=4, model=D.ATTRACTIVE, seed=5, mu=-0.2)]
    one = D.Batch(ctx, ps)
    one.fields_random(); one.init(); one.sweep(6, measure=True, dynamic=True)
    split = D.Batch(ctx, ps)
    split.fields_random(); split.init()
    for _ in range(6):
        split.sweep(1, measure=True, dynamic=True)
    for c in range(len(ps)):
        assert np.array_equal(one.fields(c), split.fields(c))
        a, g = one.dynamic_accum(c), split.dynamic_accum(c)
        assert a["count"] == 3 and g["count"] == 3
        for key in ("gk", "dos", "dss"):
            assert np.array_equal(a[key], g[key])
        assert a["rho"] == g["rho"]
        alone = D.Batch(ctx, [ps[c]])
        alone.fields_random(); alone.init(); alone.sweep(6, measure=True, dynamic=True)
        assert np.array_equal(alone.dynamic_accum(0)["gk"], a["gk"])


@pytest.mark.parametrize("side,L,beta", [(12, 60, 6.0), (21, 16, 2.0)])
def test_free_fermions_closed_form_on_the_gpu(ctx, side, L, beta):
    # U = 0 at the production size (12 x 12, beta 6, 60 slices) and the largest lattice (21 x 21,
    # N = 441): G(t,0) and G(0,t) have closed forms in the plane-wave basis at every slice
    mu = -0.3
    n = side * side
    b = D.Batch(ctx, [D.make_params(side, L, beta, pace=10 if side == 12 else 4, U=0.0, mu=mu, seed=1)], record=True)
    b.fields_random(); b.init()
    b.sweep(1, record=True, measure=True, dynamic=True)
    lat = O.lattice(side)
    K = -lat["hopping"] + mu * np.eye(n)
    eps, V = np.linalg.eigh(K)
    dtau = beta / L
    worst = 0.0
    for tau in range(1, L + 1):
        gt0 = (V * (np.exp(-tau * dtau * eps) / (1 + np.exp(-beta * eps)))) @ V.T
        g0t = -(V * (np.exp(-(beta - tau * dtau) * eps) / (1 + np.exp(-beta * eps)))) @ V.T
        worst = max(worst, np.max(np.abs(b.dynamic_green(0, tau - 1, 1, 0) - gt0)),
                    np.max(np.abs(b.dynamic_green(0, tau - 1, -1, 1) - g0t)))
    print(f"[gpu-dynamic] free fermions {side}x{side}: max error {worst:.2e}", flush=True)
    assert worst < 1e-10
    if side % 2:
        return   # the k-star construction of src/lattice/square.cpp:42-66 is off the lattice's momenta for odd sides
    acc = b.dynamic_accum(0)
    ks = b.kstars(0)
    ek = -2 * (np.cos(ks[:, 0]) + np.cos(ks[:, 1])) + mu
    for t in range(1, L):
        assert np.max(np.abs(acc["gk"][:, t] - np.exp(-t * dtau * ek) / (1 + np.exp(-beta * ek)))) < 1e-10


def test_dynamic_accessor_errors(ctx):
    b = D.Batch(ctx, [D.make_params(4, 20, 2.0, pace=5, seed=3)])
    b.fields_random(); b.init(); b.sweep(2, measure=True, dynamic=True)
    with pytest.raises(D.DqmcError):
        b.dynamic_green(0, 0, 1, 0)                 # batch created without record buffers
    with pytest.raises(D.DqmcError):
        b.dynamic_accum(0, k_index=[0, 6])          # 4x4 has 6 k-stars: index 6 is out of range
    a = b.dynamic_accum(0, k_index=[5, 0])          # any subset, any order
    full = b.dynamic_accum(0)
    assert np.array_equal(a["gk"][0], full["gk"][5]) and np.array_equal(a["gk"][1], full["gk"][0])
    r = D.Batch(ctx, [D.make_params(4, 20, 2.0, pace=5, seed=3)], record=True)
    r.fields_random(); r.init(); r.sweep(1, record=True, dynamic=True)
    for bad in [(20, 1, 0), (0, 2, 0), (0, 1, 2)]:
        with pytest.raises(D.DqmcError):
            r.dynamic_green(0, *bad)
