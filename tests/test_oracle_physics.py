"""Physics pins of the oracle's full sweep + equal-time measurement:
U = 0 closed forms, exact Hubbard-Stratonovich enumeration on a 2x2 lattice,
and sweep bookkeeping invariants (src/dqmc_walker.cpp, src/measure/measure_methods.cpp)."""
import itertools

import numpy as np
import pytest
import scipy.linalg

import _oracle as O


def free_green(side, beta, t=1.0, mu=0.0):
    lat = O.lattice(side)
    n = side * side
    K = -t * lat["hopping"] + mu * np.eye(n)
    return np.linalg.inv(np.eye(n) + scipy.linalg.expm(-beta * K))


@pytest.mark.parametrize("side,mu", [(4, 0.0), (4, -0.4), (6, 0.3)])
def test_free_fermions_every_slice(side, mu):
    beta, L = 3.0, 30
    c = O.Chain(side, L, beta, pace=7, model=0, U=0.0, mu=mu)
    c.seed(5); c.fields_random(); c.init()
    G0 = free_green(side, beta, mu=mu)
    c.sweep_forward()
    for slot in range(L):
        assert np.max(np.abs(c.slice_green(slot, 1) - G0)) < 1e-11
        assert np.max(np.abs(c.slice_green(slot, -1) - G0)) < 1e-11
    c.measure()
    s, cnt = c.measure_accum()
    vals = s / cnt
    d = np.diag(G0)
    assert abs(vals[2] - np.mean((1 - d) ** 2)) < 1e-11                      # double occupancy
    lat = O.lattice(side)
    kin = 2 * sum(G0[i, lat["nn"][i, dd]] for i in range(side * side) for dd in (0, 1)) * 2 / (side * side)
    assert abs(vals[3] - kin) < 1e-11                                          # kinetic energy
    assert abs(vals[1] - (2 - 2 * np.trace(G0) / (side * side))) < 1e-11       # filling


def _exact_enumeration(side, L, beta, U, mu, model):
    lat = O.lattice(side)
    n = side * side
    dtau = beta / L
    K = -lat["hopping"] + mu * np.eye(n)
    expK = scipy.linalg.expm(-dtau * K)
    alpha = np.arccosh(np.exp(0.5 * dtau * U))
    cfgs = np.array(list(itertools.product([-1.0, 1.0], repeat=L * n))).reshape(-1, L, n)
    out = {}
    Ms = {}
    for spin in (+1, -1):
        sv = spin if model == 0 else 1.0
        M = np.broadcast_to(np.eye(n), (cfgs.shape[0], n, n)).copy()
        for t in range(L):
            B = np.exp(sv * alpha * cfgs[:, t, :])[:, :, None] * expK[None]
            M = B @ M
        Ms[spin] = M
    Gu = np.linalg.inv(np.eye(n)[None] + Ms[1])
    Gd = np.linalg.inv(np.eye(n)[None] + Ms[-1])
    w = np.linalg.det(np.eye(n)[None] + Ms[1]) * np.linalg.det(np.eye(n)[None] + Ms[-1])
    if model == 1:
        # attractive HS coupling carries exp(-alpha sum s): the e^{2 alpha s} factor of its ratio
        w = w * np.exp(-alpha * cfgs.sum(axis=(1, 2)))
    du = np.diagonal(Gu, axis1=1, axis2=2); dd = np.diagonal(Gd, axis1=1, axis2=2)
    docc = ((1 - du) * (1 - dd)).mean(axis=1)
    kin = np.zeros(len(w))
    for i in range(n):
        for dr in (0, 1):
            j = lat["nn"][i, dr]
            kin += Gu[:, i, j] + Gd[:, i, j]
    kin *= 2.0 / n
    loc = (du + dd - 2 * du * dd).mean(axis=1)
    Z = w.sum()
    out["double_occupancy"] = (w * docc).sum() / Z
    out["kinetic_energy"] = (w * kin).sum() / Z
    out["local_spin_corr"] = (w * loc).sum() / Z
    out["filling"] = (w * (2 - (np.trace(Gu, axis1=1, axis2=2) + np.trace(Gd, axis1=1, axis2=2)) / n)).sum() / Z
    return out


@pytest.mark.parametrize("model,U,mu", [(0, 4.0, 0.0), (1, 4.0, -0.3)])
def test_exact_enumeration_2x2(model, U, mu):
    side, L, beta = 2, 4, 1.0
    exact = _exact_enumeration(side, L, beta, U, mu, model)
    c = O.Chain(side, L, beta, pace=2, model=model, U=U, mu=mu)
    c.seed(2024); c.fields_random(); c.init()
    for _ in range(100):
        c.sweep_forward(); c.sweep_backward()
    bins = []
    for b in range(40):
        c.clear_measure()
        for _ in range(100):
            c.sweep_forward(); c.measure(); c.sweep_backward(); c.measure()
        s, cnt = c.measure_accum()
        bins.append(O.normalize_bin(s, cnt))
    mean, err = O.bin_stats(bins)
    for name, idx in [("filling", 1), ("double_occupancy", 2), ("kinetic_energy", 3), ("local_spin_corr", 5)]:
        tol = 4.5 * err[idx] + 1e-12
        assert abs(mean[idx] - exact[name]) < tol, (name, mean[idx], exact[name], err[idx])


def test_sweep_bookkeeping_and_determinism():
    a = O.Chain(4, 20, 2.0, pace=6, U=4.0)   # L % pace != 0: partial last chunk
    b = O.Chain(4, 20, 2.0, pace=6, U=4.0)
    for c in (a, b):
        c.seed(3); c.fields_random(); c.init()
    for _ in range(3):
        for c in (a, b):
            c.sweep_forward(); c.sweep_backward()
        assert np.array_equal(a.fields(), b.fields())
        assert np.array_equal(a.accepts(), b.accepts())
    assert a.wrap_error < 1e-8
    # after a backward sweep G is the fresh G(0), stored in slot L-1
    assert np.array_equal(a.green(1), a.slice_green(19, 1))


def test_seed_draw_order_fields():
    # fields are drawn t-outer, i-inner, one bernoulli(0.5) each
    c = O.Chain(4, 10, 1.0)
    c.seed(42); c.fields_random()
    bits = O.bernoulli_stream(42, 160, 0.5)
    assert np.array_equal((c.fields() > 0).astype(np.uint8).ravel(), bits)
