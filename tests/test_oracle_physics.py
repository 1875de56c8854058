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


# ----------------------------------------------------------------------------
# The equal-time measurement formulas against many-body traces (no Wick's theorem)
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("model,cb,momentum", [(0, False, -1), (0, True, 0), (1, False, -1), (1, True, 1),
                                               (0, False, 1)])
def test_equal_time_observables_vs_fock_space(model, cb, momentum):
    """For one field configuration on the 2x2 lattice, every equal-time observable of
    Measure::Methods (src/measure/measure_methods.cpp:35-255) evaluated by the oracle on the
    exact G(tau) of each slice equals the expectation value of the operator it stands for,
    Tr[B(beta,tau) O B(tau,0)] / Tr B(beta,0) in the 16-state Fock space of each spin (the weight
    of a fixed field is a product over the spins, so cross-spin terms are products of per-spin
    expectations; same-spin two-body terms are traced as operator products):
      filling 1/N sum <n>, double occupancy 1/N sum <n_up n_dn>, kinetic -2t/N sum_{i, +x, +y}
      sum_s <c+_{i+d} c_i>, n(k) 1/(2N) sum_s sum_ij cos(k r_ij) <c+_i c_j>, local moment
      1/N sum <(n_up - n_dn)^2>, SDW and CDW 1/N^2 sum_ij cos(Q r_ij) <m_i m_j> with m = n_up -/+ n_dn,
      P_s 1/N sum_ij <Delta+_i Delta_j>, Delta_j = c_{j,dn} c_{j,up}.
    A flipped sign of a cross-spin term in SDW or CDW fails this check.  The direction
    convention of the kinetic bond (G(i, i+d) rather than G(i+d, i)) is not something any
    lattice's physics fixes (the two agree on average, and on 2x2 sum to the same value), so it
    stays a reading of src/measure/measure_methods.cpp:89-96."""
    from test_oracle_dynamic import Fock, slice_matrices
    side, L, beta = 2, 6, 1.2
    c = O.Chain(side, L, beta, pace=3, model=model, checkerboard=cb, U=3.0, mu=-0.35, momentum=momentum)
    c.seed(19); c.fields_random(); c.init()
    n = c.N
    F = Fock(n)
    lat = O.lattice(side)
    nn, disp, four = lat["nn"], lat["disp"], lat["fourier"]
    k = lat["points"][2] if momentum < 0 else momentum
    lifted = {sp: [F.lift(B) for B in slice_matrices(c, sp)] for sp in (1, -1)}

    def prod(sp, t2, t1):
        M = np.eye(F.dim)
        for t in range(t1 + 1, t2 + 1):
            M = lifted[sp][t - 1] @ M
        return M

    Z = {sp: np.trace(prod(sp, L, 0)) for sp in (1, -1)}
    num = [F.cd[i] @ F.c[i] for i in range(n)]
    exact = np.zeros(9)
    c.clear_measure()
    for tau in range(1, L + 1):
        ev = {sp: (lambda A, sp=sp: np.trace(prod(sp, L, tau) @ A @ prod(sp, tau, 0)) / Z[sp]) for sp in (1, -1)}
        hop = {sp: np.array([[ev[sp](F.cd[i] @ F.c[j]) for j in range(n)] for i in range(n)]) for sp in (1, -1)}  # <c+_i c_j>
        nn2 = {sp: np.array([[ev[sp](num[i] @ num[j]) for j in range(n)] for i in range(n)]) for sp in (1, -1)}
        occ = {sp: np.diag(hop[sp]) for sp in (1, -1)}
        G = {sp: np.eye(n) - hop[sp].T for sp in (1, -1)}   # G(i,j) = <c_i c+_j>
        c.set_slice(tau - 1, G[1], G[-1], 1.0)
        f = np.array([[four[disp[i, j], k] for j in range(n)] for i in range(n)])
        exact[1] += (occ[1] + occ[-1]).sum() / n
        exact[2] += (occ[1] * occ[-1]).sum() / n
        exact[3] += -2.0 * sum(hop[sp][nn[i, d], i] for sp in (1, -1) for i in range(n) for d in (0, 1)) / n
        exact[4] += sum((f * hop[sp]).sum() for sp in (1, -1)) / (2 * n)
        exact[5] += sum(nn2[1][i, i] + nn2[-1][i, i] - 2 * occ[1][i] * occ[-1][i] for i in range(n)) / n
        cross = np.outer(occ[1], occ[-1]) + np.outer(occ[-1], occ[1])
        exact[6] += (f * (nn2[1] + nn2[-1] - cross)).sum() / n ** 2
        exact[7] += (f * (nn2[1] + nn2[-1] + cross)).sum() / n ** 2
        exact[8] += (hop[1] * hop[-1]).sum() / n
        exact[0] += 1.0
    c.measure()
    s, cnt = c.measure_accum()
    assert np.all(cnt == L)
    assert np.max(np.abs(s - exact)) < 1e-11, (s, exact)
    assert abs(exact[6]) > 1e-3 and abs(exact[8]) > 1e-3   # the structure factors are not trivially zero
