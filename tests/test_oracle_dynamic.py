"""Pins of the oracle's time-displaced path: DqmcWalker::sweep_for_dynamic_greens
(src/dqmc_walker.cpp:424-524), the dynamic recompute (include/utils/numerical_stable.hpp:159-179,
:246-325) and MeasureHandler::dynamic_measure (src/measure/measure_methods.cpp:264-461).

* U = 0 closed forms in the plane-wave basis: G(tau,0) = B^tau (1 + B^L)^-1,
  G(0,tau) = -(1 + B^L)^-1 B^(L-tau), G(k,tau), D(tau) and DSS(tau).
* Many-body brute force on a 2x2 lattice (4 sites, 16 Fock states per spin): for one field
  configuration the time-ordered correlators Tr[B(beta,tau) O1 B(tau,0) O2] / Tr[B(beta,0)]
  with the Fock-space lift of each slice matrix give G(tau,0), G(0,tau) and every dynamic
  observable (the current-current correlation of rho_s included) without Wick's theorem.
* The reference's SVD-stack formulas, the orthogonal-U form of the recompute and QR stacks
  give the same Green's functions to rounding.
"""
import itertools

import numpy as np
import pytest

import _oracle as O


def slice_matrices(c, spin):
    n = c.N
    return [c.mult(0, t, spin, np.eye(n)) for t in range(1, c.L + 1)]


def chain_product(Bs, t2, t1):
    """B(t2, t1) = B_t2 ... B_{t1+1}."""
    M = np.eye(Bs[0].shape[0])
    for t in range(t1 + 1, t2 + 1):
        M = Bs[t - 1] @ M
    return M


# ----------------------------------------------------------------------------
# U = 0: plane waves
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("general", [False, True])
def test_free_fermions_closed_form(general):
    side, L, beta, mu = 4, 60, 6.0, -0.3
    n = side * side
    c = O.Chain(side, L, beta, pace=7, model=0, U=0.0, mu=mu)
    c.set_dynamic_formula(general)
    c.seed(3); c.fields_random(); c.init()
    c.sweep_dynamic()
    lat = O.lattice(side)
    K = -lat["hopping"] + mu * np.eye(n)
    eps, V = np.linalg.eigh(K)
    dtau = beta / L
    for tau in range(1, L + 1):
        gt0 = (V * (np.exp(-tau * dtau * eps) / (1 + np.exp(-beta * eps)))) @ V.T
        g0t = -(V * (np.exp(-(beta - tau * dtau) * eps) / (1 + np.exp(-beta * eps)))) @ V.T
        for spin in (1, -1):
            assert np.max(np.abs(c.dynamic_green(tau - 1, spin, 0) - gt0)) < 1e-10, tau
            assert np.max(np.abs(c.dynamic_green(tau - 1, spin, 1) - g0t)) < 1e-10, tau
    # observables: G(k,tau) = e^{-tau eps_k} / (1 + e^{-beta eps_k}), D(tau) its k-average,
    # DSS(tau) = D(tau) D(beta - tau) / 2 for tau >= 1 and n(1-n)/2 per site at tau = 0 (slot L-1)
    c.dynamic_measure()
    acc = c.dynamic_accum()
    assert acc["count"] == 1.0 and acc["sign"] == 1.0
    ks = lat["kstars"]
    ek = -2 * (np.cos(ks[:, 0]) + np.cos(ks[:, 1])) + mu
    allk = 2 * np.pi * np.array(list(itertools.product(range(side), repeat=2))) / side
    e_all = -2 * (np.cos(allk[:, 0]) + np.cos(allk[:, 1])) + mu
    fd = lambda tt, e: np.exp(-tt * e) / (1 + np.exp(-beta * e))
    for t in range(L):
        tt = t * dtau
        if t == 0:
            gk_exp = 1 / (1 + np.exp(-beta * ek))       # G(0,0) in k space
        else:
            gk_exp = fd(tt, ek)
        assert np.max(np.abs(acc["gk"][:, t] - gk_exp)) < 1e-10, t
        d_t = np.mean(fd(tt, e_all)) if t > 0 else np.mean(1 / (1 + np.exp(-beta * e_all)))
        assert abs(acc["dos"][t] - d_t) < 1e-10, t
        if t > 0:
            dss = 0.5 * np.mean(fd(tt, e_all)) * np.mean(fd(beta - tt, e_all))
        else:
            g = np.mean(1 / (1 + np.exp(-beta * e_all)))
            dss = 0.5 * g * (1 - g)
        assert abs(acc["dss"][t] - dss) < 1e-10, t


def test_free_fermions_checkerboard_direct():
    # checkerboard slice matrix (not the exact exp(-dtau K)): direct products at small beta
    side, L, beta = 4, 10, 1.0
    c = O.Chain(side, L, beta, pace=4, model=0, checkerboard=True, U=0.0, mu=0.2)
    c.seed(1); c.fields_random(); c.init()
    c.sweep_dynamic()
    Bs = slice_matrices(c, 1)
    n = c.N
    G00 = np.linalg.inv(np.eye(n) + chain_product(Bs, L, 0))
    for tau in range(1, L + 1):
        gt0 = chain_product(Bs, tau, 0) @ G00
        g0t = -G00 @ chain_product(Bs, L, tau)
        assert np.max(np.abs(c.dynamic_green(tau - 1, 1, 0) - gt0)) < 1e-11
        assert np.max(np.abs(c.dynamic_green(tau - 1, 1, 1) - g0t)) < 1e-11


# ----------------------------------------------------------------------------
# many-body brute force on 2x2
# ----------------------------------------------------------------------------
class Fock:
    """Fermion Fock space of n modes; state bit m = occupation of mode m, creation operators
    ordered by increasing mode index."""

    def __init__(self, n):
        self.n = n
        self.dim = 1 << n
        self.c = []
        for m in range(n):
            a = np.zeros((self.dim, self.dim))
            for s in range(self.dim):
                if s >> m & 1:
                    sgn = (-1) ** bin(s & ((1 << m) - 1)).count("1")
                    a[s ^ (1 << m), s] = sgn
            self.c.append(a)
        self.cd = [a.T.copy() for a in self.c]
        self.subsets = [[m for m in range(n) if s >> m & 1] for s in range(self.dim)]

    def lift(self, B):
        """Gamma(B): Gamma(B) c+_a Gamma(B)^-1 = sum_b B[b,a] c+_b, so <S'|Gamma(B)|S> = det B[S',S]."""
        out = np.zeros((self.dim, self.dim))
        for s in range(self.dim):
            S = self.subsets[s]
            for t in range(self.dim):
                T = self.subsets[t]
                if len(T) != len(S):
                    continue
                out[t, s] = 1.0 if not S else np.linalg.det(B[np.ix_(T, S)])
        return out


def fock_correlators(c, side):
    """Per-configuration dynamic observables of the chain's current fields by traces in Fock space."""
    n, L = c.N, c.L
    F = Fock(n)
    lat = O.lattice(side)
    nn = lat["nn"]
    fourier = lat["fourier"]            # (site index of displacement, k-star)
    disp = lat["disp"]
    res = {}
    lifted = {}
    for spin in (1, -1):
        lifted[spin] = [F.lift(B) for B in slice_matrices(c, spin)]

    def prod(spin, t2, t1):
        M = np.eye(F.dim)
        for t in range(t1 + 1, t2 + 1):
            M = lifted[spin][t - 1] @ M
        return M

    Z = {sp: np.trace(prod(sp, L, 0)) for sp in (1, -1)}

    def corr(spin, tau, A, Bop):
        """<A(tau) Bop(0)> within one spin species."""
        return np.trace(prod(spin, L, tau) @ A @ prod(spin, tau, 0) @ Bop) / Z[spin]

    def mean(spin, tau, A):
        return np.trace(prod(spin, L, tau) @ A @ prod(spin, tau, 0)) / Z[spin]

    # G(tau,0)(i,j) = <c_i(tau) c+_j(0)>, G(0,tau)(i,j) = -<c+_j(tau) c_i(0)>
    for spin in (1, -1):
        for tau in range(1, L + 1):
            res[("gt0", spin, tau)] = np.array([[corr(spin, tau, F.c[i], F.cd[j]) for j in range(n)] for i in range(n)])
            res[("g0t", spin, tau)] = np.array([[-corr(spin, tau, F.cd[j], F.c[i]) for j in range(n)] for i in range(n)])
    # measurement entry t uses tau = t for t >= 1 and tau = beta (slot L-1) for t = 0
    taus = [L] + list(range(1, L))
    nk = fourier.shape[1]
    gk = np.zeros((nk, L)); dos = np.zeros(L); dss = np.zeros(L); rho = 0.0
    num = [F.cd[i] @ F.c[i] for i in range(n)]
    cur = [F.cd[nn[r, 0]] @ F.c[r] - F.cd[r] @ F.c[nn[r, 0]] for r in range(n)]   # j_x(r) / (i t) per spin
    for t, tau in enumerate(taus):
        g_av = 0.0
        for spin in (1, -1):
            g_av = g_av + 0.5 * (res[("gt0", spin, tau)] if t > 0 else -res[("g0t", spin, L)])
        # at t = 0 the reference uses G(beta,beta) = G(0,0) = -G(0,beta)
        for k in range(nk):
            gk[k, t] = sum(g_av[j, i] * fourier[disp[i, j], k] for i in range(n) for j in range(n)) / n
        dos[t] = np.trace(g_av) / n
        # <Sz_i(tau) Sz_i(0)>, Sz = (n_up - n_dn) / 2
        v = 0.0
        for i in range(n):
            for s1 in (1, -1):
                for s2 in (1, -1):
                    if s1 == s2:
                        v += s1 * s2 * corr(s1, tau, num[i], num[i])
                    else:
                        v += s1 * s2 * mean(s1, tau, num[i]) * mean(s2, 0, num[i])
        dss[t] = 0.25 * v / n
        # sum_ij ff(i,j) <j_x(j,tau) j_x(i,0)>,  j_x = i t sum_spin (c+_{r+x} c_r - c+_r c_{r+x})
        for i in range(n):
            for j in range(n):
                d = disp[i, j]
                rx, ry = d % side, d // side
                ff = fourier[rx, 1] - fourier[ry, 1]
                w = 0.0
                for s1 in (1, -1):
                    for s2 in (1, -1):
                        if s1 == s2:
                            w += corr(s1, tau, cur[j], cur[i])
                        else:
                            w += mean(s1, tau, cur[j]) * mean(s2, 0, cur[i])
                rho += -1.0 * ff * w    # (i t)^2 = -t^2 with t = 1
    res["gk"] = gk; res["dos"] = dos; res["dss"] = dss; res["rho"] = 0.25 * rho / (n * n)
    return res


@pytest.mark.parametrize("model,cb,general", [(0, False, False), (0, False, True), (1, False, False),
                                              (0, True, False), (1, True, True)])
def test_dynamic_path_vs_fock_space(model, cb, general):
    side, L, beta = 2, 7, 1.4
    c = O.Chain(side, L, beta, pace=3, model=model, checkerboard=cb, U=3.0, mu=-0.3)
    c.set_dynamic_formula(general)
    c.seed(11); c.fields_random(); c.init()
    # move the chain off its initial fields, back to slice 0 (the state a dynamic sweep starts from)
    c.sweep_forward(); c.sweep_backward()
    c.clear_dynamic()
    c.sweep_dynamic()
    ref = fock_correlators(c, side)
    for spin in (1, -1):
        for tau in range(1, L + 1):
            assert np.max(np.abs(c.dynamic_green(tau - 1, spin, 0) - ref[("gt0", spin, tau)])) < 1e-11, (spin, tau)
            assert np.max(np.abs(c.dynamic_green(tau - 1, spin, 1) - ref[("g0t", spin, tau)])) < 1e-11, (spin, tau)
    c.dynamic_measure()
    acc = c.dynamic_accum()
    cs = acc["sign"]
    assert abs(cs) == 1.0 and acc["count"] == 1.0
    assert np.max(np.abs(acc["gk"] / cs - ref["gk"])) < 1e-11
    assert np.max(np.abs(acc["dos"] / cs - ref["dos"])) < 1e-11
    assert np.max(np.abs(acc["dss"] / cs - ref["dss"])) < 1e-11
    assert abs(acc["rho"] / cs - ref["rho"]) < 1e-11
    assert abs(ref["rho"]) > 1e-4    # the stiffness term is not trivially zero here


# ----------------------------------------------------------------------------
# recompute formulas and stack factorizations
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("model", [0, 1])
def test_formulas_and_factorizations_agree(model):
    kw = dict(pace=10, model=model, U=4.0, mu=-0.2)
    chains = [O.Chain(4, 80, 4.0, factorization=f, **kw) for f in (0, 0, 2)]
    chains[1].set_dynamic_formula(True)
    for c in chains:
        c.seed(42); c.fields_random(); c.init()
        c.sweep_forward(); c.sweep_backward()
        c.sweep_dynamic()
        assert c.wrap_error < 1e-7
    a = chains[0]
    for b in chains[1:]:
        for slot in range(0, 80, 3):
            for which in (0, 1):
                assert np.max(np.abs(a.dynamic_green(slot, 1, which) - b.dynamic_green(slot, 1, which))) < 1e-9
            assert np.max(np.abs(a.slice_green(slot, -1) - b.slice_green(slot, -1))) < 1e-9


def test_dynamic_sweep_leaves_a_consistent_chain():
    # after the dynamic sweep the chain sits at slice L with the left stack full: the next
    # (updating) backward sweep must end on the G(0) a fresh start from its fields gives
    c = O.Chain(4, 40, 3.0, pace=6, model=0, U=4.0, mu=0.0)
    c.seed(9); c.fields_random(); c.init()
    f0 = c.fields()
    c.sweep_dynamic()
    assert np.array_equal(c.fields(), f0)          # no field updates
    c.sweep_backward()
    g = c.green(1)
    d = O.Chain(4, 40, 3.0, pace=6, model=0, U=4.0, mu=0.0)
    d.set_fields(c.fields()); d.init()
    assert np.max(np.abs(g - d.green(1))) < 1e-10
