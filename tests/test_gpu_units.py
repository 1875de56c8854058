"""GPU parity of the sweep's building blocks run one at a time through their single-kernel
entry points (dqmc_k_wrap, dqmc_k_rank1, dqmc_k_udv_push, dqmc_k_greens_recompute) against
the oracle and against identities the mathematics fixes.

Bars (FP64; the device GEMMs accumulate in another order than the oracle's loops):
  B-matrix products and wraps      max |dM| <= 1e-12 x max |M_oracle|   (ModelBase mult_*);
                                   in mixed precision (FP32 matrix cores) <= 1e-5 x max |M_oracle|
  single-flip update               max |dG| <= 1e-13 against the oracle, 1e-11 against the
                                   Sherman-Morrison closed form G' = G [1 + Delta (1 - G)]^-1
  UDV push (pre-pivoted QR)        U, V within 1e-10, D within 1e-12 relative, of the oracle's
                                   pre-pivoted QR of A U_top D_top (SvdStack::push with that
                                   factorization); U^T U = 1 within 1e-13
  Green's function recompute       1e-10 against the direct inverse at high temperature and
                                   against the reference's SVD stack with its one-sided formula
                                   (compute_greens_00_bb) at beta = 8
"""
import numpy as np
import pytest

import _oracle as O
import dqmc_amd as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return D.Context(0)


def chain(side, L, beta, model=0, cb=False, U=4.0, mu=0.0, seed=7):
    orc = O.Chain(side, L, beta, pace=5, model=model, checkerboard=cb, U=U, mu=mu)
    orc.seed(seed)
    orc.fields_random()
    p = D.make_params(side, L, beta, pace=5, model=model, checkerboard=cb, U=U, mu=mu, seed=seed)
    return orc, p, orc.fields().astype(np.int8)


WRAP_CASES = [
    # side, L, beta, model, cb, U, mu
    (4, 12, 2.0, 0, False, 4.0, -0.3),
    (6, 10, 1.5, 0, True, 6.0, 0.2),
    (4, 8, 1.0, 1, False, 3.0, -0.5),
    (4, 8, 1.0, 1, True, 3.0, 0.1),
    (12, 20, 2.0, 0, False, 4.0, 0.0),
    (12, 20, 2.0, 0, True, 4.0, 0.0),
    (16, 8, 1.0, 0, False, 4.0, -0.2),   # N = 256 dense: the Kronecker form at its largest tile
    (8, 8, 1.0, 1, False, 4.0, -0.5),
]


def oracle_op(orc, op, t, spin, m):
    if op <= 4:
        return orc.mult(op, t, spin, m)
    if op == 5:   # B M B^-1 (DqmcWalker::wrap_from_0_to_beta)
        return orc.mult(3, t, spin, orc.mult(0, t, spin, m))
    return orc.mult(2, t, spin, orc.mult(1, t, spin, m))   # B^-1 M B (wrap_from_beta_to_0)


@pytest.mark.parametrize("case", WRAP_CASES, ids=[f"s{c[0]}m{c[3]}cb{int(c[4])}" for c in WRAP_CASES])
def test_wrap_and_products_match_oracle(ctx, case):
    side, L, beta, model, cb, U, mu = case
    print(f"[gpu-units] wrap {case}", flush=True)
    orc, p, f = chain(side, L, beta, model, cb, U, mu)
    n = side * side
    rng = np.random.default_rng(side + 10 * model + int(cb))
    ek = None if cb else orc.expK(0)
    ik = None if cb else orc.expK(1)
    worst = 0.0
    for t in (0, 1, L // 2, L):
        for spin in (1, -1):
            for op in range(7):
                m = np.asfortranarray(rng.standard_normal((n, n)))
                want = oracle_op(orc, op, t, spin, m.copy(order="F"))
                got = ctx.k_wrap(p, f, t, spin, op, m, ek, ik)
                err = np.max(np.abs(got - want)) / np.max(np.abs(want))
                worst = max(worst, err)
                assert err < 1e-12, (t, spin, op, err)
    print(f"[gpu-units] wrap {case}: max relative error {worst:.2e}", flush=True)


@pytest.mark.parametrize("side", [4, 12])
def test_mixed_precision_products(ctx, side):
    # mixed precision: the dense B products and wraps on FP32 matrix cores with FP32
    # accumulation; bar 1e-5 of max |M| (FP32 rounding 6e-8 over N-term sums)
    orc, _, f = chain(side, 10, 2.0, U=6.0)
    p = D.make_params(side, 10, 2.0, pace=5, U=6.0, seed=7, precision=D.PREC_MIXED)
    n = side * side
    m = np.asfortranarray(np.random.default_rng(5).standard_normal((n, n)))
    worst = 0.0
    for op in range(7):
        want = oracle_op(orc, op, 4, -1, m.copy(order="F"))
        got = ctx.k_wrap(p, f, 4, -1, op, m, orc.expK(0), orc.expK(1))
        worst = max(worst, np.max(np.abs(got - want)) / np.max(np.abs(want)))
    print(f"[gpu-units] mixed precision side {side}: max relative error {worst:.2e}", flush=True)
    assert 1e-12 < worst < 1e-5      # FP32 arithmetic shows, within its bound


def test_wrap_closed_form_exp_k(ctx):
    # without an explicit exp(-dtau K) the batch's closed form is used: same products to rounding
    orc, p, f = chain(4, 10, 2.0, mu=-0.2)
    m = np.asfortranarray(np.random.default_rng(1).standard_normal((16, 16)))
    got = ctx.k_wrap(p, f, 3, 1, "wrap_up", m)
    want = oracle_op(orc, 5, 3, 1, m.copy(order="F"))
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))


def test_wrap_inverse_pairs(ctx):
    # B^-1 (B M) = M and the two wraps undo each other
    orc, p, f = chain(6, 10, 2.0, cb=True, U=5.0)
    m = np.asfortranarray(np.random.default_rng(2).standard_normal((36, 36)))
    for a, b in (("B_left", "invB_left"), ("B_right", "invB_right"), ("wrap_up", "wrap_down")):
        back = ctx.k_wrap(p, f, 4, -1, b, ctx.k_wrap(p, f, 4, -1, a, m))
        assert np.max(np.abs(back - m)) < 1e-12, (a, b)


@pytest.mark.parametrize("model", [0, 1])
def test_rank1_matches_oracle_and_sherman_morrison(ctx, model):
    print(f"[gpu-units] rank1 model {model}", flush=True)
    orc, p, f = chain(6, 8, 1.0, model=model, U=4.0, mu=-0.1)
    orc.init()
    n = 36
    rng = np.random.default_rng(3 + model)
    for t, i in ((0, 0), (3, 17), (7, 35)):
        gu = orc.green(1) + 0.05 * rng.standard_normal((n, n))
        gd = orc.green(-1) + 0.05 * rng.standard_normal((n, n))
        got_u, got_d = ctx.k_rank1(p, f, t, i, gu, gd)
        orc.set_green(1, gu); orc.set_green(-1, gd)
        orc.update_greens(t, i)
        assert np.max(np.abs(got_u - orc.green(1))) < 1e-13, (t, i)
        assert np.max(np.abs(got_d - orc.green(-1))) < 1e-13, (t, i)
        # G' = G [1 + Delta (1 - G)]^-1 with Delta = m e_i e_i^T: the Green's function of
        # (1 + Delta) B when G = (1 + B)^-1 (spin up, and spin down of the repulsive model)
        s = f[t, i]
        for sigma, g, out in ((1, gu, got_u), (-1, gd, got_d)):
            if model == 1 and sigma == -1:
                continue   # the attractive update uses the spin-up factor for both matrices
            m = np.exp(-2 * sigma * orc.alpha * s) - 1
            delta = np.zeros((n, n)); delta[i, i] = m
            closed = g @ np.linalg.inv(np.eye(n) + delta @ (np.eye(n) - g))
            assert np.max(np.abs(out - closed)) < 1e-11, (t, i, sigma)
    # spin down may be left out
    gu = orc.green(1)
    alone, none = ctx.k_rank1(p, f, 2, 5, gu)
    assert none is None
    orc.set_green(1, gu); orc.update_greens(2, 5)
    assert np.max(np.abs(alone - orc.green(1))) < 1e-13


def slice_matrices(side, L, beta, U=4.0, mu=0.0, spin=1, seed=11):
    orc, _, _ = chain(side, L, beta, U=U, mu=mu, seed=seed)
    n = side * side
    return [orc.mult(0, t, spin, np.asfortranarray(np.eye(n))) for t in range(1, L + 1)]


def push_oracle(A, top):
    """SvdStack::push with the pre-pivoted QR factorization (src/svd_stack.cpp:34-61)."""
    if top is None:
        return O.udv(A, 2)
    U0, D0, V0 = top
    U1, D1, V1 = O.udv(A @ U0 * D0, 2)
    return U1, D1, V0 @ V1


@pytest.mark.parametrize("side", [4, 12])
def test_udv_push_matches_oracle(ctx, side):
    print(f"[gpu-units] udv push side {side}", flush=True)
    Bs = slice_matrices(side, 20, 3.0)
    n = side * side
    top_g = top_o = None
    for k in range(0, 20, 4):     # each push takes a product of 4 slices, as a pace-4 stack does
        A = np.linalg.multi_dot(Bs[k:k + 4][::-1])
        top_g = ctx.k_udv_push(A, top_g)
        top_o = push_oracle(A, top_o)
        U, Dg, V = top_g
        Uo, Do, Vo = top_o
        assert np.max(np.abs(U.T @ U - np.eye(n))) < 1e-13
        assert np.all(Dg > 0)
        assert np.max(np.abs(Dg / Do - 1)) < 1e-12, k
        assert np.max(np.abs(U - Uo)) < 1e-10, k
        assert np.max(np.abs(V - Vo)) < 1e-10 * np.max(np.abs(Vo)), k
    # the stack holds the product of all 20 slices
    full = np.linalg.multi_dot(Bs[::-1])
    U, Dg, V = top_g
    assert np.max(np.abs((U * Dg) @ V.T - full)) < 1e-10 * np.max(np.abs(full))


def build_stacks(ctx, Bs, k):
    """Left stack holds B_L ... B_{k+1}, right stack B_1^T ... B_k^T (as the engine's stacks)."""
    left = right = None
    for B in Bs[k:]:
        left = ctx.k_udv_push(B, left)
    for B in Bs[:k][::-1]:
        right = ctx.k_udv_push(B.T, right)
    return left, right


@pytest.mark.parametrize("side,beta,split", [(4, 1.0, 5), (4, 1.0, 0), (4, 1.0, 10), (6, 1.5, 3)])
def test_greens_recompute_direct_inverse(ctx, side, beta, split):
    print(f"[gpu-units] greens {side} beta {beta} split {split}", flush=True)
    Bs = slice_matrices(side, 10, beta)
    n = side * side
    left, right = build_stacks(ctx, Bs, split)
    G = ctx.k_greens_recompute(left, right)
    want = np.linalg.inv(np.eye(n) + np.linalg.multi_dot([np.eye(n)] + Bs[::-1]))
    assert np.max(np.abs(G - want)) < 1e-10


@pytest.mark.parametrize("side", [4, 12])
def test_greens_recompute_low_temperature(ctx, side):
    # beta = 8, U = 4: the slice product spans many decades; the reference's own path (SVD stack,
    # compute_greens_00_bb with its full-pivoting QR solve) is the comparison, and every split of
    # the slices between the two stacks gives the same G
    print(f"[gpu-units] greens low T side {side}", flush=True)
    L = 80
    Bs = slice_matrices(side, L, 8.0, U=4.0, mu=-0.2)
    top = None
    for k in range(0, L, 10):
        A = np.linalg.multi_dot(Bs[k:k + 10][::-1])
        if top is None:
            top = O.udv(A, 0)
        else:
            U1, D1, V1 = O.udv(A @ top[0] * top[1], 0)
            top = (U1, D1, top[2] @ V1)
    want = O.greens_00_bb(*top)
    worst = 0.0
    for split in (0, 30, L):
        left, right = build_stacks(ctx, Bs, split)
        G = ctx.k_greens_recompute(left, right)
        worst = max(worst, np.max(np.abs(G - want)))
    print(f"[gpu-units] greens low T side {side}: max error {worst:.2e}", flush=True)
    assert worst < 1e-10


def test_unit_argument_errors(ctx):
    _, p, f = chain(4, 8, 1.0)
    m = np.eye(16)
    for bad in [dict(t=9), dict(t=-1), dict(op=7), dict(spin=0)]:
        kw = dict(t=1, spin=1, op=0); kw.update(bad)
        with pytest.raises(D.DqmcError):
            ctx.k_wrap(p, f, kw["t"], kw["spin"], kw["op"], m)
    with pytest.raises(D.DqmcError):
        ctx.k_rank1(p, f, 8, 0, m)          # slice out of range
    with pytest.raises(D.DqmcError):
        ctx.k_rank1(p, f, 0, 16, m)         # site out of range
    with pytest.raises(D.DqmcError):
        ctx.k_udv_push(np.eye(5))           # N must be side * side
    with pytest.raises(D.DqmcError):
        ctx.k_udv_push(np.eye(484))         # larger than 448 sites
    cb = D.make_params(4, 8, 1.0, checkerboard=True)
    with pytest.raises(D.DqmcError):
        ctx.k_wrap(cb, f, 1, 1, 0, m, np.eye(16), np.eye(16))   # exp(K) given to a checkerboard chain


def test_kron_form_selected(ctx):
    """Dense chains on lattices up to 16 x 16 multiply in Kronecker form (exp(-dtau K) = X (x) Y
    over the two lattice directions, side 2 with its doubled hopping included); larger dense
    lattices and the checkerboard break-up do not."""
    sides = [2, 4, 8, 12, 16, 18]
    ps = [D.make_params(s, 8, 1.0, pace=4, U=4.0, mu=-0.3, seed=s) for s in sides]
    ps.append(D.make_params(8, 8, 1.0, pace=4, checkerboard=True, U=4.0, seed=1))
    b = D.Batch(ctx, ps)
    got = [b.info(c).kron for c in range(len(ps))]
    b.close()
    assert got == [1, 1, 1, 1, 1, 0, 0], got


def test_kron_falls_back_for_a_nonseparable_exp_k(ctx):
    """An exp(K) that does not factor over the lattice directions (here perturbed by 1e-6) is
    multiplied densely, and the products still equal the oracle's with that matrix."""
    side, L = 6, 8
    orc, p, f = chain(side, L, 1.0, U=4.0, mu=-0.2)
    n = side * side
    rng = np.random.default_rng(5)
    pert = rng.standard_normal((n, n)) * 1e-6
    ek = orc.expK(0) + (pert + pert.T)
    ik = np.linalg.inv(ek)
    orc.set_expK(np.asfortranarray(ek), np.asfortranarray(ik))
    b = D.Batch(ctx, [p])
    b.set_expk(0, ek, ik)
    assert b.info(0).kron == 0
    b.set_expk(0, *[np.asfortranarray(x) for x in b.expk(0)])   # (unchanged matrices: still dense)
    assert b.info(0).kron == 0
    b.close()
    for op in range(7):
        m = np.asfortranarray(rng.standard_normal((n, n)))
        want = oracle_op(orc, op, 3, 1, m.copy(order="F"))
        got = ctx.k_wrap(p, f, 3, 1, op, m, ek, ik)
        assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want)), op


def test_kron_set_expk_switches_form(ctx):
    """Setting the closed-form exp(K) back on a chain that held a non-separable one restores the
    Kronecker form, and a sweep of the chain in either form agrees with the oracle's."""
    side, L = 4, 12
    b = D.Batch(ctx, [D.make_params(side, L, 2.0, pace=4, U=4.0, mu=-0.3, seed=3)])
    ek, ik = b.expk(0)
    assert b.info(0).kron == 1
    pert = ek + 1e-6 * np.eye(side * side)[::-1]
    b.set_expk(0, pert, np.linalg.inv(pert))
    assert b.info(0).kron == 0
    b.set_expk(0, ek, ik)
    assert b.info(0).kron == 1
    b.close()
