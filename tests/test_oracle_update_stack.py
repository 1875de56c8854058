"""Pins of the oracle's Metropolis ratio, rank-1 update, UDV stacks and
stabilized Green's functions (src/model/*_hubbard.cpp:112-155,
src/svd_stack.cpp, include/utils/numerical_stable.hpp) against brute force."""
import numpy as np
import pytest

import _oracle as O


def brute_M(chain, spin, model):
    """B_L ... B_1 for one spin, built from B_left(I)."""
    n = chain.N
    M = np.eye(n)
    for t in range(1, chain.L + 1):
        M = chain.mult(0, t, spin, np.eye(n)) @ M
    return M


@pytest.mark.parametrize("model", [0, 1])
def test_initial_green_is_direct_inverse(model):
    c = O.Chain(2, 8, 1.0, pace=3, model=model, U=4.0)
    c.seed(7); c.fields_random(); c.init()
    for spin in (+1, -1):
        G = np.linalg.inv(np.eye(c.N) + brute_M(c, spin, model))
        assert np.max(np.abs(c.green(spin) - G)) < 1e-13


@pytest.mark.parametrize("model", [0, 1])
def test_ratio_and_rank1_vs_brute_force(model):
    c = O.Chain(2, 6, 1.2, pace=3, model=model, U=5.0, mu=0.0 if model == 0 else -0.3)
    c.seed(11); c.fields_random(); c.init()
    n, L = c.N, c.L
    f = c.fields()
    for i in range(n):
        # slice 0 == slice L: flipping row L-1 changes the leftmost factor B_L
        Mu, Md = brute_M(c, +1, model), brute_M(c, -1, model)
        w0 = np.linalg.det(np.eye(n) + Mu) * np.linalg.det(np.eye(n) + Md)
        r = c.update_ratio(L - 1, i)
        f2 = f.copy(); f2[L - 1, i] *= -1
        c.set_fields(f2)
        Mu2, Md2 = brute_M(c, +1, model), brute_M(c, -1, model)
        c.set_fields(f)
        w1 = np.linalg.det(np.eye(n) + Mu2) * np.linalg.det(np.eye(n) + Md2)
        expect = w1 / w0
        if model == 1:   # attractive weight carries exp(-alpha sum s)
            expect *= np.exp(2 * c.alpha * f[L - 1, i])
        assert abs(r - expect) < 1e-12 * max(1.0, abs(r))
        # Sherman-Morrison update reproduces the re-inverted G
        c.update_greens(L - 1, i)
        c.flip(L - 1, i)
        f = c.fields()
        assert np.max(np.abs(c.green(+1) - np.linalg.inv(np.eye(n) + Mu2))) < 1e-12
        assert np.max(np.abs(c.green(-1) - np.linalg.inv(np.eye(n) + Md2))) < 1e-12


@pytest.mark.parametrize("fact", [0, 1, 2])
def test_udv_reconstruction(fact):
    rng = np.random.default_rng(5)
    for n in (4, 16, 33):
        A = rng.standard_normal((n, n))
        U, s, V = O.udv(A, fact)
        assert np.max(np.abs(U @ np.diag(s) @ V.T - A)) < 1e-12 * np.max(np.abs(A)) * n
        assert np.max(np.abs(U.T @ U - np.eye(n))) < 1e-13 * n
        if fact == 0:
            assert np.max(np.abs(V.T @ V - np.eye(n))) < 1e-13 * n
            assert np.all(np.diff(s) <= 0)
            assert np.allclose(s, np.linalg.svd(A, compute_uv=False), rtol=1e-12)


@pytest.mark.parametrize("fact", [0, 1, 2])
def test_udv_graded_matrix(fact):
    # column-graded input X diag(d), the shape met in the stacks ((A U) D): every
    # column, however small, is reproduced to relative accuracy
    rng = np.random.default_rng(9)
    n = 12
    X = rng.standard_normal((n, n)) + 2 * np.eye(n)
    d = np.logspace(12, -12, n)
    A = X * d[None, :]
    U, s, V = O.udv(A, fact)
    R = U @ np.diag(s) @ V.T
    colerr = np.linalg.norm(R - A, axis=0) / np.linalg.norm(A, axis=0)
    assert np.max(colerr) < 1e-12
    assert np.max(np.abs(U.T @ U - np.eye(n))) < 1e-13


def test_fullpiv_solve_and_inverse():
    rng = np.random.default_rng(2)
    for n in (3, 10, 40):
        H = rng.standard_normal((n, n)) + n * np.eye(n) * 0.1
        B = rng.standard_normal((n, n))
        assert np.allclose(O.fullpiv_solve(H, B), np.linalg.solve(H, B), atol=1e-11)
        assert np.allclose(O.inverse(H), np.linalg.inv(H), atol=1e-11)


def test_greens_00_bb_vs_direct():
    rng = np.random.default_rng(4)
    n = 10
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(3, -3, n)
    G = O.greens_00_bb(U, s, V)
    assert np.max(np.abs(G - np.linalg.inv(np.eye(n) + U @ np.diag(s) @ V.T))) < 1e-12


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("cb", [False, True])
def test_stabilized_green_matches_direct_every_slice(model, cb):
    # small beta: the direct inverse is accurate, so every stored slice must agree with it
    side = 2 if not cb else 4
    c = O.Chain(side, 12, 1.0, pace=5, model=model, checkerboard=cb, U=4.0, mu=-0.2)
    c.seed(1); c.fields_random(); c.init()
    before = c.fields()
    c.sweep_forward()
    after = c.fields()
    acc = c.accepts()
    assert np.array_equal(after, before * np.where(acc > 0, -1.0, 1.0))
    n, L = c.N, c.L
    for t in range(1, L + 1):
        # when slot t-1 is stored, rows < t hold their new values, rows >= t the old ones
        f_t = np.vstack([after[:t], before[t:]])
        c.set_fields(f_t)
        for spin in (+1, -1):
            B = [c.mult(0, s, spin, np.eye(n)) for s in range(1, L + 1)]
            M = np.eye(n)
            for s in list(range(t + 1, L + 1)) + list(range(1, t + 1)):   # B_t..B_1 B_L..B_{t+1}
                M = B[s - 1] @ M
            G = np.linalg.inv(np.eye(n) + M)
            assert np.max(np.abs(c.slice_green(t - 1, spin) - G)) < 1e-10, (t, spin)
    c.set_fields(after)


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("fact", [1, 2])
def test_green_invariant_under_factorization(model, fact):
    # SVD stacks (the reference) and column-pivoted / pre-pivoted QR stacks give the same G to rounding
    kw = dict(pace=10, model=model, U=4.0, mu=0.0)
    a = O.Chain(4, 80, 4.0, factorization=0, **kw)
    b = O.Chain(4, 80, 4.0, factorization=fact, **kw)
    for c in (a, b):
        c.seed(42); c.fields_random(); c.init()
    assert np.max(np.abs(a.green(1) - b.green(1))) < 1e-11
    for _ in range(2):
        a.sweep_forward(); b.sweep_forward(); a.sweep_backward(); b.sweep_backward()
        assert np.array_equal(a.fields(), b.fields())
        for slot in range(0, 80, 7):
            assert np.max(np.abs(a.slice_green(slot, 1) - b.slice_green(slot, 1))) < 1e-9
    assert np.max(np.abs(a.green(-1) - b.green(-1))) < 1e-10


def test_prepivot_green_at_low_temperature():
    # beta = 12, U = 8: scales spanning ~1e20; SVD and pre-pivoted QR stacks still agree
    kw = dict(pace=10, model=1, U=8.0, mu=0.3)
    a = O.Chain(6, 120, 12.0, factorization=0, **kw)
    b = O.Chain(6, 120, 12.0, factorization=2, **kw)
    for c in (a, b):
        c.seed(5); c.fields_random(); c.init()
    assert np.max(np.abs(a.green(1) - b.green(1))) < 1e-10
    a.sweep_forward(); b.sweep_forward()
    assert np.array_equal(a.fields(), b.fields())
    assert np.max(np.abs(a.green(1) - b.green(1))) < 1e-10
