"""Pins of the oracle's model layer: exp(-dtau K), the checkerboard break-up and
the five B-matrix products (src/model/*.cpp, src/checkerboard/square.cpp)."""
import numpy as np
import pytest
import scipy.linalg

import _oracle as O


def kmat(side, t, mu):
    lat = O.lattice(side)
    n = side * side
    return -t * lat["hopping"] + mu * np.eye(n)


@pytest.mark.parametrize("side,dtau,mu", [(4, 0.05, 0.0), (6, 0.1, -0.3), (8, 0.05, -0.5), (3, 0.2, 0.7)])
def test_expm_vs_scipy_and_spectral(side, dtau, mu):
    K = kmat(side, 1.0, mu)
    E = O.expm(-dtau * K)
    assert np.max(np.abs(E - scipy.linalg.expm(-dtau * K))) < 1e-14
    w, V = np.linalg.eigh(K)
    assert np.max(np.abs(E - (V * np.exp(-dtau * w)) @ V.T)) < 1e-13
    # plane waves: eigenvalues of K are -2t(cos kx + cos ky) + mu
    ks = 2 * np.pi * np.arange(side) / side
    eps = np.sort((-2 * (np.cos(ks)[:, None] + np.cos(ks)[None, :]) + mu).ravel())
    assert np.allclose(np.sort(w), eps, atol=1e-12)


def test_expm_large_norm_scaling():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((10, 10)) * 3
    assert np.max(np.abs(O.expm(A) - scipy.linalg.expm(A)) / np.max(np.abs(scipy.linalg.expm(A)))) < 1e-12


def make_chain(side=4, L=8, beta=0.8, model=0, cb=False, mu=0.0, U=4.0, seed=3):
    c = O.Chain(side, L, beta, pace=4, model=model, checkerboard=cb, U=U, mu=mu)
    c.seed(seed)
    c.fields_random()
    return c


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("cb", [False, True])
def test_B_times_invB_is_identity(model, cb):
    c = make_chain(model=model, cb=cb, mu=-0.3)
    n = c.N
    rng = np.random.default_rng(1)
    M = rng.standard_normal((n, n))
    for t in [0, 1, 5, c.L]:
        for spin in (+1, -1):
            # invB_left(B_left(M)) = M ; invB_right(B_right(M)) = M
            assert np.allclose(c.mult(2, t, spin, c.mult(0, t, spin, M)), M, atol=1e-12)
            assert np.allclose(c.mult(3, t, spin, c.mult(1, t, spin, M)), M, atol=1e-12)
            # transB_left(M) = B^T M with B = B_left(I)
            B = c.mult(0, t, spin, np.eye(n))
            assert np.allclose(c.mult(4, t, spin, M), B.T @ M, atol=1e-12)
            # B_right(M) = M B
            assert np.allclose(c.mult(1, t, spin, M), M @ B, atol=1e-12)


@pytest.mark.parametrize("model", [0, 1])
def test_B_matrix_closed_form_dense(model):
    c = make_chain(model=model)
    n = c.N
    f = c.fields()
    a = c.alpha
    dtau = 0.8 / 8
    assert abs(a - np.arccosh(np.exp(0.5 * dtau * 4.0))) < 1e-15
    expK = scipy.linalg.expm(-dtau * kmat(4, 1.0, 0.0))
    for t in [0, 3, 8]:
        row = (c.L - 1) if t == 0 else t - 1          # slice 0 is slice L (PBC)
        for spin in (+1, -1):
            sv = spin if model == 0 else 1.0          # attractive ignores the spin
            B = np.diag(np.exp(sv * a * f[row])) @ expK
            assert np.allclose(c.mult(0, t, spin, np.eye(n)), B, atol=1e-13)


@pytest.mark.parametrize("side", [4, 6, 8])
def test_checkerboard_plaquette_and_trotter_error(side):
    dtau = 0.05
    c = O.Chain(side, 8, 8 * dtau, pace=4, model=0, checkerboard=True, U=0.0, mu=0.0)
    n = c.N
    E_cb = c.mult(5, 1, 1, np.eye(n))
    E = scipy.linalg.expm(-dtau * kmat(side, 1.0, 0.0))
    err = np.max(np.abs(E_cb - E))
    if side == 4:
        assert err < 1e-14                               # on the 4-ring the two bond sets commute
    else:
        assert 1e-6 < err < 2 * dtau ** 2                # O(dtau^2) break-up error
    # exact closed form: E = E_A E_B with E_X = prod of 4x4 plaquettes exp(dtau t h_plaq)
    ch, sh = np.cosh(dtau), np.sinh(dtau)
    A2 = np.array([[ch, sh], [sh, ch]])
    P = np.kron(A2, A2)

    def sub(start):
        M = np.eye(n)
        for x in range(start, side, 2):
            for y in range(start, side, 2):
                idx = [(x % side) + side * (y % side), ((x + 1) % side) + side * (y % side),
                       (x % side) + side * ((y + 1) % side), ((x + 1) % side) + side * ((y + 1) % side)]
                M[np.ix_(idx, idx)] = P
        return M
    EA, EB = sub(0), sub(1)
    assert np.allclose(E_cb, EA @ EB, atol=1e-14)
    assert np.allclose(c.mult(6, 1, 1, np.eye(n)), EA @ EB, atol=1e-14)      # right hook, same product
    assert np.allclose(c.mult(9, 1, 1, np.eye(n)), (EA @ EB).T, atol=1e-14)  # trans hook
    assert np.allclose(c.mult(7, 1, 1, np.eye(n)) @ E_cb, np.eye(n), atol=1e-13)


def test_checkerboard_mu_sign_quirk():
    # the plaquettes carry exp(+dtau mu / 2) and every site sits in two plaquettes:
    # checkerboard expK ~ exp(+dtau mu) while dense expK ~ exp(-dtau mu)
    side, dtau, mu = 4, 0.05, -0.5
    n = side * side
    cb = O.Chain(side, 8, 8 * dtau, pace=4, checkerboard=True, U=0.0, mu=mu)
    dn = O.Chain(side, 8, 8 * dtau, pace=4, checkerboard=False, U=0.0, mu=mu)
    cb0 = O.Chain(side, 8, 8 * dtau, pace=4, checkerboard=True, U=0.0, mu=0.0)
    dn0 = O.Chain(side, 8, 8 * dtau, pace=4, checkerboard=False, U=0.0, mu=0.0)
    r_cb = cb.mult(5, 1, 1, np.eye(n)) / cb0.mult(5, 1, 1, np.eye(n))
    r_dn = dn.mult(5, 1, 1, np.eye(n))[0, 0] / dn0.mult(5, 1, 1, np.eye(n))[0, 0]
    assert np.allclose(r_cb[np.abs(cb0.mult(5, 1, 1, np.eye(n))) > 1e-12], np.exp(+dtau * mu))
    assert abs(r_dn - np.exp(-dtau * mu)) < 1e-13
