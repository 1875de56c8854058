"""ctypes binding of the oracle (oracle/refcpu.cpp).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "librefcpu.so")

OBS_NAMES = ["sign", "filling_number", "double_occupancy", "kinetic_energy", "momentum_distribution",
             "local_spin_corr", "spin_density_structure_factor", "charge_density_structure_factor",
             "s_wave_pairing_corr"]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _lib = C.CDLL(LIB_PATH)
        _lib.rc_create.restype = C.c_void_p
        _lib.rc_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_double, C.c_int, C.c_int]
        vp = [C.c_void_p]
        for name in ["rc_destroy", "rc_fields_random", "rc_init", "rc_sweep_forward", "rc_sweep_backward",
                     "rc_measure", "rc_clear_measure"]:
            getattr(_lib, name).argtypes = vp
        _lib.rc_space_size.argtypes = vp
        _lib.rc_seed.argtypes = [C.c_void_p, C.c_longlong]
        _lib.rc_rng_state.argtypes = vp
        _lib.rc_rng_state.restype = C.c_ulonglong
        _lib.rc_set_rng_state.argtypes = [C.c_void_p, C.c_ulonglong]
        for name in ["rc_set_fields", "rc_get_fields"]:
            getattr(_lib, name).argtypes = [C.c_void_p, C.c_void_p]
        for name in ["rc_get_green", "rc_set_green"]:
            getattr(_lib, name).argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        _lib.rc_get_slice_green.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        _lib.rc_get_slice_sign.argtypes = [C.c_void_p, C.c_int]
        _lib.rc_get_slice_sign.restype = C.c_double
        for name in ["rc_sign", "rc_wrap_error", "rc_alpha"]:
            getattr(_lib, name).argtypes = vp
            getattr(_lib, name).restype = C.c_double
        _lib.rc_get_accepts.argtypes = [C.c_void_p, C.c_void_p]
        _lib.rc_get_measure.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.rc_get_expK.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        _lib.rc_set_expK.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.rc_mult.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        _lib.rc_update_ratio.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lib.rc_update_ratio.restype = C.c_double
        _lib.rc_update_greens.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lib.rc_flip.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lib.rc_lattice.argtypes = [C.c_int] + [C.c_void_p] * 6
        _lib.rc_lattice.restype = C.c_int
        _lib.rc_rng_bernoulli.argtypes = [C.c_longlong, C.c_int, C.c_double, C.c_void_p]
        _lib.rc_rng_canonical.argtypes = [C.c_longlong, C.c_int, C.c_void_p]
        _lib.rc_expm.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        _lib.rc_udv.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.rc_greens_00_bb.argtypes = [C.c_int] + [C.c_void_p] * 4
        _lib.rc_inverse.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        _lib.rc_fullpiv_solve.argtypes = [C.c_int] + [C.c_void_p] * 3
        for name in ["rc_sweep_dynamic", "rc_dynamic_measure", "rc_clear_dynamic"]:
            getattr(_lib, name).argtypes = vp
        _lib.rc_set_dynamic_formula.argtypes = [C.c_void_p, C.c_int]
        _lib.rc_get_dynamic_green.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        _lib.rc_get_dynamic_measure.argtypes = [C.c_void_p, C.c_void_p]
        _lib.rc_set_stable_sign.argtypes = [C.c_void_p, C.c_int]
        _lib.rc_set_slice.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_double]
        _lib.rc_init_sign.argtypes = [C.c_void_p, C.c_int]
        _lib.rc_init_sign.restype = C.c_double
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lattice(side):
    n = side * side
    nk = (side // 2 + 1) * (side // 2 + 2) // 2
    nn = np.zeros(n * 4, np.int32)
    disp = np.zeros(n * n, np.int32)
    ks = np.zeros(nk * 2)
    four = np.zeros(n * nk)
    hop = np.zeros(n * n)
    pts = np.zeros(3, np.int32)
    got = lib().rc_lattice(side, _p(nn), _p(disp), _p(ks), _p(four), _p(hop), _p(pts))
    assert got == nk
    return dict(nn=nn.reshape(n, 4), disp=disp.reshape(n, n), kstars=ks.reshape(nk, 2),
                fourier=four.reshape(n, nk), hopping=hop.reshape(n, n, order="F"), points=pts)


def bernoulli_stream(seed, n, p):
    out = np.zeros(n, np.uint8)
    lib().rc_rng_bernoulli(seed, n, p, _p(out))
    return out


def canonical_stream(seed, n):
    out = np.zeros(n)
    lib().rc_rng_canonical(seed, n, _p(out))
    return out


def expm(a):
    a = np.asfortranarray(a, dtype=np.float64)
    out = np.zeros_like(a, order="F")
    lib().rc_expm(a.shape[0], _p(a), _p(out))
    return out


def udv(a, factorization=0):
    a = np.asfortranarray(a, dtype=np.float64)
    n = a.shape[0]
    U = np.zeros((n, n), order="F"); V = np.zeros((n, n), order="F"); s = np.zeros(n)
    lib().rc_udv(n, factorization, _p(a), _p(U), _p(s), _p(V))
    return U, s, V


def greens_00_bb(U, s, V):
    U = np.asfortranarray(U); V = np.asfortranarray(V); s = np.ascontiguousarray(s, dtype=np.float64)
    out = np.zeros_like(U, order="F")
    lib().rc_greens_00_bb(U.shape[0], _p(U), _p(s), _p(V), _p(out))
    return out


def fullpiv_solve(H, B):
    H = np.asfortranarray(H, dtype=np.float64); B = np.asfortranarray(B, dtype=np.float64)
    out = np.zeros_like(H, order="F")
    lib().rc_fullpiv_solve(H.shape[0], _p(H), _p(B), _p(out))
    return out


def inverse(A):
    A = np.asfortranarray(A, dtype=np.float64)
    out = np.zeros_like(A, order="F")
    lib().rc_inverse(A.shape[0], _p(A), _p(out))
    return out


class Chain:
    """One oracle Markov chain (Walker + Model + Lattice + SvdStacks)."""

    def __init__(self, side, time_size, beta, pace=10, model=0, checkerboard=False, t=1.0, U=4.0, mu=0.0,
                 momentum=-1, factorization=0, stable_sign=True):
        """stable_sign: the initial configuration sign from the stabilized factorization (the
        engine's definition) instead of the sign of det G (the reference's, rounding noise at low
        temperature); the two agree wherever det G is resolved (pinned in test_oracle_model)."""
        self.h = lib().rc_create(side, time_size, beta, pace, model, int(checkerboard), t, U, mu, momentum,
                                 factorization)
        lib().rc_set_stable_sign(self.h, int(bool(stable_sign)))
        self.N = side * side
        self.L = time_size

    def __del__(self):
        if getattr(self, "h", None):
            lib().rc_destroy(self.h)
            self.h = None

    def seed(self, s):
        lib().rc_seed(self.h, s)

    @property
    def rng_state(self):
        return lib().rc_rng_state(self.h)

    @rng_state.setter
    def rng_state(self, x):
        lib().rc_set_rng_state(self.h, x)

    def fields_random(self):
        lib().rc_fields_random(self.h)

    def set_fields(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.L, self.N)
        lib().rc_set_fields(self.h, _p(f))

    def fields(self):
        f = np.zeros((self.L, self.N))
        lib().rc_get_fields(self.h, _p(f))
        return f

    def init(self):
        lib().rc_init(self.h)

    def sweep_forward(self):
        lib().rc_sweep_forward(self.h)

    def sweep_backward(self):
        lib().rc_sweep_backward(self.h)

    def green(self, spin):
        g = np.zeros((self.N, self.N), order="F")
        lib().rc_get_green(self.h, spin, _p(g))
        return g

    def set_green(self, spin, g):
        g = np.asfortranarray(g, dtype=np.float64)
        lib().rc_set_green(self.h, spin, _p(g))

    def slice_green(self, slot, spin):
        g = np.zeros((self.N, self.N), order="F")
        lib().rc_get_slice_green(self.h, slot, spin, _p(g))
        return g

    def set_slice(self, slot, gu, gd, sign=1.0):
        a = np.asfortranarray(gu, dtype=np.float64); b = np.asfortranarray(gd, dtype=np.float64)
        lib().rc_set_slice(self.h, slot, _p(a), _p(b), float(sign))

    def slice_sign(self, slot):
        return lib().rc_get_slice_sign(self.h, slot)

    @property
    def sign(self):
        return lib().rc_sign(self.h)

    def init_sign(self, stable):
        """The initial sign of the last init by either definition."""
        return lib().rc_init_sign(self.h, int(bool(stable)))

    @property
    def wrap_error(self):
        return lib().rc_wrap_error(self.h)

    @property
    def alpha(self):
        return lib().rc_alpha(self.h)

    def accepts(self):
        a = np.zeros((self.L, self.N), np.uint8)
        lib().rc_get_accepts(self.h, _p(a))
        return a

    def measure(self):
        lib().rc_measure(self.h)

    def measure_accum(self):
        s = np.zeros(9); c = np.zeros(9)
        lib().rc_get_measure(self.h, _p(s), _p(c))
        return s, c

    def clear_measure(self):
        lib().rc_clear_measure(self.h)

    def expK(self, which=0):
        m = np.zeros((self.N, self.N), order="F")
        lib().rc_get_expK(self.h, which, _p(m))
        return m

    def set_expK(self, expK, invExpK):
        a = np.asfortranarray(expK); b = np.asfortranarray(invExpK)
        lib().rc_set_expK(self.h, _p(a), _p(b))

    def mult(self, op, t, spin, m):
        m = np.array(m, dtype=np.float64, order="F", copy=True)
        lib().rc_mult(self.h, op, t, spin, _p(m))
        return m

    def update_ratio(self, t, i):
        return lib().rc_update_ratio(self.h, t, i)

    def update_greens(self, t, i):
        lib().rc_update_greens(self.h, t, i)

    def flip(self, t, i):
        lib().rc_flip(self.h, t, i)

    # time-displaced path (DqmcWalker::sweep_for_dynamic_greens, MeasureHandler::dynamic_measure)
    @property
    def nk(self):
        side = int(round(self.N ** 0.5))
        return (side // 2 + 1) * (side // 2 + 2) // 2

    def sweep_dynamic(self):
        lib().rc_sweep_dynamic(self.h)

    def set_dynamic_formula(self, general):
        """False: the reference's SVD-stack formulas; True: the orthogonal-U (general UDT) form."""
        lib().rc_set_dynamic_formula(self.h, int(bool(general)))

    def dynamic_green(self, slot, spin, which):
        """which 0: G(t,0), 1: G(0,t) at slot t-1 of the last dynamic sweep."""
        g = np.zeros((self.N, self.N), order="F")
        lib().rc_get_dynamic_green(self.h, slot, spin, which, _p(g))
        return g

    def dynamic_measure(self):
        lib().rc_dynamic_measure(self.h)

    def dynamic_accum(self):
        """Sums of the dynamic observables: gk (nk x L), dos (L), dss (L), rho, sign, count."""
        nk, L = self.nk, self.L
        out = np.zeros(nk * L + 2 * L + 3)
        lib().rc_get_dynamic_measure(self.h, _p(out))
        o = nk * L
        return dict(gk=out[:o].reshape(L, nk).T.copy(), dos=out[o:o + L].copy(), dss=out[o + L:o + 2 * L].copy(),
                    rho=out[o + 2 * L], sign=out[o + 2 * L + 1], count=out[o + 2 * L + 2])

    def clear_dynamic(self):
        lib().rc_clear_dynamic(self.h)


def bin_stats(bins):
    """Observable::analyse (include/measure/observable.h:169-193): mean and sqrt(<x^2>-<x>^2)/sqrt(nb-1)."""
    b = np.asarray(bins, dtype=np.float64)
    nb = b.shape[0]
    mean = b.mean(axis=0)
    err = np.sqrt(np.maximum((b ** 2).mean(axis=0) - mean ** 2, 0.0)) / np.sqrt(nb - 1)
    return mean, err


def normalize_bin(sums, counts):
    """MeasureHandler::normalize_stats (src/measure/measure_handler.cpp:170-189)."""
    sums = np.asarray(sums, dtype=np.float64); counts = np.asarray(counts, dtype=np.float64)
    sign = sums[0] / counts[0]
    out = sums / (counts * sign)
    out[0] = abs(sign)
    return out
