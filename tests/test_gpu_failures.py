"""Numerical failures are reported, never returned as clean chains.

The reference stops the run when a decomposition fails (include/utils/linear_algebra.hpp:70-74).
The engine records what a chain met in its status bits (dqmc_chain_info.status): a non-finite
G or a singular pivot is fatal (the chain does no further sweeps), a wrap error above the
batch's limit is a warning.  Every case here feeds finite inputs whose arithmetic overflows or
cancels, so every kernel still runs to its end (no fault, no trap, no wait on data).
"""
import numpy as np
import pytest

import _oracle as O
import dqmc_amd as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return D.Context(0)


def test_overflowing_expk_fails_only_its_chain(ctx):
    """exp(-dtau K) with entries 1e200 (finite) overflows the slice products to inf: the chain's
    init reports DQMC_STATUS_NONFINITE, its sweeps are skipped, and the healthy chain in the same
    batch is the same chain it is alone."""
    print("[gpu-failures] overflowing expK", flush=True)
    side, L, beta = 4, 20, 2.0
    ps = [D.make_params(side, L, beta, pace=5, U=4.0, seed=s) for s in (3, 11)]
    b = D.Batch(ctx, ps)
    b.fields_random()
    huge = np.full((side * side, side * side), 1e200, order="F")
    b.set_expk(0, huge, huge)
    b.init()
    assert b.info(0).status & D.STATUS_NONFINITE, b.info(0).status
    assert b.info(1).status == 0
    b.sweep(2)
    b.synchronize()
    bad, good = b.info(0), b.info(1)
    assert bad.status & D.STATUS_FATAL
    assert bad.pos == 0          # no sweep ran on the failed chain
    assert good.status == 0 and good.pos == 0

    alone = D.Batch(ctx, [ps[1]])
    alone.fields_random()
    alone.init()
    alone.sweep(2)
    for spin in (1, -1):
        assert np.array_equal(alone.green(0, spin), b.green(1, spin))
    assert np.array_equal(alone.fields(0), b.fields(1))


def test_singular_recompute_is_a_numeric_error(ctx):
    """B_L = 1 and B_R = -1 make 1 + B_L B_R^T = 0: the recompute's LU meets zero pivots and
    dqmc_k_greens_recompute returns DQMC_ERR_NUMERIC instead of a matrix."""
    print("[gpu-failures] singular recompute", flush=True)
    n = 16
    eye = np.eye(n)
    one = np.ones(n)
    with pytest.raises(D.DqmcError, match=r"\(-5\)"):
        ctx.k_greens_recompute(left=(eye, one, eye), right=(-eye, one, eye))
    # a regular one still works after the failure (the status is per call)
    rng = np.random.default_rng(1)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    G = ctx.k_greens_recompute(left=(q, np.linspace(0.5, 2.0, n), eye), right=(eye, one, eye))
    B = q @ np.diag(np.linspace(0.5, 2.0, n))
    assert np.max(np.abs(G - np.linalg.inv(eye + B))) < 1e-12


def test_wrap_limit_is_a_warning(ctx):
    """A wrap-error limit below any achievable error flags DQMC_STATUS_WRAP_LIMIT on the first
    stabilization, and the chain goes on exactly as without the limit."""
    print("[gpu-failures] wrap limit", flush=True)
    p = D.make_params(6, 20, 2.0, pace=5, checkerboard=True, U=4.0, seed=5)
    runs = []
    for limit in (None, 1e-300):
        b = D.Batch(ctx, [p])
        if limit is not None:
            b.set_wrap_error_limit(limit)
        b.fields_random()
        b.init()
        b.sweep(3)
        runs.append(b)
    ref, flagged = runs
    assert ref.info(0).status == 0
    st = flagged.info(0).status
    assert st & D.STATUS_WRAP_LIMIT and not st & D.STATUS_FATAL, st
    assert flagged.info(0).pos == ref.info(0).pos
    for spin in (1, -1):
        assert np.array_equal(ref.green(0, spin), flagged.green(0, spin))


def test_recovers_oracle_after_flag(ctx):
    """The status bits do not perturb a healthy chain: with the default limit the 4x4 chain of
    the parity tests still equals the oracle (fields, accept bits, G at 1e-10)."""
    print("[gpu-failures] healthy chain under the default limit", flush=True)
    orc = O.Chain(4, 20, 2.0, pace=5, U=4.0)
    orc.seed(42); orc.fields_random(); orc.init()
    b = D.Batch(ctx, [D.make_params(4, 20, 2.0, pace=5, U=4.0, seed=42)])
    b.fields_random()
    b.set_expk(0, orc.expK(0), orc.expK(1))
    b.init()
    orc.sweep_forward(); orc.sweep_backward()
    b.sweep(2)
    assert b.info(0).status == 0
    assert np.array_equal(b.fields(0).astype(np.float64), orc.fields())
    assert max(np.max(np.abs(b.green(0, s) - orc.green(s))) for s in (1, -1)) < 1e-10
