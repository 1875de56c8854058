"""The precision workload live: 64 chains of 16x16 (N = 256) repulsive Hubbard, U = 6,
mu in {0, -0.2, -0.4, -0.6} (16 chains each), beta = 10, L = 200, stabilization pace 8, dense
exp(-dtau K), seeds 0..63, delay depth 32, in one launch.

FP64 leg: two chains of the batch (mu = 0 and mu = -0.2) against the oracle over a forward and
a backward sweep: fields, accept bits and sign identical; the recomputed G that ends each sweep
within max(1e-10, 10x the spread between the oracle's own SVD and pre-pivoted QR stacks on the
same chain) -- at beta 10, U 6 the conditioning of 1 + B...B makes two FP64 recomputes of the
same G differ by a few 1e-10 whatever the implementation.

Mixed leg (FP32 matrix-core slice products and wraps, FP32 accumulation; FP64 stacks,
recompute, delayed updates and ratios): the same 64 chains live.  At every stabilization of two
chains' upward sweeps the recomputed G is compared with an FP64 recompute of the configuration
the chain had at that moment (rows up to the slice from after the sweep, the rest from before),
and each chain's wrap error |G_wrapped - G_fresh| is read.  Measured on one MI355X: recomputed G
within 5e-5 (median) and 4.6e-3 (max) of FP64, wrap error median 0.11 and max 51, so the mixed
path does not hold 1e-6 on live chains; the bars below are that measured envelope, and chains
past the wrap-error limit must carry DQMC_STATUS_WRAP_LIMIT.  The same measurement in FP64 gives
exactly zero (the check of the comparison itself).
"""
import threading

import numpy as np
import pytest

import _oracle as O
import dqmc_amd as D
import precision as P

pytestmark = pytest.mark.gpu

SIDE, L, BETA, PACE, U = 16, 200, 10.0, 8, 6.0
N = SIDE * SIDE
MUS = [0.0, -0.2, -0.4, -0.6]
NCH = 64


@pytest.fixture(scope="module")
def ctx():
    return D.Context(0)


def mu_of(c):
    return MUS[c * len(MUS) // NCH]


def params(c, precision):
    return D.make_params(SIDE, L, BETA, pace=PACE, U=U, mu=mu_of(c), delay=32, seed=c, precision=precision)


def test_fp64_batch_matches_oracle(ctx):
    print("[gpu-precision-live] FP64 64 x 16x16 batch vs oracle (chains 0, 16)", flush=True)
    b = D.Batch(ctx, [params(c, D.PREC_FP64) for c in range(NCH)], record=True)
    b.fields_random()
    b.init()
    chains = [0, 16]
    orcs = {}
    for c in chains:
        for fact in (0, 2):
            o = O.Chain(SIDE, L, BETA, pace=PACE, U=U, mu=mu_of(c), factorization=fact)
            o.seed(c); o.fields_random()
            o.set_expK(b.expk(c)[0], b.expk(c)[1])   # the closed-form exp(-dtau K) the batch uses
            orcs[c, fact] = o
    # the oracle chains run in parallel threads (ctypes releases the GIL)
    th = [threading.Thread(target=o.init) for o in orcs.values()]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for k, up in enumerate((True, False)):
        b.sweep(1, record=True)
        th = [threading.Thread(target=(o.sweep_forward if up else o.sweep_backward)) for o in orcs.values()]
        for t in th:
            t.start()
        b.synchronize()
        for t in th:
            t.join()
        for c in chains:
            o, q = orcs[c, 0], orcs[c, 2]
            assert np.array_equal(b.fields(c).astype(np.float64), o.fields()), (c, k)
            assert np.array_equal(b.accepts(c), o.accepts()), (c, k)
            assert np.array_equal(q.accepts(), o.accepts()), (c, k)
            spread = max(np.max(np.abs(q.green(spin) - o.green(spin))) for spin in (1, -1))
            dg = max(np.max(np.abs(b.green(c, spin) - o.green(spin))) for spin in (1, -1))
            assert dg < max(1e-10, 10.0 * spread), (c, k, dg, spread)
            assert b.info(c).sign == o.sign
            print(f"[gpu-precision-live] chain {c} sweep {k}: fields/accepts/sign identical, max|dG| {dg:.2e} "
                  f"(oracle SVD vs QR stacks {spread:.2e}), wrap error {b.info(c).wrap_error:.2e} (oracle "
                  f"{o.wrap_error:.2e})", flush=True)
    assert all(b.info(c).status & D.STATUS_FATAL == 0 for c in range(NCH))


def stabilization_errors(ctx, precision, chains=(0, 16), sweeps=4):
    """max|dG| of the recomputed G at every stabilization of the upward sweeps against an FP64
    recompute of the configuration the chain had at that moment (precision.py)."""
    return P.stabilization_errors(ctx, [params(c, precision) for c in chains], sweeps=sweeps)


def test_fp64_stabilization_check_is_exact(ctx):
    print("[gpu-precision-live] FP64 recompute check", flush=True)
    e = stabilization_errors(ctx, D.PREC_FP64, chains=(0,), sweeps=2)
    assert e.size == 2 * 25 and np.all(e == 0.0), e.max()


def test_mixed_live_error_is_measured_and_flagged(ctx):
    print("[gpu-precision-live] mixed 64 x 16x16 live", flush=True)
    b = D.Batch(ctx, [params(c, D.PREC_MIXED) for c in range(NCH)])
    b.fields_random()
    b.init()
    b.sweep(4)
    b.synchronize()
    we = np.array([b.info(c).wrap_error for c in range(NCH)])
    st = np.array([b.info(c).status for c in range(NCH)])
    dg = stabilization_errors(ctx, D.PREC_MIXED)
    print(f"[gpu-precision-live] mixed: recomputed-G error vs FP64 median {np.median(dg):.2e} max {dg.max():.2e}; "
          f"wrap error median {np.median(we):.2e} max {we.max():.2e}; {int(np.sum(st & D.STATUS_WRAP_LIMIT > 0))} of "
          f"{NCH} chains flagged", flush=True)
    assert np.all((st & D.STATUS_FATAL) == 0)
    # every chain past the limit is flagged, none below it is
    over = we > D.DEFAULT_WRAP_LIMIT
    assert np.array_equal(over, (st & D.STATUS_WRAP_LIMIT) > 0)
    # the measured envelope: FP32 rounding is visible (not FP64-accurate) and bounded at the recompute
    assert 1e-7 < np.median(dg) < 1e-3 and dg.max() < 2e-2, (np.median(dg), dg.max())
    assert np.median(we) > 1e-6   # the 1e-6 target is not met on live chains
