"""Mixed precision (FP32 matrix-core slice products and wraps with FP32
accumulation; FP64 UDV stacks, recompute, delayed updates and ratios) against
the FP64 engine on the same seeds and fields.  On a frozen configuration both
runs visit identical fields, so the difference is the arithmetic alone.

FP32 accumulation rounds every wrap at ~2^-24 relative, and the error grows
over the wraps between stabilizations, so G carries 1e-5..1e-3 absolute
error (not the 1e-6 of an FP64-accurate product); the equal-time observables
agree to ~1e-6 relative.  The bars below state that measured envelope.  On a
live chain the Metropolis updates amplify the error further between
stabilizations (at beta 10, U 6, pace 8 the wrap error reaches O(0.1)), which
is why this comparison freezes the fields."""
import numpy as np
import pytest

import dqmc_amd as D

pytestmark = pytest.mark.gpu

TOL_G = {8: 2e-4, 16: 5e-3}    # measured 4.3e-5 (8x8, beta 4) and 7.4e-4 (16x16, beta 10, U 6)
TOL_OBS = 1e-4


@pytest.fixture(scope="module")
def ctx():
    return D.Context(0)


@pytest.mark.parametrize("side,L,beta,U,mu,pace", [(8, 80, 4.0, 4.0, 0.0, 10), (16, 200, 10.0, 6.0, -0.2, 8)])
def test_mixed_precision_green_error_on_frozen_fields(ctx, side, L, beta, U, mu, pace):
    print(f"[gpu-precision] {side}x{side} L={L} beta={beta} U={U}", flush=True)
    p64 = D.make_params(side, L, beta, pace=pace, U=U, mu=mu, seed=11)
    pmx = D.make_params(side, L, beta, pace=pace, U=U, mu=mu, seed=11, precision=D.PREC_MIXED)
    b = D.Batch(ctx, [p64, pmx], record=True)
    b.fields_random()
    b.init()
    # thermalize the FP64 chain, then copy its fields to the mixed chain
    # (a frozen sweep later visits the same configuration in both)
    warm = D.Batch(ctx, [p64])
    warm.fields_random(); warm.init(); warm.sweep(20)
    f = warm.fields(0)
    b.set_fields(0, f); b.set_fields(1, f)
    b.init()
    b.sweep(2, measure=True, record=True, frozen=True)
    worst = 0.0
    for slot in range(L):
        for spin in (1, -1):
            worst = max(worst, np.max(np.abs(b.slice_green(0, slot, spin) - b.slice_green(1, slot, spin))))
    s64, c64 = b.measure_accum(0)
    smx, cmx = b.measure_accum(1)
    o64 = D.normalize_bin(s64, c64)
    omx = D.normalize_bin(smx, cmx)
    rel = np.abs(omx - o64) / np.maximum(np.abs(o64), 1e-12)
    print(f"[gpu-precision] max|dG| = {worst:.3e}; observable rel. errors {np.array2string(rel, precision=2)}",
          flush=True)
    assert worst < TOL_G[side], worst
    assert np.all(rel[1:] < TOL_OBS), rel
