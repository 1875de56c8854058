"""The no-allocation recompute (dqmc_k_greens_recompute_noalloc: caller device buffers, static
scratch) gives the result of dqmc_k_greens_recompute, the host-buffer entry point it is a
variant of, on two calls in a row, the second with a smaller lattice (the static scratch keeps
the first call's contents).  Bar (FP64): max |dG| <= 1e-13 x max |G| (the same kernels on the
same operands; bit-identical in practice), status 0.  N above the static scratch's limit is
DQMC_ERR_ARG with a message that names the limit."""
import numpy as np
import pytest

import dqmc_amd as D
from _hipmem import DevBuf, synchronize

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return D.Context(0)


def tops(ctx, N, seed):
    """Two stack tops of moderately conditioned slice products (U orthogonal, D graded)."""
    rng = np.random.default_rng(seed)
    mats = [np.eye(N) + 0.3 * rng.standard_normal((N, N)) for _ in range(4)]
    left = ctx.k_udv_push(mats[0])
    left = ctx.k_udv_push(mats[1], top=left)
    right = ctx.k_udv_push(mats[2])
    right = ctx.k_udv_push(mats[3], top=right)
    return left, right


def run_noalloc(ctx, N, left, right):
    dl = None if left is None else tuple(DevBuf(x) for x in left)
    dr = None if right is None else tuple(DevBuf(x) for x in right)
    G = DevBuf(np.full((N, N), np.nan))
    st = DevBuf(np.full(1, -1, np.int32))
    ctx.k_greens_recompute_noalloc(N, None if dl is None else tuple(b.ptr for b in dl),
                                   None if dr is None else tuple(b.ptr for b in dr), G.ptr, st.ptr)
    synchronize()
    return G.get((N, N)), int(st.get()[0])


@pytest.mark.parametrize("sides", [(12, 6), (16, 4)])
def test_noalloc_recompute_equals_the_existing_entry_point(ctx, sides):
    print(f"[gpu-noalloc] sides {sides}", flush=True)
    for k, side in enumerate(sides):   # two calls in a row, the second smaller
        N = side * side
        left, right = tops(ctx, N, seed=100 + side)
        for lt, rt in [(left, right), (left, None), (None, right)]:
            want = ctx.k_greens_recompute(left=lt, right=rt)
            got, st = run_noalloc(ctx, N, lt, rt)
            assert st == 0, (side, st)
            assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want)), (k, side)


def test_noalloc_rejects_lattices_beyond_its_scratch(ctx):
    print("[gpu-noalloc] size limit", flush=True)
    N = 17 * 17
    assert N > 256
    eye, one = DevBuf(np.eye(N)), DevBuf(np.ones(N))
    G = DevBuf(np.zeros((N, N)))
    with pytest.raises(D.DqmcError, match="exceeds its static scratch"):
        ctx.k_greens_recompute_noalloc(N, (eye.ptr, one.ptr, eye.ptr), None, G.ptr)
    # a call after the refused one still works
    left, right = tops(ctx, 64, seed=5)
    got, st = run_noalloc(ctx, 64, left, right)
    want = ctx.k_greens_recompute(left=left, right=right)
    assert st == 0 and np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))


def test_noalloc_reports_a_singular_recompute(ctx):
    print("[gpu-noalloc] singular", flush=True)
    N = 36
    e, one = np.eye(N), np.ones(N)
    _, st = run_noalloc(ctx, N, (e, one, e), (-e, one, e))
    assert st & D.STATUS_FATAL
