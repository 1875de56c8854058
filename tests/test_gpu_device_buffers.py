"""The building blocks on caller-provided device buffers (dqmc_kd_* with a workspace), the
per-site plugin calls on a batch chain's resident state (dqmc_chain_*), and the reference's
own walker written against the C++ facade on top of them.

Bars (FP64):
  dqmc_kd_wrap / kd_udv_push / kd_greens_recompute   bit-identical to the host-buffer entry
                                   points dqmc_k_wrap / k_udv_push / k_greens_recompute (the
                                   same kernels on the same operands)
  update ratio                     within 1e-13 relative of the oracle's get_update_ratio
  single-flip update               max |dG| <= 1e-13 against the oracle
  the reference's Metropolis loop and walker (src/dqmc_walker.cpp:109-412) through the
  facade: accept bits, fields and sign identical to the oracle's, G within 1e-10 after each
  sweep (the sweeps end on a recomputed G)
The caller's device buffers are hipMalloc'ed through the engine's own HIP runtime (_hipmem).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import dqmc_amd as D
from _hipmem import DevBuf, synchronize

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WALKER = os.path.join(ROOT, "dqmc-framework_amd", "lib", "reference_walker")


@pytest.fixture(scope="module")
def ctx():
    return D.Context(0)



def chain(side, L, beta, model=0, cb=False, U=4.0, mu=0.0, seed=7):
    orc = O.Chain(side, L, beta, pace=5, model=model, checkerboard=cb, U=U, mu=mu)
    orc.seed(seed)
    orc.fields_random()
    p = D.make_params(side, L, beta, pace=5, model=model, checkerboard=cb, U=U, mu=mu, seed=seed)
    return orc, p, orc.fields().astype(np.int8)


CASES = [(4, 12, 2.0, 0, False, 4.0, -0.3), (6, 10, 1.5, 0, True, 6.0, 0.2), (4, 8, 1.0, 1, False, 3.0, -0.5)]


@pytest.mark.parametrize("case", CASES, ids=[f"s{c[0]}m{c[3]}cb{int(c[4])}" for c in CASES])
def test_kd_wrap_equals_host_entry(ctx, case):
    side, L, beta, model, cb, U, mu = case
    print(f"[gpu-kd] wrap {case}", flush=True)
    orc, p, f = chain(side, L, beta, model, cb, U, mu)
    N = side * side
    ws = D.Workspace(ctx, p)
    df = DevBuf(f.reshape(-1))   # fields are row-major (t, i) at t*N+i
    rng = np.random.default_rng(3)
    for op in range(7):
        for t in (0, 1, L):
            M = rng.standard_normal((N, N))
            want = ctx.k_wrap(p, f, t, +1 if op % 2 == 0 else -1, op, M)
            dM = DevBuf(M)
            ws.wrap(df.ptr, t, +1 if op % 2 == 0 else -1, op, dM.ptr)
            synchronize()
            assert np.array_equal(dM.get((N, N)), want), (op, t)
    if not cb:
        # the workspace's exp(-dtau K) is the closed form the batches use (oracle Pade within 1e-14)
        eK, iK = ws.expk()
        assert np.max(np.abs(eK - orc.expK(0))) < 1e-13
        assert np.max(np.abs(iK - orc.expK(1))) < 1e-13


@pytest.mark.parametrize("model", [0, 1])
def test_kd_ratio_rank1_flip_match_oracle(ctx, model):
    print(f"[gpu-kd] ratio / rank1 / flip model {model}", flush=True)
    side, L = 4, 8
    orc, p, f = chain(side, L, 1.0, model=model, U=4.0, mu=-0.2)
    orc.init()
    N = side * side
    ws = D.Workspace(ctx, p)
    df = DevBuf(f.reshape(-1))   # fields are row-major (t, i) at t*N+i
    gu, gd = DevBuf(orc.green(1)), DevBuf(orc.green(-1))
    r = DevBuf(np.zeros(1))
    for (t, i) in [(0, 0), (3, 5), (7, 15), (2, 9)]:
        ws.update_ratio(df.ptr, t, i, gu.ptr, gd.ptr, r.ptr)
        synchronize()
        want = orc.update_ratio(t, i)
        assert abs(r.get()[0] - want) <= 1e-13 * abs(want), (t, i, r.get()[0], want)
        ws.rank1(df.ptr, t, i, gu.ptr, gd.ptr if model == 0 else None)
        ws.flip_field(df.ptr, t, i)
        synchronize()
        orc.update_greens(t, i)
        orc.flip(t, i)
        assert np.max(np.abs(gu.get((N, N)) - orc.green(1))) < 1e-13
        if model == 0:
            assert np.max(np.abs(gd.get((N, N)) - orc.green(-1))) < 1e-13
        assert np.array_equal(df.get().reshape(L, N).astype(np.float64), orc.fields())


def test_kd_push_and_recompute_equal_host_entries(ctx):
    print("[gpu-kd] push / recompute", flush=True)
    N = 36
    p = D.make_params(6, 10, 1.0, pace=5)
    ws = D.Workspace(ctx, p)
    rng = np.random.default_rng(11)
    A = rng.standard_normal((N, N))
    B = rng.standard_normal((N, N))
    U0, D0, V0 = ctx.k_udv_push(A)
    U1, D1, V1 = ctx.k_udv_push(B, top=(U0, D0, V0))
    dA, dB = DevBuf(A), DevBuf(B)
    out = [DevBuf(np.zeros((N, N))), DevBuf(np.zeros(N)), DevBuf(np.zeros((N, N)))]
    st = DevBuf(np.full(1, -1, np.int32))
    ws.udv_push(dA.ptr, None, *[o.ptr for o in out], st.ptr)
    synchronize()
    assert st.get()[0] == 0
    assert np.array_equal(out[0].get((N, N)), U0) and np.array_equal(out[1].get(), D0)
    assert np.array_equal(out[2].get((N, N)), V0)
    top = [DevBuf(out[0].get((N, N))), DevBuf(out[1].get()), DevBuf(out[2].get((N, N)))]
    ws.udv_push(dB.ptr, tuple(t.ptr for t in top), *[o.ptr for o in out], st.ptr)
    synchronize()
    assert np.array_equal(out[0].get((N, N)), U1) and np.array_equal(out[1].get(), D1)
    assert np.array_equal(out[2].get((N, N)), V1)
    # recompute from (U1, D1, V1) left and (U0, D0, V0) right, and from one stack alone
    G = DevBuf(np.zeros((N, N)))
    left = tuple(DevBuf(x) for x in (U1, D1, V1))
    right = tuple(DevBuf(x) for x in (U0, D0, V0))
    for lt, rt, hl, hr in [(left, right, (U1, D1, V1), (U0, D0, V0)), (left, None, (U1, D1, V1), None),
                           (None, right, None, (U0, D0, V0))]:
        ws.greens_recompute(None if lt is None else tuple(t.ptr for t in lt),
                            None if rt is None else tuple(t.ptr for t in rt), G.ptr, st.ptr)
        synchronize()
        assert st.get()[0] == 0
        assert np.array_equal(G.get((N, N)), ctx.k_greens_recompute(left=hl, right=hr))
    # 1 + B_L B_R^T = 0: the status comes back nonzero, on the device
    eye, one = DevBuf(np.eye(N)), DevBuf(np.ones(N))
    neg = DevBuf(-np.eye(N))
    ws.greens_recompute((eye.ptr, one.ptr, eye.ptr), (neg.ptr, one.ptr, eye.ptr),
                        G.ptr, st.ptr)
    synchronize()
    assert st.get()[0] & D.STATUS_FATAL


@pytest.mark.parametrize("model", [0, 1])
def test_chain_plugin_calls_run_the_reference_metropolis_loop(ctx, model):
    """DqmcWalker::metropolis_update (src/dqmc_walker.cpp:173-197) written against the chain's
    per-site calls, with the wraps on the resident G, reproduces the oracle's decisions and G
    when both take their draws from the same minstd_rand0 stream."""
    print(f"[gpu-kd] reference Metropolis loop, model {model}", flush=True)
    side, L, beta, seed = 4, 10, 1.5, 21
    orc = O.Chain(side, L, beta, pace=5, model=model, U=4.0, mu=-0.1)
    orc.seed(seed); orc.fields_random(); orc.init()
    b = D.Batch(ctx, [D.make_params(side, L, beta, pace=5, model=model, U=4.0, mu=-0.1, seed=seed)])
    b.fields_random()
    b.init()
    N = side * side
    draws = iter(O.canonical_stream(12345, 3 * L * N))
    G = {s: orc.green(s) for s in (1, -1)}
    for s in (1, -1):
        b.set_green(0, s, G[s])
    for t in range(1, 4):
        eff_t = t - 1
        # wrap_from_0_to_beta(t - 1): G <- B(t) G B(t)^-1 on both spins (the attractive chain
        # keeps one matrix on the device, G_dn = G_up)
        for s in ((1, -1) if model == 0 else (1,)):
            b.mult_green(0, t, s, 0)
            b.mult_green(0, t, s, 3)
        for s in (1, -1):
            orc.set_green(s, orc.mult(3, t, s, orc.mult(0, t, s, orc.green(s))))
        for i in range(N):
            r = b.update_ratio(0, eff_t, i)
            ro = orc.update_ratio(eff_t, i)
            assert abs(r - ro) <= 1e-11 * max(1.0, abs(ro)), (t, i, r, ro)
            if next(draws) < min(1.0, abs(r)):
                b.update_green(0, eff_t, i)
                b.flip_field(0, eff_t, i)
                orc.update_greens(eff_t, i)
                orc.flip(eff_t, i)
        for s in (1, -1):
            assert np.max(np.abs(b.green(0, s) - orc.green(s))) < 1e-11
        assert np.array_equal(b.fields(0).astype(np.float64), orc.fields())


@pytest.mark.parametrize("cfg", [
    # side L beta pace model cb U mu seed
    (4, 20, 2.0, 5, 0, 0, 4.0, 0.0, 42),
    (4, 16, 2.0, 5, 1, 1, 4.0, -0.3, 7),
    (6, 12, 1.5, 4, 0, 1, 5.0, 0.2, 3),
])
def test_reference_walker_through_the_facade_matches_oracle(cfg):
    print(f"[gpu-kd] facade walker {cfg}", flush=True)
    side, L, beta, pace, model, cb, U, mu, seed = cfg
    out = subprocess.run([WALKER] + [str(x) for x in cfg] + ["4"], check=True, capture_output=True, text=True,
                         timeout=600).stdout.strip().splitlines()
    orc = O.Chain(side, L, beta, pace=pace, model=model, checkerboard=bool(cb), U=U, mu=mu)
    orc.seed(seed); orc.fields_random(); orc.init()
    N = side * side
    for k, line in enumerate(out):
        got = json.loads(line)
        if k % 2 == 0:
            orc.sweep_forward()
        else:
            orc.sweep_backward()
        assert np.array_equal(np.array(got["accepts"], np.uint8).reshape(L, N), orc.accepts()), k
        assert np.array_equal(np.array(got["fields"]).reshape(L, N), orc.fields()), k
        assert got["sign"] == orc.sign, k
        for name, s in (("green_up", 1), ("green_dn", -1)):
            g = np.array(got[name]).reshape(N, N).T
            assert np.max(np.abs(g - orc.green(s))) < 1e-10, (k, name)
        assert got["wrap_error"] < max(1e-9, 10.0 * orc.wrap_error)
    assert len(out) == 4
