"""Workgroup GEMMs of the engine (vector FP64, MFMA FP64, MFMA FP32) against
an FP64 numpy product of the same operands, through dqmc_k_gemm."""
import numpy as np
import pytest

import dqmc_amd as D

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16), (64, 64, 64), (100, 36, 144), (144, 144, 144), (7, 65, 3), (256, 256, 256), (400, 400, 400),
          # outputs of at most 16 rows take 16 x 128 tiles: several column tiles, partial slabs
          (16, 144, 144), (16, 300, 37), (9, 130, 20), (1, 257, 5)]


@pytest.fixture(scope="module")
def ctx():
    return D.Context(0)


@pytest.mark.parametrize("variant", [0, 1, 2, 5, 6])
@pytest.mark.parametrize("shape", SHAPES, ids=[f"{m}x{n}x{k}" for m, n, k in SHAPES])
@pytest.mark.parametrize("tA,tB", [(False, False), (True, False), (False, True)])
def test_gemm_variants(ctx, variant, shape, tA, tB):
    print(f"[gpu-gemm] variant {variant} shape {shape} tA {tA} tB {tB}", flush=True)
    M, Nc, K = shape
    rng = np.random.default_rng(M * 7 + Nc * 3 + K + variant)
    A = rng.standard_normal((K, M) if tA else (M, K))
    B = rng.standard_normal((Nc, K) if tB else (K, Nc))
    C0 = rng.standard_normal((M, Nc))
    rs, ms, cs = rng.uniform(0.5, 2, M), rng.uniform(0.5, 2, K), rng.uniform(0.5, 2, Nc)
    opA = A.T if tA else A
    opB = B.T if tB else B
    ref = rs[:, None] * (opA @ np.diag(ms) @ opB) * cs[None, :] + 0.5 * C0
    got = ctx.gemm(variant, A, B, C0, transA=tA, transB=tB, beta=0.5, rs=rs, ms=ms, cs=cs)
    scale = np.abs(rs[:, None]) * (np.abs(opA) @ np.diag(ms) @ np.abs(opB)) * cs[None, :] + 0.5 * np.abs(C0)
    err = np.max(np.abs(got - ref) / scale)
    tol = 1e-14 * K if variant < 2 else 2e-6 * np.sqrt(K)
    assert err < tol, err


LDS_SHAPES = [(16, 144, 16), (32, 144, 32), (144, 36, 32), (7, 65, 3), (64, 100, 64), (16, 400, 16), (9, 300, 20)]


@pytest.mark.parametrize("variant", [3, 4])
@pytest.mark.parametrize("shape", LDS_SHAPES, ids=[f"{m}x{n}x{k}" for m, n, k in LDS_SHAPES])
@pytest.mark.parametrize("tA,tB", [(False, False), (True, False), (False, True)])
def test_gemm_lds_operands(ctx, variant, shape, tA, tB):
    """The MFMA GEMM with A (3) or A and B (4) in LDS, as the LU, QR and
    block-update callers use it."""
    M, Nc, K = shape
    if variant == 4 and M * K + K * Nc > 13824:
        pytest.skip("operands exceed the LDS panel")
    print(f"[gpu-gemm] variant {variant} shape {shape} tA {tA} tB {tB}", flush=True)
    rng = np.random.default_rng(M * 5 + Nc * 11 + K + variant)
    A = rng.standard_normal((K, M) if tA else (M, K))
    B = rng.standard_normal((Nc, K) if tB else (K, Nc))
    C0 = rng.standard_normal((M, Nc))
    ms = rng.uniform(0.5, 2, K)
    opA = A.T if tA else A
    opB = B.T if tB else B
    ref = opA @ np.diag(ms) @ opB - 1.5 * C0
    got = ctx.gemm(variant, A, B, C0, transA=tA, transB=tB, beta=-1.5, ms=ms)
    scale = np.abs(opA) @ np.diag(ms) @ np.abs(opB) + 1.5 * np.abs(C0)
    assert np.max(np.abs(got - ref) / scale) < 1e-14 * K


WIDE_SHAPES = [(144, 144, 144), (256, 256, 256), (100, 121, 144), (160, 160, 37), (161, 150, 200), (400, 400, 400),
               (196, 196, 196), (17, 300, 5), (129, 128, 16)]


@pytest.mark.parametrize("variant", [5, 6])
@pytest.mark.parametrize("shape", WIDE_SHAPES, ids=[f"{m}x{n}x{k}" for m, n, k in WIDE_SHAPES])
@pytest.mark.parametrize("tA,tB", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("beta", [0.0, -0.75])
def test_wide_gemm_equals_tiled(ctx, variant, shape, tA, tB, beta):
    """The wide-tile GEMM sums every output element in the same order as the 64 x 64 tiles and
    scales it the same way, so the engine may use either: the results are equal bit for bit
    (with and without scalings, transposes and a C term)."""
    print(f"[gpu-gemm] wide variant {variant} shape {shape} tA {tA} tB {tB} beta {beta}", flush=True)
    M, Nc, K = shape
    rng = np.random.default_rng(M * 13 + Nc * 5 + K + variant)
    A = rng.standard_normal((K, M) if tA else (M, K))
    B = rng.standard_normal((Nc, K) if tB else (K, Nc))
    C0 = rng.standard_normal((M, Nc))
    for scal in (False, True):
        kw = dict(rs=rng.uniform(0.5, 2, M), ms=rng.uniform(0.5, 2, K), cs=rng.uniform(0.5, 2, Nc)) if scal else {}
        ref = ctx.gemm(1, A, B, C0, transA=tA, transB=tB, beta=beta, **kw)
        got = ctx.gemm(variant, A, B, C0, transA=tA, transB=tB, beta=beta, **kw)
        assert np.array_equal(got, ref), np.max(np.abs(got - ref))


def test_gemm_plain(ctx):
    rng = np.random.default_rng(0)
    A = rng.standard_normal((80, 80)); B = rng.standard_normal((80, 80))
    for v in (0, 1):
        assert np.allclose(ctx.gemm(v, A, B), A @ B, atol=1e-12)
