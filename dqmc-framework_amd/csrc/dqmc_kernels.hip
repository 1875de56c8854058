// HIP kernels of the DQMC engine for gfx950 (MI355X).
//
// Execution model: one 256-thread workgroup owns one Markov chain for a whole
// launch.  A launch sweeps every chain of a batch (chains may differ in N, L
// and model) through n single-direction sweeps: slice wraps, delayed-update
// Metropolis, slice-product accumulation, UDV-stack stabilization and the
// equal-time measurements all run inside the workgroup, with the chain's
// state (G, stacks, fields, random state) resident in HBM between launches.
// Per-chain small vectors (diagonals, scalings, field row, pivots) live in LDS.
//
// Reference semantics followed (file:line of dqmc-framework):
//   sweeps        src/dqmc_walker.cpp:246-412
//   Metropolis    src/dqmc_walker.cpp:173-197, src/model/repulsive_hubbard.cpp:112-155,
//                 src/model/attractive_hubbard.cpp:112-157
//   B products    src/model/repulsive_hubbard.cpp:158-243, src/model/model_base.cpp:7-35,
//                 src/checkerboard/square.cpp:66-226
//   stacks        src/svd_stack.cpp:34-87 (UDV by pivoted QR instead of SVD; the
//                 Green's function is independent of the factorization)
//   recompute     include/utils/numerical_stable.hpp:132-152, :186-239
//   measurements  src/measure/measure_methods.cpp:22-255
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chain_desc.h"

#define NT DQMC_NT
#define NMAX DQMC_NMAX

namespace dqmc {

// ---------------------------------------------------------------------------
// LDS layout of one workgroup
// ---------------------------------------------------------------------------
#define NB 16                    // panel width of the blocked LU and QR
#define GLD 80                   // row stride of the GEMM slabs in LDS (64 + 16)
#define PANEL_DOUBLES DQMC_PANEL_DOUBLES   // an LU/QR panel (up to NMAX x NB) plus its NB x NB blocks
// panel-region layout of the LU / QR: P (NMAX x NB), two NB x NB blocks, then the
// reduction scratch of the register-resident panel factorizations
#define PANEL_BLK (NMAX * NB)
#define PANEL_RED (NMAX * NB + 2 * NB * NB)
#define RED_SLOTS (2 * 4 * NB)      // [parity][wave][slot]
#define RED_ROWJ (RED_SLOTS)        // [parity][NB]: row j of the panel
#define RED_CAND (RED_ROWJ + 2 * NB)   // [parity][wave][NB]: pivot candidate rows (LU)
#define RED_CVAL (RED_CAND + 2 * 4 * NB)   // [parity][wave]: candidate |value|, then row index
#define RED_GRAM (RED_CVAL + 2 * 2 * 4)    // [wave][128]: Gram partials (QR)
#define RED_END (RED_GRAM + 4 * 128 + 128)
static_assert(PANEL_RED + RED_END <= DQMC_PANEL_DOUBLES, "panel scratch exceeds the LDS panel");

struct Lds {
    double As[16][GLD];        // GEMM A slab (k-major, 16 deep; rows padded so that the four
    double Bs[16][GLD];        // lane groups of an MFMA operand read fall on disjoint LDS banks)
    double du[NMAX], dd[NMAX]; // diagonals of the current G (delayed updates, measurements)
    double scp[NMAX], scm[NMAX];   // exp(+alpha s_i), exp(-alpha s_i) of the current field row
    double v0[NMAX], v1[NMAX], v2[NMAX], v3[NMAX];  // stabilization / QR scratch vectors
    int perm[NMAX];
    int frow[NMAX];            // field row of the current slice
    int swp_row[2 * NB], swp_src[2 * NB];   // the row interchanges of one LU panel, composed
    double panel[PANEL_DOUBLES];  // LU / QR panel, or the delayed-update panels X, Y
    double red[16];
    double gam[2];
    int ivar[4];
    int status;                   // DQMC_STATUS_* bits met by the running launch (thread 0 / panel lanes set them)
    unsigned long long prof[2];   // sub-phase ticks (walk, panels) of the running launch
    int prof_on;
    // spin range of this workgroup: [sp_lo, sp_hi) = [0, nspin), or one spin of a pair launch
    int sp_lo, sp_hi;
    int pair, me;                 // pair launch; this workgroup's spin in it
    unsigned xseq;                // exchanges of this launch so far (pair launch)
    int xbroken;                  // the partner stopped answering: no further waits
    int join;                     // pair launch: 1 paired, 0 this workgroup runs the chain alone, -1 it leaves
};

// one LDS object per workgroup, at namespace scope so that separately
// compiled (noinline) phases address it with ds_* instructions
__shared__ Lds g_lds;

// sub-phase timer: thread 0 adds the 100 MHz wall-clock ticks between start()
// and stop() to g_lds.prof[slot] when the launch profiles
struct SubClock {
    unsigned long long t0 = 0;
    __device__ void start() { if (g_lds.prof_on && threadIdx.x == 0) t0 = wall_clock64(); }
    __device__ void stop(int slot) { if (g_lds.prof_on && threadIdx.x == 0) g_lds.prof[slot] += wall_clock64() - t0; }
};

__device__ __forceinline__ int tid() { return threadIdx.x; }
__device__ __forceinline__ int lane() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave() { return threadIdx.x >> 6; }

// Cross-lane moves without the LDS crossbar: DPP within 16-lane rows and the
// gfx950 row/half swaps across them.  dpp_d<CTRL>: the value of a partner lane
// (0xB1: lane ^ 1, 0x4E: lane ^ 2, 0x141: lane 7 - i within 8, 0x140: lane
// 15 - i within 16, 0x128: lane ^ 8).  swap16 / swap32 return, for a pair of
// values (x, y), x with its odd rows (upper half) exchanged for the even rows
// (lower half) of y, and y with the exchanged rows: with x = y every lane sees
// both rows of its row pair (both halves of the wavefront).
__device__ __forceinline__ double pack_d(unsigned lo, unsigned hi) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int CTRL> __device__ __forceinline__ double dpp_d(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
    return pack_d(lo, hi);
}
template <int CTRL> __device__ __forceinline__ int dpp_i(int v) { return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, false); }
__device__ __forceinline__ void swap16_d(double& x, double& y) {
    const unsigned long long bx = (unsigned long long)__double_as_longlong(x), by = (unsigned long long)__double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)bx, (unsigned)by, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(bx >> 32), (unsigned)(by >> 32), false, false);
    x = pack_d(lo[0], hi[0]);
    y = pack_d(lo[1], hi[1]);
}
__device__ __forceinline__ void swap32_d(double& x, double& y) {
    const unsigned long long bx = (unsigned long long)__double_as_longlong(x), by = (unsigned long long)__double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)bx, (unsigned)by, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(bx >> 32), (unsigned)(by >> 32), false, false);
    x = pack_d(lo[0], hi[0]);
    y = pack_d(lo[1], hi[1]);
}
// the value of lane ^ 16 / lane ^ 32
__device__ __forceinline__ double xor16_d(double v) { double a = v, b = v; swap16_d(a, b); return (threadIdx.x & 16) ? a : b; }
__device__ __forceinline__ double xor32_d(double v) { double a = v, b = v; swap32_d(a, b); return (threadIdx.x & 32) ? a : b; }
__device__ __forceinline__ int xor16_i(int v) {
    const auto r = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
    return (int)((threadIdx.x & 16) ? r[0] : r[1]);
}
__device__ __forceinline__ int xor32_i(int v) {
    const auto r = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    return (int)((threadIdx.x & 32) ? r[0] : r[1]);
}

__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    double a = v, b = v;
    swap16_d(a, b);
    v = a + b;
    a = v; b = v;
    swap32_d(a, b);
    return a + b;
}
__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_d<0xB1>(v));
    v = fmax(v, dpp_d<0x4E>(v));
    v = fmax(v, dpp_d<0x141>(v));
    v = fmax(v, dpp_d<0x140>(v));
    double a = v, b = v;
    swap16_d(a, b);
    v = fmax(a, b);
    a = v; b = v;
    swap32_d(a, b);
    return fmax(a, b);
}
// maximum that propagates NaN from either operand (fmax drops it), so that a
// non-finite Green's function cannot report a small error
__device__ __forceinline__ double nan_max(double a, double b) { return (a != a || a > b) ? a : b; }
__device__ __forceinline__ double wave_max_nan(double v) {
    v = nan_max(v, dpp_d<0xB1>(v));
    v = nan_max(v, dpp_d<0x4E>(v));
    v = nan_max(v, dpp_d<0x141>(v));
    v = nan_max(v, dpp_d<0x140>(v));
    double a = v, b = v;
    swap16_d(a, b);
    v = nan_max(a, b);
    a = v; b = v;
    swap32_d(a, b);
    return nan_max(a, b);
}
// a pivot the elimination cannot divide by safely: zero, subnormal, inf or NaN
__device__ __forceinline__ bool bad_pivot(double piv) {
    const double a = fabs(piv);
    return !(a >= 2.2250738585072014e-308 && a <= 1.7976931348623157e308);
}
// argmax with smallest index on ties (the result does not depend on the pairing order)
__device__ __forceinline__ void argmax_step(double& v, int& idx, double ov, int oi) {
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
}
__device__ __forceinline__ void wave_argmax(double& v, int& idx) {
    argmax_step(v, idx, dpp_d<0xB1>(v), dpp_i<0xB1>(idx));
    argmax_step(v, idx, dpp_d<0x4E>(v), dpp_i<0x4E>(idx));
    argmax_step(v, idx, dpp_d<0x141>(v), dpp_i<0x141>(idx));
    argmax_step(v, idx, dpp_d<0x140>(v), dpp_i<0x140>(idx));
    argmax_step(v, idx, xor16_d(v), xor16_i(idx));
    argmax_step(v, idx, xor32_d(v), xor32_i(idx));
}
// Reduce-scatter of 16 values over a wavefront: afterwards lane L holds the
// wavefront sum of slot (L >> 2) & 15.  Across the halves and rows one swap
// exchanges a pair of slots at once (the lower lanes keep the first of the
// pair, the upper lanes the second); within a row the partner across bit 3 is
// lane ^ 8 and across bit 2 lane 7 - i of the same 8 (any partner across the
// bit that shares the higher bits gives the same sums).
__device__ __forceinline__ double wave_reduce16(const double (&v)[16]) {
    const int L = lane();
    double a8[8], a4[4], a2[2];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        double x = v[i], y = v[i + 8];
        swap32_d(x, y);
        a8[i] = x + y;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double x = a8[i], y = a8[i + 4];
        swap16_d(x, y);
        a4[i] = x + y;
    }
    const bool h3 = L & 8, h2 = L & 4;
#pragma unroll
    for (int i = 0; i < 2; ++i) a2[i] = (h3 ? a4[i + 2] : a4[i]) + dpp_d<0x128>(h3 ? a4[i] : a4[i + 2]);
    double a1 = (h2 ? a2[1] : a2[0]) + dpp_d<0x141>(h2 ? a2[0] : a2[1]);
    a1 += dpp_d<0x4E>(a1);
    a1 += dpp_d<0xB1>(a1);
    return a1;
}


// uniform values in scalar registers: function arguments of device functions
// arrive in vector registers and would otherwise count as divergent
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ bool uni(bool v) { return __builtin_amdgcn_readfirstlane((int)v) != 0; }
__device__ __forceinline__ double uni(double v) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(bits & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(bits >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <class T> __device__ __forceinline__ T* uni(T* p) {
    const unsigned long long u = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(u & 0xffffffffull));
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return (T*)(((unsigned long long)hi << 32) | lo);
}

// value of lane l (uniform l) of a double, through the scalar unit
__device__ __forceinline__ double readlane_d(double v, int l) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sum over the workgroup of n values at once (4 per LDS pass); all threads get the result
__device__ void wg_sum8(double* v, int n, Lds&) {
    Lds& s = g_lds;
    for (int q = 0; q < n; ++q) v[q] = wave_sum(v[q]);
    for (int base = 0; base < n; base += 4) {
        const int m = (n - base) < 4 ? (n - base) : 4;
        if (lane() == 0)
            for (int q = 0; q < m; ++q) s.red[q * 4 + wave()] = v[base + q];
        __syncthreads();
        for (int q = 0; q < m; ++q) v[base + q] = s.red[q * 4 + 0] + s.red[q * 4 + 1] + s.red[q * 4 + 2] + s.red[q * 4 + 3];
        __syncthreads();
    }
}
__device__ double wg_max(double v, Lds&) {
    Lds& s = g_lds;
    v = wave_max_nan(v);
    if (lane() == 0) s.red[wave()] = v;
    __syncthreads();
    double r = nan_max(nan_max(s.red[0], s.red[1]), nan_max(s.red[2], s.red[3]));
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Random numbers: minstd_rand0 and the canonical double of
// std::bernoulli_distribution (two engine draws per decision, C++ standard
// [rand.util.canonical] with k = 2).  Evaluated without contraction so that
// the decision bits equal those of the CPU generator.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t minstd_next(uint64_t x) {
    const uint64_t p = 16807ull * x;             // < 2^46
    uint64_t r = (p & 0x7fffffffull) + (p >> 31); // p mod (2^31-1), one fold
    if (r >= 0x7fffffffull) r -= 0x7fffffffull;
    return r;
}
__device__ __forceinline__ double canonical(uint64_t& x) {
#pragma clang fp contract(off)
    const double R = 2147483646.0;
    x = minstd_next(x);
    const double a = (double)(x - 1);
    x = minstd_next(x);
    const double b = (double)(x - 1);
    double sum = 0.0, tmp = 1.0;
    sum = sum + a * tmp;
    tmp = tmp * R;
    sum = sum + b * tmp;
    tmp = tmp * R;
    double ret = sum / tmp;
    if (ret >= 1.0) ret = 0x1.fffffffffffffp-1;
    return ret;
}

// minstd_rand0 is a Lehmer generator, x_{n+j} = 16807^j x_n mod (2^31 - 1), so
// the draws of a whole Metropolis block can be formed at once: the canonical
// double of site p of a block uses draws 2p+1 and 2p+2 after the block's start.
struct MinstdPowers {
    uint64_t v[2 * 32 + 2];
};
constexpr MinstdPowers make_minstd_powers() {
    MinstdPowers t{};
    uint64_t x = 1;
    for (int i = 0; i < 2 * 32 + 2; ++i) { t.v[i] = x; x = (x * 16807ull) % 0x7fffffffull; }
    return t;
}
__constant__ MinstdPowers kMinstdPow = make_minstd_powers();

__device__ __forceinline__ uint64_t mulmod_m31(uint64_t x, uint64_t y) {   // x, y < 2^31 - 1
    const uint64_t p = x * y;                                          // < 2^62
    uint64_t r = (p & 0x7fffffffull) + (p >> 31);                      // < 2^32
    r = (r & 0x7fffffffull) + (r >> 31);                               // <= 2^31
    if (r >= 0x7fffffffull) r -= 0x7fffffffull;
    return r;
}
// the canonical double of site p of a block starting at state x (two draws per site),
// with the same operations as canonical()
__device__ __forceinline__ double block_canonical(uint64_t x, int p) {
#pragma clang fp contract(off)
    const double R = 2147483646.0;
    const double a = (double)(mulmod_m31(x, kMinstdPow.v[2 * p + 1]) - 1);
    const double b = (double)(mulmod_m31(x, kMinstdPow.v[2 * p + 2]) - 1);
    double sum = 0.0, tmp = 1.0;
    sum = sum + a * tmp;
    tmp = tmp * R;
    sum = sum + b * tmp;
    tmp = tmp * R;
    double ret = sum / tmp;
    if (ret >= 1.0) ret = 0x1.fffffffffffffp-1;
    return ret;
}

// ---------------------------------------------------------------------------
// Workgroup GEMM (FP64):  C = rs .* (opA(A) diag(ms) opB(B)) .* cs + beta*C
// column-major; 64x64 output tiles, 4x4 register micro-tiles per thread,
// 16-deep k-slabs staged through LDS.  rs/ms/cs may be null (=1).
// C must not alias A or B.
// ---------------------------------------------------------------------------
__device__ __noinline__ void wg_gemm(int M, int Nc, int K, const double* A, int lda, bool tA, const double* B, int ldb, bool tB,
                        double* C, int ldc, const double* rs, const double* ms, const double* cs, double beta, Lds&) {
    Lds& s = g_lds;
    const int tx = tid() & 15, ty = tid() >> 4;
    const int tm = (M + 63) >> 6, tn = (Nc + 63) >> 6;
    for (int tile = 0; tile < tm * tn; ++tile) {
        const int r0 = (tile % tm) << 6, c0 = (tile / tm) << 6;
        double acc[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
        for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = tid() + NT * q;
                int m, k;
                if (!tA) { m = e & 63; k = e >> 6; } else { k = e & 15; m = e >> 4; }
                const int gm = r0 + m, gk = k0 + k;
                double v = 0.0;
                if (gm < M && gk < K) {
                    v = tA ? A[(size_t)gm * lda + gk] : A[(size_t)gk * lda + gm];
                    if (ms) v *= ms[gk];
                }
                s.As[k][m] = v;
                int n;
                if (!tB) { k = e & 15; n = e >> 4; } else { n = e & 63; k = e >> 6; }
                const int gn = c0 + n, gk2 = k0 + k;
                double w = 0.0;
                if (gn < Nc && gk2 < K) w = tB ? B[(size_t)gk2 * ldb + gn] : B[(size_t)gn * ldb + gk2];
                s.Bs[k][n] = w;
            }
            __syncthreads();
#pragma unroll 4
            for (int k = 0; k < 16; ++k) {
                double a[4], b[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) a[r] = s.As[k][tx + 16 * r];
#pragma unroll
                for (int c = 0; c < 4; ++c) b[c] = s.Bs[k][ty + 16 * c];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[r][c] += a[r] * b[c];
            }
            __syncthreads();
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int gn = c0 + ty + 16 * c;
            if (gn >= Nc) continue;
            const double csc = cs ? cs[gn] : 1.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = r0 + tx + 16 * r;
                if (gm >= M) continue;
                double v = acc[r][c] * csc;
                if (rs) v *= rs[gm];
                double* cp = C + (size_t)gn * ldc + gm;
                *cp = (beta != 0.0) ? v + beta * (*cp) : v;
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// The same contract on the matrix cores.  Each of the 4 wavefronts owns a
// 32x32 quarter of the 64x64 tile as 2x2 blocks of v_mfma_f64_16x16x4f64;
// 16x16 blocks outside the output are skipped.
// Operand lanes: A[row = lane%16][k = lane/16], B[k = lane/16][col = lane%16];
// result register i of lane l holds D[l/16 + 4i][l%16] (FP64 16x16x4 layout).
// With lowp the tile is rounded to FP32 and multiplied by
// v_mfma_f32_16x16x4f32 with FP32 accumulation (result D[4(l/16)+i][l%16]).
// Pipelining: the loads of the next 16-deep k-slab (or of the next tile's
// first slab) are issued before the matrix-core work on the current slab, so
// their latency hides behind it.  aL / bL say that A / B live in LDS; otherwise
// they are read with global (not flat) loads, which keeps the LDS wait counter
// free of them.  The scalings rs, ms, cs, when given, live in LDS, and C in
// global memory: a flat access could only be waited for together with every
// load in flight.
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
#define AS_G __attribute__((address_space(1)))
#define AS_L __attribute__((address_space(3)))
// An LDS row address as an opaque register value: the element offsets of the
// reads along the row then fold into the LDS instructions' offset fields
// (otherwise the compiler reassociates the row base with the array base and
// materializes one address register per read).
__device__ __forceinline__ const AS_L double* lds_row(const double* p) {
    unsigned v = (unsigned)(uintptr_t)(const AS_L double*)p;
    asm volatile("" : "+v"(v));
    return (const AS_L double*)(uintptr_t)v;
}
#define AS_C __attribute__((address_space(4)))
// A chain's descriptor, read through the constant address space: its fields are
// uniform, so they load through the scalar cache (a generic reference would make
// every field read a flat load that also holds up the next LDS wait)
typedef const AS_C ChainDesc& Desc;
template <bool lds> struct MemSpace { typedef const AS_G double* cptr; };
template <> struct MemSpace<true> { typedef const AS_L double* cptr; };

#define KS 16            // slab depth of the GEMM
template <int V> struct IntC { static constexpr int value = V; };
// Two-ahead slab pipeline.  A step first reads every operand of its 16-deep
// slab from LDS into registers and releases the slab at a barrier; its MFMAs
// are then issued in one basic block together with the LDS store of the next
// slab (whose global loads were issued one step earlier) and the global loads
// of the slab after that, so the matrix cores run while the slab traffic is
// issued.  Out-of-range elements are read from offset 0 and zeroed by a bit
// mask at the store, without branches.  The (tile, slab) cursors are kept in
// scalar registers and advanced by additions; each thread's element offsets
// within a slab are computed once.  C is prefetched before the last step of a
// tile when beta != 0.
// TM = 64: 64x64 tiles, wavefront (w & 1, w >> 1) owns the 32x32 quarter; TM = 16 (outputs of at
// most 16 rows): 16x128 tiles, wavefront w owns columns 32w..32w+31, so all four wavefronts work.
template <bool lowp, bool aL, bool bL, int TM>
__device__ __forceinline__ void gemm_body(int M, int Nc, int K, const double* A_, int lda, bool tA, const double* B_,
                                          int ldb, bool tB, double* C, int ldc, const double* rs, const double* ms_,
                                          const double* cs, double beta, double alpha) {
    Lds& s = g_lds;
    M = uni(M); Nc = uni(Nc); K = uni(K); lda = uni(lda); ldb = uni(ldb); ldc = uni(ldc);
    tA = uni(tA); tB = uni(tB); beta = uni(beta); alpha = uni(alpha);
    A_ = uni(A_); B_ = uni(B_); C = uni(C); rs = uni(rs); ms_ = uni(ms_); cs = uni(cs);
    typedef typename MemSpace<aL>::cptr APtr;
    typedef typename MemSpace<bL>::cptr BPtr;
    const APtr A = (APtr)A_;
    const BPtr B = (BPtr)B_;
    const bool has_ms = ms_ != nullptr, has_rs = rs != nullptr, has_cs = cs != nullptr;
    // the scalings live in LDS (null tests above on the generic pointers); C is in global memory
    const AS_L double* ms = has_ms ? (const AS_L double*)ms_ : (const AS_L double*)&s.As[0][0];
    const AS_L double* rsl = (const AS_L double*)rs;
    const AS_L double* csl = (const AS_L double*)cs;
    AS_G double* const Cg = (AS_G double*)C;
    const int l = lane();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int TN = TM == 64 ? 64 : 128;              // tile columns
    constexpr int SQA = TM * KS / NT, SQB = TN * KS / NT;  // slab elements per thread (A, B)
    constexpr int RBM = TM == 64 ? 2 : 1;                // 16-row blocks per wavefront
    // slab layout in the As/Bs space (2 x 16 x GLD doubles): TM = 64: A and B at row stride GLD;
    // TM = 16: B (16 x 128) at row stride 144, then A (16 x 16) at row stride 16 (both strides
    // are 32 mod 64 dwords, so the lane groups of an operand read fall on disjoint banks)
    constexpr int GA = TM == 64 ? GLD : 16, GB = TM == 64 ? GLD : 144;
    static_assert(TM == 64 || KS * GB + KS * GA <= 2 * 16 * GLD, "skinny slabs exceed the GEMM staging space");
    double* const Asl = TM == 64 ? &s.As[0][0] : &s.As[0][0] + KS * GB;
    double* const Bsl = TM == 64 ? &s.Bs[0][0] : &s.As[0][0];
    const int wr = TM == 64 ? (w & 1) * 32 : 0, wc = TM == 64 ? (w >> 1) * 32 : w * 32;
    if (M <= 0 || Nc <= 0 || K <= 0) { __syncthreads(); return; }
    float* Af = reinterpret_cast<float*>(Asl);  // [16][2 GA] floats in lowp mode
    float* Bf = reinterpret_cast<float*>(Bsl);
    const int t = tid();
    // this thread's elements of a slab (SQA of A, SQB of B) are affine in their index q:
    //   A element q at (row ma0 + q*dma, depth ka0 + q*dka), offset oa0 + q*doa from the slab origin,
    //   B element q at (depth kb0 + q*dkb, column nb0 + q*dnb), offset ob0 + q*dob
    const int ma0 = tA ? t / KS : t % TM, ka0 = tA ? t % KS : t / TM;
    const int dma = tA ? NT / KS : 0, dka = tA ? 0 : NT / TM;
    const int kb0 = tB ? t / TN : t % KS, nb0 = tB ? t % TN : t / KS;
    const int dkb = tB ? NT / TN : 0, dnb = tB ? 0 : NT / KS;
    const int oa0 = tA ? ma0 * lda + ka0 : ka0 * lda + ma0, doa = tA ? dma * lda : dka * lda;
    const int ob0 = tB ? kb0 * ldb + nb0 : nb0 * ldb + kb0, dob = tB ? dkb * ldb : dnb * ldb;
    // slab origin (r0, c0, k0) -> offsets of its first element in A and B (strides
    // chosen once, so that the per-step offsets need no select on the transposes)
    const int sar = tA ? lda : 1, sak = tA ? 1 : lda, sbc = tB ? 1 : ldb, sbk = tB ? ldb : 1;
    auto a_base = [&](int r0, int k0) { return r0 * sar + k0 * sak; };
    auto b_base = [&](int c0, int k0) { return c0 * sbc + k0 * sbk; };
    double va[SQA], vb[SQB], msv[SQA];
    unsigned okm = 0;   // bit q: A element q of the loaded slab in range; bit SQA + q: B element q
    auto load = [&](int r0, int c0, int k0) {
        const int ba = a_base(r0, k0), bb = b_base(c0, k0);
        const int mr = M - r0, nr = Nc - c0, kr = K - k0;
        unsigned ok = 0;
#pragma unroll
        for (int q = 0; q < SQA; ++q) {
            const bool okA = (ma0 + q * dma < mr) & (ka0 + q * dka < kr);
            va[q] = A[okA ? ba + oa0 + q * doa : 0];
            ok |= (unsigned)okA << q;
        }
#pragma unroll
        for (int q = 0; q < SQB; ++q) {
            const bool okB = (nb0 + q * dnb < nr) & (kb0 + q * dkb < kr);
            vb[q] = B[okB ? bb + ob0 + q * dob : 0];
            ok |= (unsigned)okB << (SQA + q);
        }
        okm = ok;
    };
    // the K scalings of the loaded slab (depth origin k0s), read ahead of its store
    auto load_ms = [&](int k0s) {
#pragma unroll
        for (int q = 0; q < SQA; ++q) msv[q] = ms[has_ms ? min(k0s + ka0 + q * dka, K - 1) : 0];
    };
    // loaded slab -> LDS
    auto store = [&]() {
#pragma unroll
        for (int q = 0; q < SQA; ++q) {
            const int m = ma0 + q * dma, k = ka0 + q * dka;
            double v = ((okm >> q) & 1u) ? va[q] : 0.0;
            v = has_ms ? v * msv[q] : v;
            if (lowp) Af[k * 2 * GA + m] = (float)v;
            else Asl[k * GA + m] = v;
        }
#pragma unroll
        for (int q = 0; q < SQB; ++q) {
            const int kbq = kb0 + q * dkb, n = nb0 + q * dnb;
            const double x = ((okm >> (SQA + q)) & 1u) ? vb[q] : 0.0;
            if (lowp) Bf[kbq * 2 * GB + n] = (float)x;
            else Bsl[kbq * GB + n] = x;
        }
    };
    // cursors: (r0, c0, k0) the slab multiplied in this step, sk the depth origin of
    // the slab held in registers, (lr, lc, lk) the next slab to load; loads past the
    // last slab read in-range elements of A and are never used
    int r0 = 0, c0 = 0, k0 = 0;
    int lr = 0, lc = 0, lk = 0;
    auto advance = [&](int& r, int& c, int& k) {
        k += KS;
        if (k >= K) { k = 0; r += TM; if (r >= M) { r = 0; c += TN; } }
    };
    load(lr, lc, lk);
    advance(lr, lc, lk);
    load_ms(0);
    store();
    int sk = lk;
    load(lr, lc, lk);
    advance(lr, lc, lk);
    dbl4 acc[2][2];
    flt4 accf[2][2];
    double cv[2][2][4];
    // element e of acc[rb][cb]: row rb*16 + l%16, column cb*16 + (lowp ? 4(l/16) + e : l/16 + 4e)
    const int ocl = (l & 15) + (lowp ? 4 * (l >> 4) : (l >> 4)) * ldc;
    while (c0 < Nc) {
        if (k0 == 0) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
                    accf[a][b] = flt4{0.f, 0.f, 0.f, 0.f};
                }
        }
        const int rows = M - r0 - wr, cols = Nc - c0 - wc;   // extent of this wave's quarter
        const bool last = k0 + KS >= K;
        const int cbase = (c0 + wc) * ldc + r0 + wr;
        if (last && beta != 0.0) {   // C of the tile, in flight during the last step
#pragma unroll
            for (int rb = 0; rb < RBM; ++rb)
#pragma unroll
                for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int row = rb * 16 + (l & 15);
                        const int col = cb * 16 + (lowp ? 4 * (l >> 4) + e : (l >> 4) + 4 * e);
                        const bool ok = (row < rows) & (col < cols);   // (others are never stored)
                        cv[rb][cb][e] = Cg[ok ? cbase + ocl + (cb * 16 + (lowp ? e : 4 * e)) * ldc + rb * 16 : 0];
                    }
        }
        __syncthreads();   // the slab of this step is in LDS
        double oa[4][2], ob[4][2];
        float fa[4][2], fb[4][2];
#pragma unroll
        for (int kk = 0; kk < KS / 4; ++kk) {
            const int k = kk * 4 + (l >> 4);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (lowp) {
                    if (h < RBM) fa[kk][h] = Af[k * 2 * GA + wr + 16 * h + (l & 15)];
                    fb[kk][h] = Bf[k * 2 * GB + wc + 16 * h + (l & 15)];
                } else {
                    if (h < RBM) oa[kk][h] = Asl[k * GA + wr + 16 * h + (l & 15)];
                    ob[kk][h] = Bsl[k * GB + wc + 16 * h + (l & 15)];
                }
            }
        }
        load_ms(sk);
        __syncthreads();   // operands in registers: the slab buffer is free
        // the tile is formed transposed (B^T A^T): lane l then holds C rows l%16 of its
        // block, so the C reads and writes of the epilogue run down columns
        auto step = [&](auto RBc, auto CBc) {
            constexpr int RB = decltype(RBc)::value, CB = decltype(CBc)::value;
#pragma unroll
            for (int kk = 0; kk < KS / 4; ++kk)
#pragma unroll
                for (int a = 0; a < RB; ++a)
#pragma unroll
                    for (int b = 0; b < CB; ++b) {
                        if (lowp) accf[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[kk][b], fa[kk][a], accf[a][b], 0, 0, 0);
                        else acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ob[kk][b], oa[kk][a], acc[a][b], 0, 0, 0);
                    }
            store();
            sk = lk;
            load(lr, lc, lk);
            // issue order: each MFMA followed by a share of the slab traffic, so that
            // the stores, loads and their address arithmetic run under the matrix cores
            constexpr int NM = RB * CB * (KS / 4);
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
                __builtin_amdgcn_sched_group_barrier(0x002, 6, 0);   // VALU
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);   // LDS write
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);   // global load
            }
        };
        if (RBM == 2 && rows > 16 && cols > 16) step(IntC<RBM>(), IntC<2>());
        else if (RBM == 2 && rows > 16 && cols > 0) step(IntC<RBM>(), IntC<1>());
        else if (rows > 0 && cols > 16) step(IntC<1>(), IntC<2>());
        else if (rows > 0 && cols > 0) step(IntC<1>(), IntC<1>());
        else step(IntC<0>(), IntC<0>());
        advance(lr, lc, lk);
        if (last) {
            // tile done: the scalings, then the stores
            double rsv[2] = {1.0, 1.0}, csv[2][4] = {{1.0, 1.0, 1.0, 1.0}, {1.0, 1.0, 1.0, 1.0}};
            if (has_cs) {
#pragma unroll
                for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int col = cb * 16 + (lowp ? 4 * (l >> 4) + e : (l >> 4) + 4 * e);
                        csv[cb][e] = csl[col < cols ? c0 + wc + col : 0];
                    }
            }
            if (has_rs) {
#pragma unroll
                for (int rb = 0; rb < RBM; ++rb) {
                    const int row = rb * 16 + (l & 15);
                    rsv[rb] = rsl[row < rows ? r0 + wr + row : 0];
                }
            }
#pragma unroll
            for (int rb = 0; rb < RBM; ++rb)
#pragma unroll
                for (int cb = 0; cb < 2; ++cb) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int row = rb * 16 + (l & 15);
                        const int col = cb * 16 + (lowp ? 4 * (l >> 4) + e : (l >> 4) + 4 * e);
                        if (row >= rows || col >= cols) continue;
                        double v = (lowp ? (double)accf[rb][cb][e] : acc[rb][cb][e]) * alpha;
                        if (has_cs) v *= csv[cb][e];
                        if (has_rs) v *= rsv[rb];
                        Cg[cbase + ocl + (cb * 16 + (lowp ? e : 4 * e)) * ldc + rb * 16] =
                            (beta != 0.0) ? v + beta * cv[rb][cb][e] : v;
                    }
                }
        }
        advance(r0, c0, k0);
    }
    __syncthreads();
}

template <bool lowp, bool aL, bool bL>
__device__ __noinline__ void wg_gemm_mfma(int M, int Nc, int K, const double* A, int lda, bool tA, const double* B,
                                          int ldb, bool tB, double* C, int ldc, const double* rs, const double* ms,
                                          const double* cs, double beta, Lds&, double alpha = 1.0) {
    if (uni(M) <= 16) gemm_body<lowp, aL, bL, 16>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, alpha);
    else gemm_body<lowp, aL, bL, 64>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, alpha);
}

// ---------------------------------------------------------------------------
// Wide tiles for the N x N x N products (FP64, A and B in global memory).  The
// 64x64 tiles above give each wavefront 2x2 blocks, i.e. 16 MFMAs per slab step
// between two workgroup barriers, and re-read A and B once per tile row and
// column.  Here each wavefront keeps RB x CB blocks of 16x16 in accumulators
// (a (32 RB) x (32 CB) workgroup tile: 160 x 160 with RB = CB = 5 covers an
// N = 144 matrix in one tile, 128 x 128 tiles cover N = 256 in four), so a step
// issues RB CB KW / 4 MFMAs per wavefront and A and B are read once per tile.  The
// slabs (KW = 16 deep, 8 for the 160 x 160 tile, whose accumulators leave fewer
// registers for staging) are double-buffered in the LDS panel, which the callers
// leave free.
// Every output element sums its products in the same order as gemm_body (16-deep
// slabs in k order, four 16x16x4 MFMAs each, formed transposed), and the
// epilogue applies alpha, cs, rs and beta C in the same order, so the two
// GEMMs give bit-identical results.
// ---------------------------------------------------------------------------
template <int RB, int CB, int KW>
__device__ __noinline__ void wg_gemm_wide(int M, int Nc, int K, const double* A_, int lda, bool tA, const double* B_,
                                          int ldb, bool tB, double* C, int ldc, const double* rs, const double* ms_,
                                          const double* cs, double beta, double alpha) {
    Lds& s = g_lds;
    M = uni(M); Nc = uni(Nc); K = uni(K); lda = uni(lda); ldb = uni(ldb); ldc = uni(ldc);
    tA = uni(tA); tB = uni(tB); beta = uni(beta); alpha = uni(alpha);
    const AS_G double* const A = (const AS_G double*)uni(A_);
    const AS_G double* const B = (const AS_G double*)uni(B_);
    AS_G double* const Cg = (AS_G double*)uni(C);
    rs = uni(rs); ms_ = uni(ms_); cs = uni(cs);
    const bool has_ms = ms_ != nullptr, has_rs = rs != nullptr, has_cs = cs != nullptr;
    // the scalings live in LDS (null tests above on the generic pointers)
    const AS_L double* ms = has_ms ? (const AS_L double*)ms_ : (const AS_L double*)s.panel;
    const AS_L double* rsl = (const AS_L double*)rs;
    const AS_L double* csl = (const AS_L double*)cs;
    constexpr int TM = 32 * RB, TN = 32 * CB;      // workgroup tile
    constexpr int SA = TM + 16, SB = TN + 16;      // LDS row strides of the k-major slabs (16 mod 32 doubles)
    constexpr int SQA = TM * KW / NT, SQB = TN * KW / NT;   // slab elements per thread (KW-deep slabs)
    constexpr int BUF = KW * (SA + SB);            // one slab buffer (A then B)
    static_assert(KW == 8 || KW == 16, "slab depth");
    static_assert(SQA == RB * (KW / 8) && SQB == CB * (KW / 8), "slab elements per thread");
    static_assert(2 * BUF <= PANEL_DOUBLES, "two wide slabs exceed the LDS panel");
    static_assert(SQA + SQB <= 32, "slab range bits");
    if (M <= 0 || Nc <= 0 || K <= 0) { __syncthreads(); return; }
    const int l = lane();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wr = (w & 1) * 16 * RB, wc = (w >> 1) * 16 * CB;   // this wavefront's quarter of the tile
    const int t = tid();
    // Slab element q of this thread, branch-free for either layout: along the operand's
    // contiguous dimension in memory the lanes read consecutive doubles.
    //   rows contiguous (A, B^T):  row = t%32 + 32 (q % R),   depth = t/32 + 8 (q / R)
    //   depth contiguous (A^T, B): row = t/KW + (NT/KW) q,   depth = t%KW
    // i.e. row = rb0 + dr1 (q % R) + dr2 (q / R), depth = kb0 + dk2 (q / R), R = RB (A) or CB (B)
    constexpr int RK = NT / KW;   // rows per thread group in the depth-contiguous map
    const int ra0 = tA ? t / KW : t & 31, ka0 = tA ? t % KW : t >> 5;
    const int dra1 = tA ? RK : 32, dra2 = tA ? RK * RB : 0, dka2 = tA ? 0 : 8;
    const int nb0 = tB ? t & 31 : t / KW, kb0 = tB ? t >> 5 : t % KW;
    const int dnb1 = tB ? 32 : RK, dnb2 = tB ? 0 : RK * CB, dkb2 = tB ? 8 : 0;
    // memory strides of a row and of a depth step: op(A)(m, k) at A[m sar + k sak], op(B)(k, n) at B[n sbn + k sbk]
    const int sar = tA ? lda : 1, sak = tA ? 1 : lda, sbn = tB ? 1 : ldb, sbk = tB ? ldb : 1;
    const int oa0 = ra0 * sar + ka0 * sak, ob0 = nb0 * sbn + kb0 * sbk;
    const int oa1 = dra1 * sar, oa2 = dra2 * sar + dka2 * sak;
    const int ob1 = dnb1 * sbn, ob2 = dnb2 * sbn + dkb2 * sbk;
    const int la0 = ka0 * SA + ra0, la1 = dra1, la2 = dra2 + dka2 * SA;
    const int lb0 = kb0 * SB + nb0, lb1 = dnb1, lb2 = dnb2 + dkb2 * SB;
    double va[SQA], vb[SQB], msv[SQA];
    unsigned okm = 0;   // bit q: A element q of the loaded slab in range; bit SQA + q: B element q
    auto load = [&](int r0, int c0, int k0) {
        const int mr = M - r0, nr = Nc - c0, kr = K - k0;
        const int ba = r0 * sar + k0 * sak, bb = c0 * sbn + k0 * sbk;
        unsigned ok = 0;
#pragma unroll
        for (int q = 0; q < SQA; ++q) {
            const int i1 = q % RB, i2 = q / RB;
            const bool okA = (ra0 + i1 * dra1 + i2 * dra2 < mr) & (ka0 + i2 * dka2 < kr);
            va[q] = A[okA ? ba + oa0 + i1 * oa1 + i2 * oa2 : 0];
            ok |= (unsigned)okA << q;
        }
#pragma unroll
        for (int q = 0; q < SQB; ++q) {
            const int i1 = q % CB, i2 = q / CB;
            const bool okB = (nb0 + i1 * dnb1 + i2 * dnb2 < nr) & (kb0 + i2 * dkb2 < kr);
            vb[q] = B[okB ? bb + ob0 + i1 * ob1 + i2 * ob2 : 0];
            ok |= (unsigned)okB << (SQA + q);
        }
        okm = ok;
    };
    // the K scalings of the loaded slab (depth origin k0s), read ahead of its store
    auto load_ms = [&](int k0s) {
#pragma unroll
        for (int q = 0; q < SQA; ++q) msv[q] = ms[has_ms ? min(k0s + ka0 + (q / RB) * dka2, K - 1) : 0];
    };
    auto store = [&](AS_L double* buf) {
#pragma unroll
        for (int q = 0; q < SQA; ++q) {
            const int i1 = q % RB, i2 = q / RB;
            double v = ((okm >> q) & 1u) ? va[q] : 0.0;
            v = has_ms ? v * msv[q] : v;
            buf[la0 + i1 * la1 + i2 * la2] = v;
        }
#pragma unroll
        for (int q = 0; q < SQB; ++q) {
            const int i1 = q % CB, i2 = q / CB;
            buf[KW * SA + lb0 + i1 * lb1 + i2 * lb2] = ((okm >> (SQA + q)) & 1u) ? vb[q] : 0.0;
        }
    };
    AS_L double* const buf0 = (AS_L double*)s.panel;
    int r0 = 0, c0 = 0, k0 = 0;
    int lr = 0, lc = 0, lk = 0;
    auto advance = [&](int& r, int& c, int& k) {
        k += KW;
        if (k >= K) { k = 0; r += TM; if (r >= M) { r = 0; c += TN; } }
    };
    // double-buffered slabs: a step reads buffer (step & 1) while it stores the next slab into
    // the other one (read in the previous step, which every wavefront has finished at the
    // step's barrier); the slab after that is loaded into registers meanwhile
    load(lr, lc, lk);
    advance(lr, lc, lk);
    load_ms(0);
    store(buf0);
    int sk = lk;
    load(lr, lc, lk);
    advance(lr, lc, lk);
    int par = 0;
    dbl4 acc[RB][CB];
    while (c0 < Nc) {
        if (k0 == 0) {
#pragma unroll
            for (int a = 0; a < RB; ++a)
#pragma unroll
                for (int b = 0; b < CB; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
        }
        const bool last = k0 + KW >= K;
        AS_L double* const cur = buf0 + par * BUF;
        AS_L double* const nxt = buf0 + (par ^ 1) * BUF;
        load_ms(sk);
        __syncthreads();   // this step's slab is in `cur`; `nxt` is no longer read
#pragma unroll
        for (int kk = 0; kk < KW / 4; ++kk) {
            const int k = kk * 4 + (l >> 4);
            double oa[RB], ob[CB];
#pragma unroll
            for (int h = 0; h < RB; ++h) oa[h] = cur[k * SA + wr + 16 * h + (l & 15)];
#pragma unroll
            for (int h = 0; h < CB; ++h) ob[h] = cur[KW * SA + k * SB + wc + 16 * h + (l & 15)];
#pragma unroll
            for (int a = 0; a < RB; ++a)
#pragma unroll
                for (int b = 0; b < CB; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ob[b], oa[a], acc[a][b], 0, 0, 0);
            if (kk == 0) store(nxt);   // the next slab, under the matrix cores
        }
        sk = lk;
        load(lr, lc, lk);
        advance(lr, lc, lk);
        par ^= 1;
        if (last) {
            // tile done: element e of acc[rb][cb] is C(r0 + wr + 16 rb + l%16, c0 + wc + 16 cb + l/16 + 4e)
            const int rows = M - r0 - wr, cols = Nc - c0 - wc;
            const int cbase = (c0 + wc) * ldc + r0 + wr + (l & 15) + (l >> 4) * ldc;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                const int row = rb * 16 + (l & 15);
                const double rsv = has_rs ? rsl[row < rows ? r0 + wr + row : 0] : 1.0;
                double cv[CB][4];
                if (beta != 0.0) {
#pragma unroll
                    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int col = cb * 16 + (l >> 4) + 4 * e;
                            const bool ok = (row < rows) & (col < cols);
                            cv[cb][e] = Cg[ok ? cbase + (cb * 16 + 4 * e) * ldc + rb * 16 : 0];
                        }
                }
#pragma unroll
                for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int col = cb * 16 + (l >> 4) + 4 * e;
                        if (row >= rows || col >= cols) continue;
                        double v = acc[rb][cb][e] * alpha;
                        if (has_cs) v *= csl[c0 + wc + col];
                        if (has_rs) v *= rsv;
                        Cg[cbase + (cb * 16 + 4 * e) * ldc + rb * 16] = (beta != 0.0) ? v + beta * cv[cb][e] : v;
                    }
            }
        }
        advance(r0, c0, k0);
    }
    __syncthreads();
}

// the FP64 product of two global operands: wide tiles when the output and the depth are large
// enough to fill them, the 64x64 tiles otherwise (both give the same bits)
__device__ __forceinline__ void gemm64_global(int M, int Nc, int K, const double* A, int lda, bool tA, const double* B,
                                              int ldb, bool tB, double* C, int ldc, const double* rs,
                                              const double* ms, const double* cs, double beta, double alpha) {
    Lds& s = g_lds;
    M = uni(M); Nc = uni(Nc); K = uni(K);
    if (M >= 96 && Nc >= 96 && K >= 32) {
        // the tile shape that pads the output least (measured on one CU: N = 144 122 -> 69 us with
        // 160 x 160 tiles, N = 256 394 -> 273 us and N = 196 304 -> 228 us with 128 x 128 tiles,
        // N = 400 1783 -> 1400 us with 160 x 160 tiles)
        const int p128 = ((M + 127) / 128) * ((Nc + 127) / 128) * 128 * 128;
        const int p160 = ((M + 159) / 160) * ((Nc + 159) / 160) * 160 * 160;
        if (p160 < p128) wg_gemm_wide<5, 5, 8>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, alpha);
        else wg_gemm_wide<4, 4, 16>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, alpha);
        return;
    }
    wg_gemm_mfma<false, false, false>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, s, alpha);
}
// C <- C - V op(T) (V^T C) for a block reflector (V: m x jb at ld ldv, T: jb x jb at ld NB,
// op(T) = T^T when tT; C: m x n at ld ldc; W1, W2: jb x n scratch at ld NB), the three
// products in one call so that the register save and restore of a call is paid once
__device__ __noinline__ void wg_reflect(const double* V, int ldv, const double* T, bool tT, double* C, int ldc, int m,
                                        int n, int jb, double* W1, double* W2, Lds&) {
    // W1 = V^T C and W2 = op(T) W1 have jb <= 16 rows (16-row tiles), C -= V W2 has m rows
    for (int i = 0; i < 2; ++i) {
        const int K = i == 0 ? m : jb;
        const double* A = i == 1 ? T : V;
        const int lda = i == 1 ? NB : ldv;
        const bool tA = i == 0 ? true : tT;
        const double* B = i == 0 ? C : W1;
        const int ldb = i == 0 ? ldc : NB;
        double* const Co = i == 0 ? W1 : W2;
        gemm_body<false, false, false, 16>(jb, n, K, A, lda, tA, B, ldb, false, Co, NB, nullptr, nullptr, nullptr, 0.0,
                                           1.0);
    }
    gemm_body<false, false, false, 64>(m, n, jb, V, ldv, false, W2, NB, false, C, ldc, nullptr, nullptr, nullptr, 1.0,
                                       -1.0);
}

// GEMMs of the engine: FP64 on the matrix cores for everything that stays
// double precision; the slice products and wraps drop to FP32 MFMA when the
// chain runs in mixed precision.  gemm64: A and B in global memory;
// gemm64_lg: A in LDS; gemm64_ll: both in LDS.
__device__ __forceinline__ void gemm64(int M, int Nc, int K, const double* A, int lda, bool tA, const double* B,
                                       int ldb, bool tB, double* C, int ldc, const double* rs, const double* ms,
                                       const double* cs, double beta, Lds&, double alpha = 1.0) {
    gemm64_global(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, alpha);
}
__device__ __forceinline__ void gemm64_lg(int M, int Nc, int K, const double* A, int lda, bool tA, const double* B,
                                          int ldb, bool tB, double* C, int ldc, double beta, Lds&, double alpha = 1.0) {
    Lds& s = g_lds;
    wg_gemm_mfma<false, true, false>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, nullptr, nullptr, nullptr, beta, s,
                                     alpha);
}
__device__ __forceinline__ void gemm64_ll(int M, int Nc, int K, const double* A, int lda, bool tA, const double* B,
                                          int ldb, bool tB, double* C, int ldc, double beta, Lds&, double alpha = 1.0) {
    Lds& s = g_lds;
    wg_gemm_mfma<false, true, true>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, nullptr, nullptr, nullptr, beta, s,
                                    alpha);
}
// FP32 matrix cores for this product: every B product in DQMC_PREC_MIXED, only the
// propagations of G (wraps, and G(t,0), G(0,t) of the time-displaced sweep) in
// DQMC_PREC_FP32_WRAP, only the slice products accumulated into the stacks in
// DQMC_PREC_FP32_PRODUCT
__device__ __forceinline__ bool lowp_product(Desc d, bool wrap) {
    const int p = d.precision;
    return p == DQMC_PREC_MIXED || (p == DQMC_PREC_FP32_WRAP && wrap) || (p == DQMC_PREC_FP32_PRODUCT && !wrap);
}
__device__ __forceinline__ void gemmB(Desc d, bool wrap, int M, int Nc, int K, const double* A, int lda, bool tA,
                                      const double* B, int ldb, bool tB, double* C, int ldc, const double* rs,
                                      const double* ms, const double* cs, Lds&) {
    Lds& s = g_lds;
    if (lowp_product(d, wrap))
        wg_gemm_mfma<true, false, false>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, 0.0, s);
    else gemm64_global(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, 0.0, 1.0);
}

// Elementwise streams over e in [0, n): st(e, ld(e)), with U loads of each
// thread issued before any of its stores, so a pass costs about one memory
// latency per U elements instead of one per element (the compiler cannot
// reorder a load above a store that may alias it).
template <int U = 8, class Ld, class St>
__device__ __forceinline__ void wg_stream(int n, Ld ld, St st) {
    int e0 = tid();
    for (; e0 + (U - 1) * NT < n; e0 += U * NT) {
        double v[U];
#pragma unroll
        for (int q = 0; q < U; ++q) v[q] = ld(e0 + q * NT);
#pragma unroll
        for (int q = 0; q < U; ++q) st(e0 + q * NT, v[q]);
    }
    if (e0 < n) {   // the rest in one more round: indices clamped for the loads, stores predicated
        double v[U];
#pragma unroll
        for (int q = 0; q < U; ++q) v[q] = ld(min(e0 + q * NT, n - 1));
#pragma unroll
        for (int q = 0; q < U; ++q)
            if (e0 + q * NT < n) st(e0 + q * NT, v[q]);
    }
}
// two streams of the same length with the loads of both in flight in every round
template <int U = 8, class Ld1, class St1, class Ld2, class St2>
__device__ __forceinline__ void wg_stream2(int n, Ld1 ld1, St1 st1, Ld2 ld2, St2 st2) {
    for (int e0 = tid(); e0 < n; e0 += U * NT) {
        double v[U], w[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int e = min(e0 + q * NT, n - 1);   // (clamped in the last round; not stored)
            v[q] = ld1(e);
            w[q] = ld2(e);
        }
#pragma unroll
        for (int q = 0; q < U; ++q)
            if (e0 + q * NT < n) { st1(e0 + q * NT, v[q]); st2(e0 + q * NT, w[q]); }
    }
}
// Division by a wave-uniform divisor 1 <= d <= 448 without the integer divide
// sequence: q = floor(e m / 2^32) with m = ceil(2^32 / d).  With m d = 2^32 + r,
// 0 <= r < d, the product overshoots e / d by e r / (d 2^32), below 1 / d for
// e < 2^32 / 448, so q = floor(e / d) for every 0 <= e < 9.5e6 (the indices
// divided here stay below N^2 <= 448^2).
struct FastDiv {
    uint64_t m;
    int d;
    __device__ explicit FastDiv(int d_) : m(((1ull << 32) + (uint64_t)d_ - 1) / (uint64_t)d_), d(d_) {}
    __device__ __forceinline__ int div(int e) const { return (int)(((uint64_t)(uint32_t)e * m) >> 32); }
};
__device__ __forceinline__ const AS_G double* gptr(const double* p) { return (const AS_G double*)p; }
__device__ __forceinline__ AS_G double* gptr(double* p) { return (AS_G double*)p; }

// dst <- src, both in global memory
__device__ __noinline__ void wg_copy(double* dst_, const double* src_, int n2) {
    AS_G double* dst = gptr(uni(dst_));
    const AS_G double* src = gptr(uni(src_));
    wg_stream<16>(uni(n2), [&](int e) { return src[e]; }, [&](int e, double v) { dst[e] = v; });
    __syncthreads();
}
__device__ void wg_identity(double* M_, int N) {
    AS_G double* const M = gptr(M_);
    const FastDiv dN(N);
    for (int e = tid(); e < N * N; e += NT) M[e] = (e == dN.div(e) * (N + 1)) ? 1.0 : 0.0;
    __syncthreads();
}
__device__ void wg_scale_rows(double* M, int N, const double* sc) {
    const FastDiv dN(N);
    for (int e = tid(); e < N * N; e += NT) M[e] *= sc[e - dN.div(e) * N];
    __syncthreads();
}
__device__ void wg_scale_cols(double* M, int N, const double* sc) {
    const FastDiv dN(N);
    for (int e = tid(); e < N * N; e += NT) M[e] *= sc[dN.div(e)];
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Checkerboard: one sublattice pass of 4x4 plaquette rotations (in place).
// Plaquette at (x,y) with x,y = start mod 2 acts on the sites
// {(x,y), (x+1,y), (x,y+1), (x+1,y+1)}; P = [[a,b,b,c],[b,a,c,b],[b,c,a,b],[c,b,b,a]].
// ---------------------------------------------------------------------------
// sites of plaquette p = px + h py of one sublattice (start 0: even corners, 1: odd)
__device__ __forceinline__ void plaq_sites_h(int p, int side, int start, const FastDiv& dh, int idx[4]) {
    const int py = dh.div(p), px = p - py * dh.d;
    const int x = start + 2 * px, y = start + 2 * py;   // <= side - 1
    const int x1 = x + 1 == side ? 0 : x + 1, y1 = y + 1 == side ? 0 : y + 1;
    idx[0] = x + side * y; idx[1] = x1 + side * y; idx[2] = x + side * y1; idx[3] = x1 + side * y1;
}
__device__ __forceinline__ void plaq_sites(int p, int side, int start, int idx[4]) {
    const int h = side >> 1;
    const int x = start + 2 * (p % h), y = start + 2 * (p / h);
    const int x1 = (x + 1) % side, y1 = (y + 1) % side;
    idx[0] = x + side * y; idx[1] = x1 + side * y; idx[2] = x + side * y1; idx[3] = x1 + side * y1;
}
__device__ __forceinline__ void plaq_apply(const AS_C double* P, const double v[4], double o[4]) {
    const double a = P[0], b = P[1], c = P[2];
    o[0] = a * v[0] + b * v[1] + b * v[2] + c * v[3];
    o[1] = b * v[0] + a * v[1] + c * v[2] + b * v[3];
    o[2] = b * v[0] + c * v[1] + a * v[2] + b * v[3];
    o[3] = c * v[0] + b * v[1] + b * v[2] + a * v[3];
}
__device__ void wg_cb_left(double* M, int N, int side, const AS_C double* P, int start) {
    const int np = N >> 2;
    for (int e = tid(); e < np * N; e += NT) {
        const int p = e % np, col = e / np;
        int idx[4]; plaq_sites(p, side, start, idx);
        double* c = M + (size_t)col * N;
        double v[4] = {c[idx[0]], c[idx[1]], c[idx[2]], c[idx[3]]}, o[4];
        plaq_apply(P, v, o);
        c[idx[0]] = o[0]; c[idx[1]] = o[1]; c[idx[2]] = o[2]; c[idx[3]] = o[3];
    }
    __syncthreads();
}
__device__ void wg_cb_right(double* M, int N, int side, const AS_C double* P, int start) {
    const int np = N >> 2;
    for (int e = tid(); e < np * N; e += NT) {
        const int row = e % N, p = e / N;
        int idx[4]; plaq_sites(p, side, start, idx);
        double v[4] = {M[(size_t)idx[0] * N + row], M[(size_t)idx[1] * N + row], M[(size_t)idx[2] * N + row],
                       M[(size_t)idx[3] * N + row]}, o[4];
        plaq_apply(P, v, o);
        M[(size_t)idx[0] * N + row] = o[0]; M[(size_t)idx[1] * N + row] = o[1];
        M[(size_t)idx[2] * N + row] = o[2]; M[(size_t)idx[3] * N + row] = o[3];
    }
    __syncthreads();
}

// Fused checkerboard products.  A left product mixes rows within each
// column, a right product mixes columns within each row, so a sequence
// [pre-scale, sublattice pass, sublattice pass, post-scale] runs on blocks of
// whole columns (left) or whole rows (right) staged in LDS: one read and one
// write of the matrix instead of one per pass.  starts[] lists the sublattice
// order (1 = B plaquettes at odd (x,y), 0 = A at even).
__device__ void cb_pass_lds_left(double* T, int ld, int ncols, int side, const AS_C double* P, int start) {
    const int h = side / 2, np = h * h;
    const FastDiv dnp(np), dh(h);
    for (int e = tid(); e < np * ncols; e += NT) {
        const int col = dnp.div(e), pq = e - col * np;
        int idx[4]; plaq_sites_h(pq, side, start, dh, idx);
        double* c = T + (size_t)col * ld;
        double v[4] = {c[idx[0]], c[idx[1]], c[idx[2]], c[idx[3]]}, o[4];
        plaq_apply(P, v, o);
        c[idx[0]] = o[0]; c[idx[1]] = o[1]; c[idx[2]] = o[2]; c[idx[3]] = o[3];
    }
    __syncthreads();
}
// T holds nr rows x N columns, element (r, c) at c * nr + r
__device__ void cb_pass_lds_right(double* T, int nr, int side, const AS_C double* P, int start) {
    const int h = side / 2, np = h * h;
    const FastDiv dnr(nr), dh(h);
    for (int e = tid(); e < np * nr; e += NT) {
        const int pq = dnr.div(e), r = e - pq * nr;
        int idx[4]; plaq_sites_h(pq, side, start, dh, idx);
        double v[4] = {T[idx[0] * nr + r], T[idx[1] * nr + r], T[idx[2] * nr + r], T[idx[3] * nr + r]}, o[4];
        plaq_apply(P, v, o);
        T[idx[0] * nr + r] = o[0]; T[idx[1] * nr + r] = o[1]; T[idx[2] * nr + r] = o[2]; T[idx[3] * nr + r] = o[3];
    }
    __syncthreads();
}
// M <- post .* E_s2 E_s1 (pre .* M)  (row scalings; pre/post may be null)
__device__ __noinline__ void wg_cb_left_fused(double* M, int N, int side, const AS_C double* P, int s1, int s2, const double* pre,
                                 const double* post, Lds&) {
    Lds& s = g_lds;
    N = uni(N);
    AS_G double* Mg = gptr(uni(M));
    // scalings live in LDS; test the generic pointers for null (a null generic
    // pointer cast to the LDS address space is not the LDS null)
    const bool has_pre = uni(pre) != nullptr, has_post = uni(post) != nullptr;
    const AS_L double* prel = (const AS_L double*)uni(pre);
    const AS_L double* postl = (const AS_L double*)uni(post);
    const FastDiv dN(N);
    const int cb = max(1, min(N, PANEL_DOUBLES / N));
    for (int c0 = 0; c0 < N; c0 += cb) {
        const int nc = min(cb, N - c0);
        AS_G double* blk = Mg + (size_t)c0 * N;
        wg_stream<16>(nc * N, [&](int e) { return blk[e]; },
                  [&](int e, double v) { s.panel[e] = has_pre ? v * prel[e - dN.div(e) * N] : v; });
        __syncthreads();
        cb_pass_lds_left(s.panel, N, nc, side, P, s1);
        cb_pass_lds_left(s.panel, N, nc, side, P, s2);
        for (int e = tid(); e < nc * N; e += NT) {
            const double v = s.panel[e];
            blk[e] = has_post ? v * postl[e - dN.div(e) * N] : v;
        }
        __syncthreads();
    }
}
// M <- (M .* pre) E_s1 E_s2 .* post  (column scalings)
__device__ __noinline__ void wg_cb_right_fused(double* M, int N, int side, const AS_C double* P, int s1, int s2, const double* pre,
                                  const double* post, Lds&) {
    Lds& s = g_lds;
    N = uni(N);
    AS_G double* Mg = gptr(uni(M));
    // scalings live in LDS; test the generic pointers for null (a null generic
    // pointer cast to the LDS address space is not the LDS null)
    const bool has_pre = uni(pre) != nullptr, has_post = uni(post) != nullptr;
    const AS_L double* prel = (const AS_L double*)uni(pre);
    const AS_L double* postl = (const AS_L double*)uni(post);
    const int rb = max(1, min(N, PANEL_DOUBLES / N));
    for (int r0 = 0; r0 < N; r0 += rb) {
        const int nr = min(rb, N - r0);
        const FastDiv dnr(nr);
        wg_stream<16>(nr * N, [&](int e) { const int c = dnr.div(e), r = e - c * nr; return Mg[(size_t)c * N + r0 + r]; },
                  [&](int e, double v) { s.panel[e] = has_pre ? v * prel[dnr.div(e)] : v; });
        __syncthreads();
        cb_pass_lds_right(s.panel, nr, side, P, s1);
        cb_pass_lds_right(s.panel, nr, side, P, s2);
        for (int e = tid(); e < nr * N; e += NT) {
            const int c = dnr.div(e), r = e - c * nr;
            const double v = s.panel[e];
            Mg[(size_t)c * N + r0 + r] = has_post ? v * postl[c] : v;
        }
        __syncthreads();
    }
}
// whole wrap in LDS when the matrix fits: up  G <- B G B^-1, down  G <- B^-1 G B
__device__ __noinline__ void wg_cb_wrap_lds(Desc d, bool up, double* G_, const double* ep_, const double* en_,
                                            Lds&) {
    Lds& s = g_lds;
    const int N = uni(d.N), side = uni(d.side);
    // the scalings live in LDS and G in global memory (no generic accesses)
    const AS_L double* const ep = (const AS_L double*)uni(ep_);
    const AS_L double* const en = (const AS_L double*)uni(en_);
    AS_G double* const G = gptr(uni(G_));
    wg_stream(N * N, [&](int e) { return G[e]; }, [&](int e, double v) { s.panel[e] = v; });
    __syncthreads();
    const FastDiv dN(N);
    if (up) {
        // B_left: E (B then A), rows * ep;  invB_right: E^-1 (B then A) on columns, columns * en
        cb_pass_lds_left(s.panel, N, N, side, d.P, 1);
        cb_pass_lds_left(s.panel, N, N, side, d.P, 0);
        for (int e = tid(); e < N * N; e += NT) s.panel[e] *= ep[e - dN.div(e) * N];
        __syncthreads();
        cb_pass_lds_right(s.panel, N, side, d.Pi, 1);
        cb_pass_lds_right(s.panel, N, side, d.Pi, 0);
        for (int e = tid(); e < N * N; e += NT) G[e] = s.panel[e] * en[dN.div(e)];
    } else {
        // B_right: columns * ep, E (A then B);  invB_left: rows * en, E^-1 (A then B)
        for (int e = tid(); e < N * N; e += NT) s.panel[e] *= ep[dN.div(e)];
        __syncthreads();
        cb_pass_lds_right(s.panel, N, side, d.P, 0);
        cb_pass_lds_right(s.panel, N, side, d.P, 1);
        for (int e = tid(); e < N * N; e += NT) s.panel[e] *= en[e - dN.div(e) * N];
        __syncthreads();
        cb_pass_lds_left(s.panel, N, N, side, d.Pi, 0);
        cb_pass_lds_left(s.panel, N, N, side, d.Pi, 1);
        for (int e = tid(); e < N * N; e += NT) G[e] = s.panel[e];
    }
    __syncthreads();
}

// Field-row scalings exp(+alpha s_i) / exp(-alpha s_i) into LDS.
__device__ void load_field_row(Desc d, int e, Lds&) {
    Lds& s = g_lds;
    for (int i = tid(); i < d.N; i += NT) {
        const int f = ((const AS_G int8_t*)d.field)[(size_t)e * d.N + i];
        s.frow[i] = f;
        s.scp[i] = f > 0 ? d.ev[0] : d.ev[1];
        s.scm[i] = f > 0 ? d.ev[1] : d.ev[0];
    }
    __syncthreads();
}
// after Metropolis flipped entries of the LDS field row
__device__ void refresh_scalings(Desc d, Lds&) {
    Lds& s = g_lds;
    for (int i = tid(); i < d.N; i += NT) {
        const int f = s.frow[i];
        s.scp[i] = f > 0 ? d.ev[0] : d.ev[1];
        s.scm[i] = f > 0 ? d.ev[1] : d.ev[0];
    }
    __syncthreads();
}
// exp(+sigma_eff alpha s) for spin index sp (0 up, 1 down); attractive ignores the spin
__device__ __forceinline__ const double* vpos(Desc d, int sp, Lds&) {
    Lds& s = g_lds;
    return (d.model == 1 || sp == 0) ? s.scp : s.scm;
}
__device__ __forceinline__ const double* vneg(Desc d, int sp, Lds&) {
    Lds& s = g_lds;
    return (d.model == 1 || sp == 0) ? s.scm : s.scp;
}

// B(t) products on an N x N matrix M (in place; W is N*N scratch for the dense path).
// op: 0 B_left, 1 B_right, 2 invB_left, 3 invB_right, 4 transB_left
// (wrap: the product propagates G, which selects its arithmetic in the FP32 modes)
// ---------------------------------------------------------------------------
// Dense B products in Kronecker form.  On the periodic square lattice the
// hopping matrix is a sum over the two directions, K = Kx (+) Ky, so
// F = exp(-s K) factors: F[(x,y),(x',y')] = X[x][x'] Y[y][y'] (site x + side y; the
// chemical potential is a scalar that the host folds into X).  A column v of M,
// read as the side x side matrix V[x'][y'] = v[x' + side y'], maps to
// (F v) = X V Y^T: two 16 x 16 x 16 products (8 MFMAs, factors zero padded to 16)
// instead of a length-N dot product per element, which cuts an N x N x N
// product by a factor N / (2 side) in work.  The host takes X and Y from the
// chain's exp(K) itself and uses this form only when X (x) Y reproduces every
// entry of it within 1e-13 of its largest, so the products equal the dense
// ones to rounding.
// MFMA layouts (16x16x4 FP64): A operand lane l holds A[l%16][4s + l/16], B operand
// B[4s + l/16][l%16], result element e of lane l is D[l/16 + 4e][l%16], so a result
// is the B operand of the next product directly (k step s = element s).
// ---------------------------------------------------------------------------
// fragment of a padded 16 x 16 factor for depth step s: as an A operand (row l%16, depth 4s + l/16)
// when tr is false, or of its transpose; the same index gives the B operand of the transpose
__device__ __forceinline__ double kron_frag(const AS_G double* F, int s, bool tr) {
    const int l = lane();
    return tr ? F[(l & 15) * 16 + 4 * s + (l >> 4)] : F[(4 * s + (l >> 4)) * 16 + (l & 15)];
}

// W = diag(rpost) op(F) diag(rpre) M, op(F) = X (x) Y or its transpose; one wavefront per column
// (columns w, w + 4, ...; four of them in flight together).  rpre / rpost live in LDS or are null.
__device__ __noinline__ void kron_left(const double* M_, double* W_, int N, int side, const double* X_,
                                       const double* Y_, bool tr, const double* rpre_, const double* rpost_) {
    N = uni(N); side = uni(side); tr = uni(tr);
    const AS_G double* const M = gptr(uni(M_));
    AS_G double* const W = gptr(uni(W_));
    const AS_G double* const X = gptr(uni(X_));
    const AS_G double* const Y = gptr(uni(Y_));
    rpre_ = uni(rpre_); rpost_ = uni(rpost_);
    const bool has_pre = rpre_ != nullptr, has_post = rpost_ != nullptr;
    const AS_L double* const rpre = (const AS_L double*)rpre_;
    const AS_L double* const rpost = (const AS_L double*)rpost_;
    const int l = lane(), w = wave();
    const int i = l & 15, kq = l >> 4;
    // step 1, U = V Y'^T: B operand (Y'^T)[4s + l/16][l%16] = Y'[l%16][4s + l/16]; step 2, W = X' U: A operand X'
    double xf[4], yf[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) { xf[s4] = kron_frag(X, s4, tr); yf[s4] = kron_frag(Y, s4, tr); }
    // this lane's element of V for depth step s: V[i][4s + kq] = v[i + side (4s + kq)]
    int vo[4];
    bool vok[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
        vok[s4] = i < side && 4 * s4 + kq < side;
        vo[s4] = vok[s4] ? i + side * (4 * s4 + kq) : 0;
    }
    double vpre[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) vpre[s4] = has_pre ? rpre[vo[s4]] : 1.0;
    // this lane's output sites: W[x = kq + 4e][y = i] at x + side y
    int wo[4];
    bool wok[4];
    double wpost[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int x = kq + 4 * e;
        wok[e] = x < side && i < side;
        wo[e] = wok[e] ? x + side * i : 0;
        wpost[e] = has_post ? rpost[wo[e]] : 1.0;
    }
    constexpr int CH = 4;
    double v[CH][4], vn[CH][4];
    // the loads of the next group of CH columns are issued before this group's products
    auto load_group = [&](int j0, double (&dst)[CH][4]) {
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = min(j0 + 4 * c, N - 1);   // (a column past N repeats the last; not stored)
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) dst[c][s4] = M[(size_t)j * N + vo[s4]];
        }
    };
    load_group(w, vn);
    for (int j0 = w; j0 < N; j0 += 4 * CH) {
#pragma unroll
        for (int c = 0; c < CH; ++c)
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) v[c][s4] = vn[c][s4];
        if (j0 + 4 * CH < N) load_group(j0 + 4 * CH, vn);
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            dbl4 u = dbl4{0.0, 0.0, 0.0, 0.0}, r = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const double a = vok[s4] ? (has_pre ? v[c][s4] * vpre[s4] : v[c][s4]) : 0.0;
                u = __builtin_amdgcn_mfma_f64_16x16x4f64(a, yf[s4], u, 0, 0, 0);
            }
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) r = __builtin_amdgcn_mfma_f64_16x16x4f64(xf[s4], u[s4], r, 0, 0, 0);
            const int j = j0 + 4 * c;
            if (j < N) {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (wok[e]) W[(size_t)j * N + wo[e]] = has_post ? r[e] * wpost[e] : r[e];
            }
        }
    }
    __syncthreads();
}

// W = M diag(cpre) (X (x) Y) diag(cpost), by blocks of 16 rows of M: T = M_R (X (x) 1) from global
// memory into LDS (wavefront w takes the column groups y = w, w + 4, w + 8, w + 12), then
// W_R = T (1 (x) Y) (wavefront w takes x' = w, w + 4, ...) into LDS, written out along columns.
// The global loads of a block (16 per lane) are issued together, and those of the next block
// while the current one runs its second pass and write-out.  Row r of a block sits at stride
// LDR = 17 doubles in LDS (16 rows plus one, so a wavefront's lanes that read one row at
// different columns fall on different banks).
#define KRON_LDR 17
__device__ __noinline__ void kron_right(const double* M_, double* W_, int N, int side, const double* X_,
                                        const double* Y_, const double* cpre_, const double* cpost_, Lds&) {
    Lds& s = g_lds;
    N = uni(N); side = uni(side);
    const AS_G double* const M = gptr(uni(M_));
    AS_G double* const W = gptr(uni(W_));
    const AS_G double* const X = gptr(uni(X_));
    const AS_G double* const Y = gptr(uni(Y_));
    cpre_ = uni(cpre_); cpost_ = uni(cpost_);
    const bool has_pre = cpre_ != nullptr, has_post = cpost_ != nullptr;
    const AS_L double* const cpre = (const AS_L double*)cpre_;
    const AS_L double* const cpost = (const AS_L double*)cpost_;
    AS_L double* const T = (AS_L double*)s.panel;            // 16 rows x N: column c at c LDR
    // the pass-2 output, 16 rows x N: column c = x' + side y' at y' bw + x' LDR, with the stride bw
    // between the side-column groups 4 mod 32 doubles (a lane group stores 16 columns side apart)
    const int bw = side * KRON_LDR + ((4 - side * KRON_LDR) & 31);
    AS_L double* const Wl = T + (size_t)N * KRON_LDR;
    const FastDiv dside(side);
    const int l = lane(), w = wave();
    const int i = l & 15, kq = l >> 4;
    // B operands: X[4s + l/16][l%16] (pass 1) and Y[4s + l/16][l%16] (pass 2)
    double xb[4], yb[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) { xb[s4] = kron_frag(X, s4, true); yb[s4] = kron_frag(Y, s4, true); }
    const bool dok[4] = {4 * 0 + kq < side, 4 * 1 + kq < side, 4 * 2 + kq < side, 4 * 3 + kq < side};
    // this lane's pass-1 elements: column group y = w + 4 g, depth step s: column 4s + kq + side y
    int co[4][4];
    bool cok[4][4];
    double cp[4][4];
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const int y = w + 4 * g;
            cok[g][s4] = y < side && dok[s4];
            co[g][s4] = cok[g][s4] ? 4 * s4 + kq + side * y : 0;
            cp[g][s4] = has_pre ? cpre[co[g][s4]] : 1.0;
        }
    double v[4][4];
    auto load_block = [&](int r0) {
        const int row = min(r0 + i, N - 1);   // (rows past N repeat the last; not stored)
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) v[g][s4] = M[(size_t)co[g][s4] * N + row];
    };
    load_block(0);
    for (int r0 = 0; r0 < N; r0 += 16) {
        // pass 1: T[r][(x', y)] = sum_x M[r][(x, y)] cpre[(x, y)] X[x][x']
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int y = w + 4 * g;
            if (y < side) {
                dbl4 t = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) {
                    const double a = cok[g][s4] ? (has_pre ? v[g][s4] * cp[g][s4] : v[g][s4]) : 0.0;
                    t = __builtin_amdgcn_mfma_f64_16x16x4f64(a, xb[s4], t, 0, 0, 0);
                }
                // element e: T[r0 + kq + 4e][(x' = i, y)]
                if (i < side) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) T[(i + side * y) * KRON_LDR + kq + 4 * e] = t[e];
                }
            }
        }
        if (r0 + 16 < N) load_block(r0 + 16);   // the next block's loads run under pass 2
        __syncthreads();
        // pass 2: W[r][(x', y')] = sum_y T[r][(x', y)] Y[y][y'] cpost[(x', y')]
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int xp = w + 4 * g;
            if (xp < side) {
                double a[4];
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) a[s4] = dok[s4] ? T[(xp + side * (4 * s4 + kq)) * KRON_LDR + i] : 0.0;
                dbl4 r = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) r = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s4], yb[s4], r, 0, 0, 0);
                // element e: W[r0 + kq + 4e][(x', y' = i)]
                if (i < side) {
                    const int c = xp + side * i;
                    const double cq = has_post ? cpost[c] : 1.0;
#pragma unroll
                    for (int e = 0; e < 4; ++e) Wl[i * bw + xp * KRON_LDR + kq + 4 * e] = has_post ? r[e] * cq : r[e];
                }
            }
        }
        __syncthreads();
        const int nr = min(16, N - r0);
        for (int q = tid(); q < 16 * N; q += NT) {
            const int c = q >> 4, rr = q & 15, yq = dside.div(c);
            if (rr < nr) W[(size_t)c * N + r0 + rr] = Wl[yq * bw + (c - yq * side) * KRON_LDR + rr];
        }
    }
    __syncthreads();
}

__device__ void wg_bmul(Desc d, int op, int sp, double* M, double* W, Lds&, bool wrap = false) {
    Lds& s = g_lds;
    const int N = d.N;
    const double* ep = vpos(d, sp, s);
    const double* en = vneg(d, sp, s);
    if (d.cb) {
        switch (op) {
            case 0: wg_cb_left_fused(M, N, d.side, d.P, 1, 0, nullptr, ep, s); break;
            case 1: wg_cb_right_fused(M, N, d.side, d.P, 0, 1, ep, nullptr, s); break;
            case 2: wg_cb_left_fused(M, N, d.side, d.Pi, 0, 1, en, nullptr, s); break;
            case 3: wg_cb_right_fused(M, N, d.side, d.Pi, 1, 0, nullptr, en, s); break;
            case 4: wg_cb_left_fused(M, N, d.side, d.P, 0, 1, ep, nullptr, s); break;
        }
        return;
    }
    if (d.kron && !lowp_product(d, wrap)) {
        // X, Y of exp(K) at kron, of exp(K)^-1 at kron + 512
        const double* kx = d.kron;
        const double* ky = d.kron + 256;
        const double* ix = d.kron + 512;
        const double* iy = d.kron + 768;
        // in place: a column (left) or block of rows (right) is read whole before it is written,
        // and the loads issued ahead read only columns or rows not yet written
        switch (op) {
            case 0: kron_left(M, M, N, d.side, kx, ky, false, nullptr, ep); break;
            case 1: kron_right(M, M, N, d.side, kx, ky, ep, nullptr, s); break;
            case 2: kron_left(M, M, N, d.side, ix, iy, false, en, nullptr); break;
            case 3: kron_right(M, M, N, d.side, ix, iy, nullptr, en, s); break;
            case 4: kron_left(M, M, N, d.side, kx, ky, true, ep, nullptr); break;
        }
        return;
    }
    switch (op) {
        case 0: gemmB(d, wrap, N, N, N, d.expK, N, false, M, N, false, W, N, ep, nullptr, nullptr, s); break;
        case 1: gemmB(d, wrap, N, N, N, M, N, false, d.expK, N, false, W, N, nullptr, ep, nullptr, s); break;
        case 2: gemmB(d, wrap, N, N, N, d.invExpK, N, false, M, N, false, W, N, nullptr, en, nullptr, s); break;
        case 3: gemmB(d, wrap, N, N, N, M, N, false, d.invExpK, N, false, W, N, nullptr, nullptr, en, s); break;
        case 4: gemmB(d, wrap, N, N, N, d.expK, N, true, M, N, false, W, N, nullptr, ep, nullptr, s); break;
    }
    wg_copy(M, W, N * N);
}

// G <- B(t) G B(t)^-1 (up) or G <- B(t)^-1 G B(t) (down) for spin index sp
__device__ void wg_wrap(Desc d, int sp, bool up, double* G, double* W, Lds&) {
    Lds& s = g_lds;
    const int N = d.N;
    if (!d.cb) {
        const double* ep = vpos(d, sp, s);
        const double* en = vneg(d, sp, s);
        if (d.kron && !lowp_product(d, true)) {
            const double* kx = d.kron;
            const double* ky = d.kron + 256;
            const double* ix = d.kron + 512;
            const double* iy = d.kron + 768;
            if (up) {   // G <- diag(ep) F G F^-1 diag(en)
                kron_left(G, W, N, d.side, kx, ky, false, nullptr, ep);
                kron_right(W, G, N, d.side, ix, iy, nullptr, en, s);
            } else {    // G <- F^-1 diag(en) G diag(ep) F
                kron_right(G, W, N, d.side, kx, ky, ep, nullptr, s);
                kron_left(W, G, N, d.side, ix, iy, false, en, nullptr);
            }
            return;
        }
        if (up) {
            gemmB(d, true, N, N, N, d.expK, N, false, G, N, false, W, N, ep, nullptr, nullptr, s);
            gemmB(d, true, N, N, N, W, N, false, d.invExpK, N, false, G, N, nullptr, nullptr, en, s);
        } else {
            gemmB(d, true, N, N, N, G, N, false, d.expK, N, false, W, N, nullptr, ep, nullptr, s);
            gemmB(d, true, N, N, N, d.invExpK, N, false, W, N, false, G, N, nullptr, en, nullptr, s);
        }
        return;
    }
    if (N * N <= PANEL_DOUBLES) { wg_cb_wrap_lds(d, up, G, vpos(d, sp, s), vneg(d, sp, s), s); return; }
    if (up) { wg_bmul(d, 0, sp, G, W, s, true); wg_bmul(d, 3, sp, G, W, s, true); }
    else { wg_bmul(d, 1, sp, G, W, s, true); wg_bmul(d, 2, sp, G, W, s, true); }
}

// ---------------------------------------------------------------------------
// LU with partial pivoting of the m x jb panel staged in LDS (s.panel, ld m),
// in place (unit-lower L below the diagonal, U on and above it).  Thread t keeps
// rows t and t + NT in registers.  Per column one workgroup barrier: every
// wavefront publishes its best pivot candidate (largest |value| among rows >= j,
// smallest row on ties) together with that row, and the owner of row j
// publishes row j, so after the barrier every thread knows the pivot row and
// does the interchange and the elimination of its own rows.  The pivot of step
// j is stored as s.perm[k0 + j] = k0 + p.  Returns the sign the interchanges
// and pivots contribute to det.
// ---------------------------------------------------------------------------
// One wavefront factors a panel of m <= 64 RPL rows: lane L holds rows L + 64 i (i < RPL) in
// registers, the pivot is a wavefront argmax (DPP and row swaps) and rows j and p are read out of
// their lanes into uniform registers, so a column step needs no barrier at all.  Same
// operations in the same order as the four-wavefront form below (bit-identical factors).
template <int RPL>
__device__ __forceinline__ double lu_panel_wave(int m, int jb, int k0) {
    Lds& s = g_lds;
    const int L = lane();
    double* const P = s.panel;
    double x[RPL][NB];
#pragma unroll
    for (int i = 0; i < RPL; ++i)
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const int r = L + 64 * i;
            x[i][c] = (r < m && c < jb) ? P[c * m + r] : 0.0;
        }
    double detsign = 1.0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < jb) {
        double bv = -1.0;
        int bi = j;
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const int r = L + 64 * i;
            if (r >= j && r < m) {
                const double v = fabs(x[i][j]);
                if (v > bv || (v == bv && r < bi)) { bv = v; bi = r; }
            }
        }
        wave_argmax(bv, bi);
        const int p = __builtin_amdgcn_readfirstlane(bi);
        const int pl = p & 63, ps = p >> 6;
        if (L == 0) s.perm[k0 + j] = k0 + p;
        // row p (lane pl, register row ps) read out lane by lane into uniform registers, and row j
        // (lane j, register row 0) only for an interchange
        double rowp[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            double v = x[0][c];
#pragma unroll
            for (int i = 1; i < RPL; ++i) v = (ps == i) ? x[i][c] : v;
            rowp[c] = readlane_d(v, pl);
        }
        if (p != j) {
            detsign = -detsign;
            double rowj[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) rowj[c] = readlane_d(x[0][c], j);
            if (L == pl)
#pragma unroll
                for (int i = 0; i < RPL; ++i)
                    if (ps == i)
#pragma unroll
                        for (int c = 0; c < NB; ++c) x[i][c] = rowj[c];
            if (L == j)
#pragma unroll
                for (int c = 0; c < NB; ++c) x[0][c] = rowp[c];
        }
        const double piv = rowp[j];
        if (piv < 0) detsign = -detsign;
        if (bad_pivot(piv) && L == 0) s.status |= DQMC_STATUS_SINGULAR;
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const int r = L + 64 * i;
            if (r > j && r < m) {
                const double l = x[i][j] / piv;
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    if (c > j) x[i][c] -= l * rowp[c];
                x[i][j] = l;
            }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
        const int r = L + 64 * i;
        if (r >= m) continue;
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < jb) P[c * m + r] = x[i][c];
    }
    return detsign;
}

__device__ __noinline__ double lu_panel(int m, int jb, int k0, Lds&) {
    Lds& s = g_lds;
    m = uni(m); jb = uni(jb); k0 = uni(k0);
    if (m <= 256) {
        double* const dsl = s.panel + PANEL_RED + RED_CVAL;
        if (wave() == 0) {
            double ds;
            if (m <= 64) ds = lu_panel_wave<1>(m, jb, k0);
            else if (m <= 128) ds = lu_panel_wave<2>(m, jb, k0);
            else if (m <= 192) ds = lu_panel_wave<3>(m, jb, k0);
            else ds = lu_panel_wave<4>(m, jb, k0);
            if (lane() == 0) dsl[0] = ds;
        }
        __syncthreads();
        const double ds = dsl[0];
        __syncthreads();
        return ds;
    }
    const int t = tid(), w = wave();
    double* const P = s.panel;
    double* const red = s.panel + PANEL_RED;
    double x[2][NB];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const int r = t + i * NT;
            x[i][c] = (r < m && c < jb) ? P[c * m + r] : 0.0;
        }
    double detsign = 1.0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < jb) {   // (a guard, not a break, so that the loop unrolls and x stays in registers)
        const int par = j & 1;
        double bv = -1.0;
        int bi = j;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = t + i * NT;
            if (r >= j && r < m) {
                const double v = fabs(x[i][j]);
                if (v > bv || (v == bv && r < bi)) { bv = v; bi = r; }
            }
        }
        wave_argmax(bv, bi);
        if (bi % NT == t) {   // this thread holds the wavefront's candidate row
            const int slot = bi / NT;
#pragma unroll
            for (int c = 0; c < NB; ++c) red[RED_CAND + (par * 4 + w) * NB + c] = slot ? x[1][c] : x[0][c];
        }
        if (lane() == 0) {
            red[RED_CVAL + (par * 2 + 0) * 4 + w] = bv;
            red[RED_CVAL + (par * 2 + 1) * 4 + w] = (double)bi;
        }
        if (t == j)
#pragma unroll
            for (int c = 0; c < NB; ++c) red[RED_ROWJ + par * NB + c] = x[0][c];
        __syncthreads();
        // the pivot: best of the 4 candidates (ties: smaller row, then lower wavefront)
        int wb = 0;
        double vb = red[RED_CVAL + (par * 2 + 0) * 4 + 0];
        int pb = (int)red[RED_CVAL + (par * 2 + 1) * 4 + 0];
#pragma unroll
        for (int q = 1; q < 4; ++q) {
            const double v = red[RED_CVAL + (par * 2 + 0) * 4 + q];
            const int pr = (int)red[RED_CVAL + (par * 2 + 1) * 4 + q];
            if (v > vb || (v == vb && pr < pb)) { vb = v; pb = pr; wb = q; }
        }
        const int p = pb;
        if (t == 0) s.perm[k0 + j] = k0 + p;
        double rowp[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) rowp[c] = red[RED_CAND + (par * 4 + wb) * NB + c];
        if (p != j) {
            detsign = -detsign;
            // interchange rows j and p (both from the published copies)
            if (t == p % NT) {
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    const double v = red[RED_ROWJ + par * NB + c];
                    if (p / NT) x[1][c] = v; else x[0][c] = v;
                }
            }
            if (t == j)
#pragma unroll
                for (int c = 0; c < NB; ++c) x[0][c] = rowp[c];
        }
        const double piv = rowp[j];
        if (piv < 0) detsign = -detsign;
        if (bad_pivot(piv) && t == 0) s.status |= DQMC_STATUS_SINGULAR;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = t + i * NT;
            if (r > j && r < m) {
                const double l = x[i][j] / piv;
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    if (c > j) x[i][c] -= l * rowp[c];
                x[i][j] = l;
            }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int r = t + i * NT;
        if (r >= m) continue;
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < jb) P[c * m + r] = x[i][c];
    }
    __syncthreads();
    return detsign;
}

// ---------------------------------------------------------------------------
// C +/-= A W with both operands in LDS: C is M x Nc in global memory (ld ldc); A is M x K4 with
// element (r, k) at A[k lda + r]; W is K4 x Nc with element (k, c) at W[k ldw + c]; K4 is the
// padded depth (a multiple of 4; both operands zero beyond the true depth).  Wavefronts
// wave_lo..3 take 2x2 groups of 16x16 blocks of C (block rows 2gr, 2gr+1, block columns 2gc,
// 2gc+1) in turn or, with `next`, as an LDS counter hands them out to whichever wavefront asks
// (a wavefront that joins late takes what is left).  No workgroup barrier inside.  The product
// is formed transposed (W^T A^T), so 16 lanes run down one column of C; each depth step reads
// two A and two W operands for four MFMAs with the depth loop unrolled, and the C loads of a
// wavefront's next group are issued before the current group's products, so their latency runs
// under the matrix cores.  Each block sums its K4/4 MFMAs in depth order whatever the grouping.
// ---------------------------------------------------------------------------
template <int K4, bool SUB>
__device__ __forceinline__ void lds_update(AS_G double* C, int ldc, int M, int Nc, const AS_L double* A, int lda,
                                           const AS_L double* W, int ldw, int wave_lo, AS_L int* next) {
    const int l = lane(), w = wave();
    const int nw = 4 - wave_lo, wi = w - wave_lo;
    const int lr = l & 15, lk = l >> 4;
    const int ngr = (M + 31) >> 5, ng = ngr * ((Nc + 31) >> 5);
    const FastDiv dg(ngr);
    auto claim = [&]() {
        int g = 0;
        if (l == 0) g = __hip_atomic_fetch_add(next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return __builtin_amdgcn_readfirstlane(g);
    };
    double cvn[2][2][4];
    auto load_group = [&](int g) {
        const int gc = dg.div(g), gr = g - gc * ngr;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int orow = min((2 * gr + a) * 16 + lr, M - 1);
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    cvn[a][b][i] = C[(size_t)min((2 * gc + b) * 16 + lk + 4 * i, Nc - 1) * ldc + orow];
        }
    };
    int g = next ? claim() : wi;
    if (g < ng) load_group(g);
    while (g < ng) {
        const int gc = dg.div(g), gr = g - gc * ngr;
        double cv[2][2][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i) cv[a][b][i] = cvn[a][b][i];
        const int gn = next ? claim() : g + nw;
        if (gn < ng) load_group(gn);
        int oa[2], ob[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            oa[a] = min((2 * gr + a) * 16 + lr, M - 1);   // (rows and columns past the end repeat the last; not stored)
            ob[a] = min((2 * gc + a) * 16 + lr, Nc - 1);
        }
        dbl4 acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
        // (operands double-buffered: the next depth step's LDS reads are issued ahead of the
        // current step's MFMAs, so their latency runs under the matrix cores)
        double ao[2][2], bo[2][2];
        ao[0][0] = A[lk * lda + oa[0]]; ao[0][1] = A[lk * lda + oa[1]];
        bo[0][0] = W[lk * ldw + ob[0]]; bo[0][1] = W[lk * ldw + ob[1]];
#pragma unroll
        for (int kk = 0; kk < K4; kk += 4) {
            const int cur = (kk >> 2) & 1;
            if (kk + 4 < K4) {
                const int ka = (kk + 4 + lk) * lda, kw = (kk + 4 + lk) * ldw;
                ao[cur ^ 1][0] = A[ka + oa[0]]; ao[cur ^ 1][1] = A[ka + oa[1]];
                bo[cur ^ 1][0] = W[kw + ob[0]]; bo[cur ^ 1][1] = W[kw + ob[1]];
            }
            __builtin_amdgcn_sched_barrier(0);   // (keeps the reads ahead of the MFMAs)
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(bo[cur][0], ao[cur][0], acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(bo[cur][1], ao[cur][0], acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(bo[cur][0], ao[cur][1], acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(bo[cur][1], ao[cur][1], acc[1][1], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int row = (2 * gr + a) * 16 + lr;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int col = (2 * gc + b) * 16 + lk + 4 * i;
                    if (row < M && col < Nc) C[(size_t)col * ldc + row] = SUB ? cv[a][b][i] - acc[a][b][i] : cv[a][b][i] + acc[a][b][i];
                }
            }
        g = gn;
    }
}

// ---------------------------------------------------------------------------
// Blocked LU with partial pivoting: A X = B, X overwrites B (N x nrhs; B must
// follow A in memory, B = A + N N, when nrhs > 0); A is destroyed (only the
// diagonal blocks of U and the U12 rows are kept: L is not read again).  Panels
// of NB columns are factored in LDS (pivot search, swaps and rank-1 updates never
// leave the CU); the row interchanges and L11^-1 are applied to the columns right
// of the panel and to B one column per thread, which also leaves the panel's U12
// rows in LDS; the rows below are then updated, C -= L21 U12, by lds_update with
// both operands in LDS (right-looking, Golub & Van Loan Alg. 3.4.2).  Back
// substitution runs by NB-row blocks: each block's rows of X are solved one column
// per thread against U11 and left in LDS, and the rows above are updated the same
// way with U's column block.  Columns go in chunks that fit the LDS panel.
// Returns the sign of det(A).
// ---------------------------------------------------------------------------
__device__ __noinline__ double wg_lu_solve(double* A, double* B, int N, int nrhs, Lds&) {
    Lds& s = g_lds;
    double detsign = 1.0;
    double* P = s.panel;                      // m x jb, ld = m
    double* T11 = s.panel + NMAX * NB;        // NB x NB
    N = uni(N);
    nrhs = uni(nrhs);
    A = uni(A);
    const AS_G double* Ag = gptr(A);
    const int ncol = N + nrhs;                // columns of [A B]
    for (int k0 = 0; k0 < N; k0 += NB) {
        const int jb = min(NB, N - k0), m = N - k0;
        const FastDiv dm(m);
        SubClock pclk;
        pclk.start();
        wg_stream(m * jb, [&](int e) { const int c = dm.div(e), r = e - c * m; return Ag[(size_t)(k0 + c) * N + k0 + r]; },
                  [&](int e, double v) { P[e] = v; });
        __syncthreads();
        detsign *= lu_panel(m, jb, k0, s);
        pclk.stop(1);
        {   // U11 back to A (one element per thread)
            const int r = tid() & (NB - 1), c = tid() >> 4;
            if (r < jb && c < jb) gptr(A)[(size_t)(k0 + c) * N + k0 + r] = P[c * m + r];
        }
        // the panel's interchanges (row k0+j <-> perm[k0+j], j = 0..jb-1 in order) composed into
        // one gather: row swp_row[i] receives the old row swp_src[i]; rows k0..k0+jb-1 come first
        // (wavefront 0: lane i tracks entry i; a ballot finds the lane holding the pivot row)
        if (wave() == 0) {
            const int i = lane();
            int row = i < jb ? k0 + i : -1, src = row;
            int nrow = jb;
            for (int j = 0; j < jb; ++j) {
                const int pr = __builtin_amdgcn_readfirstlane(s.perm[k0 + j]);
                const unsigned long long hit = __ballot(row == pr);
                int ip;
                if (hit) ip = __ffsll((long long)hit) - 1;
                else {
                    ip = nrow++;
                    if (i == ip) { row = pr; src = pr; }
                }
                const int sj = __builtin_amdgcn_readlane(src, j), sp = __builtin_amdgcn_readlane(src, ip);
                if (i == j) src = sp;
                if (i == ip) src = sj;
            }
            if (i < 2 * NB) {
                s.swp_row[i] = i < nrow ? row : -1;
                s.swp_src[i] = src;
            }
        }
        // L11^-1 (unit lower, zero outside jb x jb) into Li, row-major 16 x 16: lane j of wavefront
        // 0 forms column j by forward substitution
        double* const Li = P + (size_t)m * NB;
        // (the sums run over every q < i: y[q] = 0 for q < j adds exact zeros, and the loops keep no
        // per-lane branch; rows past jb read past the panel and are discarded)
        if (wave() == 0 && lane() < NB) {
            const int j = lane();
            double y[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                double v = i == j ? 1.0 : 0.0;
#pragma unroll
                for (int q = 0; q < i; ++q) v -= P[q * m + i] * y[q];
                y[i] = i < jb ? v : 0.0;
            }
#pragma unroll
            for (int i = 0; i < NB; ++i) Li[i * NB + j] = y[i];
        }
        __syncthreads();
        // columns right of the panel, then B, in chunks of cw: the interchanges and U12 = L11^-1 X
        // by 16-column blocks (lane l holds rows 4s + l/16 of column l%16 of a block: the B operand
        // of the product with L11^-1; each wavefront takes four blocks at a time with all their
        // loads in flight), U12 to A and to Wl, then C -= L21 U12 on the rows below the panel
        const int c0 = k0 + jb, nr = ncol - c0, mt = m - jb;
        double* const Wl = Li + NB * NB;                // NB x cw, ld cw
        const int cwmax = ((PANEL_DOUBLES - m * NB - NB * NB) / NB) & ~31;
        const int lr = lane() & 15, lk = lane() >> 4;
        double la[4];   // L11^-1 as the A operand: Li[l%16][4s + l/16]
#pragma unroll
        for (int q = 0; q < 4; ++q) la[q] = Li[lr * NB + 4 * q + lk];
        int rw[8], rs[8];   // interchange entries i = 4s + l/16 (s < 4) and 16 + 4s + l/16
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = (q < 4 ? 4 * q : 16 + 4 * (q - 4)) + lk;
            rw[q] = s.swp_row[i];
            rs[q] = max(s.swp_src[i], 0);
        }
        for (int q0 = 0; q0 < nr; q0 += cwmax) {
            const int cw = min(cwmax, nr - q0), nbk = (cw + 15) >> 4;
            for (int bb = wave() * 4; bb < nbk; bb += 16) {
                double xv[4][8];
                int cl[4];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    cl[b] = (bb + b) * 16 + lr;
                    const AS_G double* col = gptr(A + (size_t)(c0 + q0 + min(cl[b], cw - 1)) * N);
#pragma unroll
                    for (int q = 0; q < 8; ++q) xv[b][q] = col[rs[q]];
                }
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    if (cl[b] >= cw) continue;
                    AS_G double* col = gptr(A + (size_t)(c0 + q0 + cl[b]) * N);
                    dbl4 u = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int i = (q < 4 ? 4 * q : 16 + 4 * (q - 4)) + lk;
                        if (i >= jb && rw[q] >= 0) col[rw[q]] = xv[b][q];   // rows leaving the panel's rows
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const double x = (4 * q + lk < jb) ? xv[b][q] : 0.0;
                        u = __builtin_amdgcn_mfma_f64_16x16x4f64(la[q], x, u, 0, 0, 0);
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int r = lk + 4 * e;
                        if (r < jb) col[k0 + r] = u[e];
                        Wl[r * cw + cl[b]] = u[e];
                    }
                }
            }
            __syncthreads();
            // (rows below the panel exist only when jb = NB, so the depth is NB exactly)
            if (mt > 0)
                lds_update<NB, true>(gptr(A + (size_t)(c0 + q0) * N + k0 + jb), N, mt, cw, (const AS_L double*)(P + jb), m,
                                     (const AS_L double*)Wl, cw, 0, nullptr);
            __syncthreads();
        }
    }
    if (nrhs == 0) return detsign;
    // back substitution: Ucb = U[0:k0, k0:k0+NB] (element (r, c) at c k0 + r, zero for c >= jb), the
    // block's rows of X in Xl (NB x cw, ld cw, zero rows past jb)
    double* const Ucb = s.panel;
    double* const Xl = T11 + NB * NB;
    const int cwb = ((PANEL_DOUBLES - NMAX * NB - NB * NB) / NB) & ~31;
    B = uni(B);
    for (int k0 = ((N - 1) / NB) * NB; k0 >= 0; k0 -= NB) {
        const int jb = min(NB, N - k0);
        {
            const int r = tid() & (NB - 1), c = tid() >> 4;
            if (r < jb && c < jb) T11[c * NB + r] = Ag[(size_t)(k0 + c) * N + k0 + r];
        }
        if (k0 > 0) {
            const FastDiv dk(k0);
            wg_stream(k0 * NB, [&](int e) { const int c = dk.div(e), r = e - c * k0; return Ag[(size_t)(k0 + min(c, jb - 1)) * N + r]; },
                      [&](int e, double v) { Ucb[e] = dk.div(e) < jb ? v : 0.0; });
        }
        __syncthreads();
        for (int q0 = 0; q0 < nrhs; q0 += cwb) {
            const int cw = min(cwb, nrhs - q0);
            for (int c = tid(); c < cw; c += NT) {
                AS_G double* col = gptr(B + (size_t)(q0 + c) * N);
                double x[NB];
#pragma unroll
                for (int j = 0; j < NB; ++j) x[j] = (j < jb) ? col[k0 + j] : 0.0;
#pragma unroll
                for (int j = NB - 1; j >= 0; --j) {
                    if (j < jb) {
                        x[j] /= T11[j * NB + j];
#pragma unroll
                        for (int i = 0; i < j; ++i) x[i] -= T11[j * NB + i] * x[j];
                    }
                }
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    if (j < jb) col[k0 + j] = x[j];
                    Xl[j * cw + c] = x[j];
                }
            }
            __syncthreads();
            if (k0 > 0)
                lds_update<NB, true>(gptr(B + (size_t)q0 * N), N, k0, cw, (const AS_L double*)Ucb, k0,
                                     (const AS_L double*)Xl, cw, 0, nullptr);
            __syncthreads();
        }
    }
    return detsign;
}

// ---------------------------------------------------------------------------
// Householder QR of the m x jb panel staged in LDS (s.panel, ld m).  Thread t
// keeps rows t and t + NT of the panel in registers, so each column costs one
// workgroup barrier: the reduction of its norm and of its dot products with the
// columns to its right (LAPACK dgeqr2 with the reflector scaled to a unit
// leading element, tau = (alpha - x0) / alpha written as -v0 / alpha).  Then the
// Gram matrix V^T V of the reflectors and the T factor of the block reflector
// I - V T V^T (dlarft, forward columnwise) are formed.  Outputs: R (upper part)
// to Mp and the explicit V (unit lower) to M, both at (k0, k0) with ld N; tau to
// s.v1[k0..]; T to Tb (LDS, ld NB) and Tg.
// ---------------------------------------------------------------------------
// One wavefront factors a panel of m <= 64 RPL rows (lane L holds rows L + 64 i): the column's
// norm and dot products are reduce-scattered within the wavefront and read out of their lanes
// into uniform registers, so a column step needs no barrier at all.  Same Householder steps,
// Gram matrix and T factor as the four-wavefront form below (sums in a different order).
template <int RPL>
__device__ __forceinline__ void qr_panel_wave(int m, int jb, int k0, int N, AS_G double* M, AS_G double* Mp,
                                              AS_G double* Tg) {
    Lds& s = g_lds;
    const int L = lane();
    double* const P = s.panel;
    double* const red = s.panel + PANEL_RED;
    double x[RPL][NB];
#pragma unroll
    for (int i = 0; i < RPL; ++i)
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const int r = L + 64 * i;
            x[i][c] = (r < m && c < jb) ? P[c * m + r] : 0.0;
        }
    double* const tau = s.v1 + k0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < jb) {
        double part[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) part[q] = 0.0;
        // (rows outside r > j enter as zero: the same sums for finite data, and no branches, so
        // the rows' dependence chains can interleave)
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const int r = L + 64 * i;
            const double xj = (r > j && r < m) ? x[i][j] : 0.0;
            part[0] += xj * xj;
#pragma unroll
            for (int q = 1; q < NB; ++q)
                if (j + q < NB) part[q] += xj * x[i][j + q];
        }
        // row j (lane j, register row 0) and the slot totals (slot q on lane 4q) read out of their
        // lanes into uniform registers
        const double wsum = wave_reduce16(part);
        const double x0 = readlane_d(x[0][j], j);
        const double nrm = sqrt(x0 * x0 + readlane_d(wsum, 0));
        const bool nz = nrm != 0.0;
        const double alpha_r = x0 > 0 ? -nrm : nrm, v0 = x0 - alpha_r;
        const double tk_r = -v0 / alpha_r, inv_r = 1.0 / v0;   // (discarded for a zero column)
        const double tk = nz ? tk_r : 0.0, inv_v0 = nz ? inv_r : 1.0, alpha = nz ? alpha_r : x0;
        if (L == 0) tau[j] = tk;
        double wq[NB];
#pragma unroll
        for (int q = 1; q < NB; ++q)
            wq[q] = (j + q < NB) ? tk * (readlane_d(x[0][j + q], j) + inv_v0 * readlane_d(wsum, 4 * q)) : 0.0;
#pragma unroll
        for (int i = 0; i < RPL; ++i) {
            const int r = L + 64 * i;
            const bool act = r > j && r < m && tk != 0.0;
            const double v = x[i][j] * inv_v0, va = act ? v : 0.0;
            x[i][j] = act ? v : x[i][j];
#pragma unroll
            for (int q = 1; q < NB; ++q)
                if (j + q < NB) x[i][j + q] -= wq[q] * va;
        }
        const bool lj = L == j;   // lane j holds row j in x[0]
        x[0][j] = lj ? alpha : x[0][j];
#pragma unroll
        for (int q = 1; q < NB; ++q)
            if (j + q < NB) x[0][j + q] -= lj ? wq[q] : 0.0;
      }
    }
    // R to Mp, explicit V to M
#pragma unroll
    for (int i = 0; i < RPL; ++i) {
        const int r = L + 64 * i;
        if (r >= m) continue;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            if (c < jb) {
                const size_t o = (size_t)(k0 + c) * N + k0 + r;
                Mp[o] = (r <= c) ? x[i][c] : 0.0;
                M[o] = (r == c) ? 1.0 : (r > c ? x[i][c] : 0.0);
            }
        }
    }
    // Gram matrix G(l, j) = V(j, l) + sum_{r > j} V(r, l) V(r, j) (l < j), pair j (j - 1) / 2 + l,
    // in groups of 16 pairs (the pair counter is a constant of the unrolled loops, so every
    // operand is a fixed register)
    double* const gram = red + RED_GRAM;
    double part[16];
#pragma unroll
    for (int j = 1; j < NB; ++j) {
#pragma unroll
        for (int l = 0; l < j; ++l) {
            const int pidx = j * (j - 1) / 2 + l;
            // (pairs with j >= jb are formed too and not read: T takes j < jb only)
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < RPL; ++i) {
                const int r = L + 64 * i;
                const double f = (r < m && r > j) ? x[i][j] : ((r < m && r == j) ? 1.0 : 0.0);
                acc += x[i][l] * f;   // (V(r, l) V(r, j) below the diagonal, V(j, l) on it)
            }
            part[pidx & 15] = acc;
            if ((pidx & 15) == 15 || pidx == NB * (NB - 1) / 2 - 1) {
#pragma unroll
                for (int q = (pidx & 15) + 1; q < 16; ++q) part[q] = 0.0;
                const double wsum = wave_reduce16(part);
                if ((L & 3) == 0) gram[4 * 128 + (pidx & ~15) + ((L >> 2) & 15)] = wsum;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    // T: lane i builds row i, T(i, j) = -tau_j sum_{l=i}^{j-1} T(i, l) G(l, j), T(j, j) = tau_j
    // (the sums start at l = 0: T(i, l) = 0 for l < i adds exact zeros, and the loops keep no
    // per-lane branch)
    double* const Tb = s.panel + PANEL_BLK;
    if (L < NB) {
        const int i = L;
        double trow[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int l = 0; l < j; ++l) acc += trow[l] * gram[4 * 128 + j * (j - 1) / 2 + l];
            const double tj = tau[min(j, jb - 1)];
            trow[j] = (j < jb && i == j) ? tj : ((j < jb && i < j) ? -tj * acc : 0.0);
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            Tb[j * NB + i] = trow[j];
            Tg[(size_t)(k0 / NB) * NB * NB + j * NB + i] = trow[j];
        }
    }
}

__device__ __noinline__ void qr_panel(int m, int jb, int k0, int N, double* M_, double* Mp_, double* Tg_, Lds&) {
    Lds& s = g_lds;
    m = uni(m); jb = uni(jb); k0 = uni(k0); N = uni(N);
    AS_G double* M = gptr(uni(M_));
    AS_G double* Mp = gptr(uni(Mp_));
    AS_G double* Tg = gptr(uni(Tg_));
    if (m <= 256) {
        if (wave() == 0) {
            if (m <= 64) qr_panel_wave<1>(m, jb, k0, N, M, Mp, Tg);
            else if (m <= 128) qr_panel_wave<2>(m, jb, k0, N, M, Mp, Tg);
            else if (m <= 192) qr_panel_wave<3>(m, jb, k0, N, M, Mp, Tg);
            else qr_panel_wave<4>(m, jb, k0, N, M, Mp, Tg);
        }
        __syncthreads();
        return;
    }
    const int t = tid(), w = wave();
    double* const P = s.panel;
    double* const red = s.panel + PANEL_RED;
    double x[2][NB];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const int r = t + i * NT;
            x[i][c] = (r < m && c < jb) ? P[c * m + r] : 0.0;
        }
    double* const tau = s.v1 + k0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < jb) {   // (a guard, not a continue, so that the loop unrolls and x stays in registers)
        const int par = j & 1;
        // partial sums over this thread's rows r > j: slot 0 the squared norm of
        // column j below the diagonal, slot q the dot product with column j + q
        double part[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) part[q] = 0.0;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = t + i * NT;
            if (r > j && r < m) {
                const double xj = x[i][j];
                part[0] += xj * xj;
#pragma unroll
                for (int q = 1; q < NB; ++q)
                    if (j + q < NB) part[q] += xj * x[i][j + q];
            }
        }
        if (t == j)
#pragma unroll
            for (int q = 0; q < NB; ++q) red[RED_ROWJ + par * NB + q] = (j + q < NB) ? x[0][j + q] : 0.0;
        const double wsum = wave_reduce16(part);
        if ((lane() & 3) == 0) red[(par * 4 + w) * NB + ((lane() >> 2) & 15)] = wsum;
        __syncthreads();
        // workgroup totals of slot q: the 4 wavefront sums
#define QR_TOT(q) (red[(par * 4 + 0) * NB + (q)] + red[(par * 4 + 1) * NB + (q)] + \
                   red[(par * 4 + 2) * NB + (q)] + red[(par * 4 + 3) * NB + (q)])
        const double x0 = red[RED_ROWJ + par * NB];
        const double nrm = sqrt(x0 * x0 + QR_TOT(0));
        double tk = 0.0, inv_v0 = 1.0, alpha = x0;
        if (nrm != 0.0) {
            alpha = x0 > 0 ? -nrm : nrm;
            const double v0 = x0 - alpha;
            tk = -v0 / alpha;
            inv_v0 = 1.0 / v0;
        }
        if (t == 0) tau[j] = tk;
        // H_j on the columns right of j: w_q = tau (P(j, j+q) + v^T P(j+1:, j+q))
        double wq[NB];
#pragma unroll
        for (int q = 1; q < NB; ++q) wq[q] = (j + q < NB) ? tk * (red[RED_ROWJ + par * NB + q] + inv_v0 * QR_TOT(q)) : 0.0;
#undef QR_TOT
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = t + i * NT;
            if (r > j && r < m && tk != 0.0) {
                const double v = x[i][j] * inv_v0;
                x[i][j] = v;
#pragma unroll
                for (int q = 1; q < NB; ++q)
                    if (j + q < NB) x[i][j + q] -= wq[q] * v;
            }
        }
        if (t == j) {
            x[0][j] = alpha;
#pragma unroll
            for (int q = 1; q < NB; ++q)
                if (j + q < NB) x[0][j + q] -= wq[q];
        }
      }
    }
    // R to Mp, explicit V to M
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int r = t + i * NT;
        if (r >= m) continue;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            if (c < jb) {
                const size_t o = (size_t)(k0 + c) * N + k0 + r;
                Mp[o] = (r <= c) ? x[i][c] : 0.0;
                M[o] = (r == c) ? 1.0 : (r > c ? x[i][c] : 0.0);
            }
        }
    }
    // Gram matrix G(l, j) = V_l^T V_j (l < j) = V(j, l) + sum_{r > j} V(r, l) V(r, j):
    // 120 pairs in 8 groups of 16, reduce-scattered per wavefront
    double* const gram = red + RED_GRAM;
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        double part[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            part[q] = 0.0;
            const int pidx = g * 16 + q;
            // pair index p = j (j - 1) / 2 + l, 0 <= l < j < NB
            int j = 1;
            while ((j + 1) * j / 2 <= pidx) ++j;
            const int l = pidx - j * (j - 1) / 2;
            if (j < NB && j < jb) {
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const int r = t + i * NT;
                    if (r < m && r > j) part[q] += x[i][l] * x[i][j];
                    else if (r < m && r == j) part[q] += x[i][l];
                }
            }
        }
        const double wsum = wave_reduce16(part);
        if ((lane() & 3) == 0) gram[w * 128 + g * 16 + ((lane() >> 2) & 15)] = wsum;
    }
    __syncthreads();
    if (t < 128) gram[4 * 128 + t] = gram[t] + gram[128 + t] + gram[256 + t] + gram[384 + t];
    __syncthreads();
    // T: lane i of wave 0 builds row i, T(i, j) = -tau_j sum_{l=i}^{j-1} T(i, l) G(l, j), T(j, j) = tau_j
    double* const Tb = s.panel + PANEL_BLK;
    if (w == 0 && lane() < NB) {
        const int i = lane();
        double trow[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            trow[j] = 0.0;
            if (j < jb && i == j) trow[j] = tau[j];
            else if (j < jb && i < j) {
                double acc = 0.0;
#pragma unroll
                for (int l = 0; l < NB; ++l)
                    if (l >= i && l < j) acc += trow[l] * gram[4 * 128 + j * (j - 1) / 2 + l];
                trow[j] = -tau[j] * acc;
            }
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            Tb[j * NB + i] = trow[j];
            Tg[(size_t)(k0 / NB) * NB * NB + j * NB + i] = trow[j];
        }
    }
    __syncthreads();
}

// The same block reflector with C held in registers, for m <= 4 MQ rows: each wavefront takes
// whole 16-column blocks of C (all m rows, so the reduction V^T C stays in the wavefront; no
// workgroup barrier between blocks), reads C once and writes it once.  Lane l holds
// C[4s + l/16][l%16] of its block for depth step s, which is both the B operand of
// W1 = V^T C and, for s = 4 rb + e, element e of the 16x16 result block rb of V W2, so the
// update C - V W2 is elementwise in registers.  V (m x jb) is staged once in the LDS panel
// at an odd leading dimension (rows padded to a multiple of 16 with zeros, columns to 16),
// so that lanes reading one row at 16 different columns fall on different banks.  The instance
// takes MB = ceil(m / 16) 16-row blocks exactly, so its loops carry no branch on the row count.
template <int MB>
__device__ __noinline__ void wg_reflect_regs(const double* V_, int ldv, const double* T_, bool tT, double* C_, int ldc,
                                             int m, int n, int jb) {
    constexpr int MQ = 4 * MB;
    Lds& s = g_lds;
    ldv = uni(ldv); tT = uni(tT); ldc = uni(ldc); m = uni(m); n = uni(n); jb = uni(jb);
    const AS_G double* const V = gptr(uni(V_));
    const AS_G double* const T = gptr(uni(T_));
    AS_G double* const C = gptr(uni(C_));
    constexpr int mp = MB * 16, ldp = mp + 1;
    AS_L double* const Vl = (AS_L double*)s.panel;   // 16 x ldp
    {
        const FastDiv dmp(mp);
        for (int q = tid(); q < 16 * mp; q += NT) {
            const int c = dmp.div(q), r = q - c * mp;
            Vl[c * ldp + r] = (c < jb && r < m) ? V[(size_t)c * ldv + r] : 0.0;
        }
    }
    const int l = lane(), w = wave();
    const int li = l & 15, lk = l >> 4;
    // op(T) as the A operand: op(T)[l%16][4s + l/16] (T is zero outside jb x jb)
    double tf[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) tf[s4] = tT ? T[li * NB + 4 * s4 + lk] : T[(4 * s4 + lk) * NB + li];
    __syncthreads();
    const int ncb = (n + 15) >> 4;
    for (int cb = w; cb < ncb; cb += 4) {
        const int col = cb * 16 + li;
        const bool colok = col < n;
        // (rows below the last 16-row block are all inside m: one base, constant offsets; the last
        // block's rows past m repeat row m - 1.  Nothing is zeroed: a lane's column of W1, W2 and
        // C - V W2 depends on that column of C only, a column past n is not stored, and the rows
        // of V past m are zero, so the repeated rows add nothing)
        AS_G double* const cc = C + (size_t)(colok ? col : 0) * ldc;
        double c[MQ];
#pragma unroll
        for (int q = 0; q < MQ; ++q) {
            const int r = 4 * q + lk;
            c[q] = q < MQ - 4 ? cc[r] : cc[min(r, m - 1)];
        }
        // W1 = V^T C: A operand V^T[l%16][4s + l/16] = V[4s + l/16][l%16] (the operands of the next
        // four steps read under the current four's MFMAs; the fences keep the loads from all being
        // hoisted to the top)
        dbl4 w1 = dbl4{0.0, 0.0, 0.0, 0.0};
        {
            const AS_L double* const vr = lds_row((const double*)(Vl + li * ldp + lk));
            double vo[2][4];
#pragma unroll
            for (int j = 0; j < 4; ++j) vo[0][j] = vr[4 * j];
#pragma unroll
            for (int g = 0; g < MB; ++g) {
                const int cur = g & 1;
                if (g + 1 < MB)
#pragma unroll
                    for (int j = 0; j < 4; ++j) vo[cur ^ 1][j] = vr[4 * (4 * g + 4 + j)];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < 4; ++j) w1 = __builtin_amdgcn_mfma_f64_16x16x4f64(vo[cur][j], c[4 * g + j], w1, 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // W2 = op(T) W1
        dbl4 w2 = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) w2 = __builtin_amdgcn_mfma_f64_16x16x4f64(tf[s4], w1[s4], w2, 0, 0, 0);
        // C -= V W2 by 16-row blocks: A operand V[16 rb + l%16][4s + l/16]
        {
            const AS_L double* const vc = lds_row((const double*)(Vl + lk * ldp + li));
            double vo[2][4];
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) vo[0][s4] = vc[4 * s4 * ldp];
#pragma unroll
            for (int rb = 0; rb < MB; ++rb) {
                const int cur = rb & 1;
                if (rb + 1 < MB)
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) vo[cur ^ 1][s4] = vc[4 * s4 * ldp + 16 * (rb + 1)];
                __builtin_amdgcn_sched_barrier(0);
                dbl4 pv = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) pv = __builtin_amdgcn_mfma_f64_16x16x4f64(vo[cur][s4], w2[s4], pv, 0, 0, 0);
#pragma unroll
                for (int e = 0; e < 4; ++e) c[4 * rb + e] = c[4 * rb + e] - pv[e];
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (colok) {
#pragma unroll
            for (int q = 0; q < MQ; ++q) {
                const int r = 4 * q + lk;
                if (q < MQ - 4 || r < m) cc[r] = c[q];
            }
        }
    }
    __syncthreads();
}
// the instance for m <= 256 rows (16 * MB >= m)
__device__ __forceinline__ void reflect_regs(const double* V, int ldv, const double* T, bool tT, double* C, int ldc, int m,
                                             int n, int jb) {
    switch ((m + 15) >> 4) {
        case 1: wg_reflect_regs<1>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 2: wg_reflect_regs<2>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 3: wg_reflect_regs<3>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 4: wg_reflect_regs<4>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 5: wg_reflect_regs<5>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 6: wg_reflect_regs<6>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 7: wg_reflect_regs<7>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 8: wg_reflect_regs<8>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 9: wg_reflect_regs<9>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 10: wg_reflect_regs<10>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 11: wg_reflect_regs<11>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 12: wg_reflect_regs<12>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 13: wg_reflect_regs<13>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 14: wg_reflect_regs<14>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        case 15: wg_reflect_regs<15>(V, ldv, T, tT, C, ldc, m, n, jb); break;
        default: wg_reflect_regs<16>(V, ldv, T, tT, C, ldc, m, n, jb); break;
    }
}

// ---------------------------------------------------------------------------
// UDV by pre-pivoted blocked Householder QR of M (N x N):
//   columns sorted once by decreasing norm (P), then M P = Q R without
//   further pivoting (Bauer, SciPost Phys. Core 2, 011 (2020));
//   U = Q (explicit), D = |diag R|, Tl = D^-1 R P^T (dense N x N).
// Panels of NB columns are factored in LDS; the block reflector
// I - V T V^T (LAPACK dlarft, forward/columnwise) updates the trailing
// columns and forms Q with FP64 MFMA GEMMs.
// Buffers: M is overwritten by the reflectors V (unit lower, explicit),
// Mp (N x N) receives M P and then R, Wq holds 2 x NB x N of GEMM scratch and
// the NB x NB T factor of every panel.
// ---------------------------------------------------------------------------
__device__ __noinline__ void wg_qr_udv(double* M, double* Mp, int N, double* U, double* D, double* Tl, double* Wq, Lds&) {
    Lds& s = g_lds;
    double* P = s.panel;                  // m x jb panel, ld = m
    double* Wsm = Wq;                     // NB x N GEMM scratch (ld NB)
    double* Wsm2 = Wq + (size_t)NB * N;   // NB x N GEMM scratch (ld NB)
    double* Tg = Wq + (size_t)2 * NB * N; // T factors of all panels, NB x NB each
    N = uni(N);
    const AS_G double* Mg = gptr(uni(M));
    AS_G double* Mpg = gptr(uni(Mp));
    // column norms (all of a lane's <= NMAX/64 elements in flight) and the sorting permutation
    for (int j = wave(); j < N; j += 4) {
        double v[NMAX / 64];
#pragma unroll
        for (int q = 0; q < NMAX / 64; ++q) {
            const int i = lane() + 64 * q;
            v[q] = i < N ? Mg[(size_t)j * N + i] : 0.0;
        }
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < NMAX / 64; ++q) acc += v[q] * v[q];
        acc = wave_sum(acc);
        if (lane() == 0) s.v0[j] = acc;
    }
    __syncthreads();
    // rank by decreasing norm, ties by index; a NaN norm ranks as +inf so the
    // ranks always form a permutation
    for (int j = tid(); j < N; j += NT) {
        const double nj = isnan(s.v0[j]) ? INFINITY : s.v0[j];
        int r = 0;
        for (int i = 0; i < N; ++i) {
            const double ni = isnan(s.v0[i]) ? INFINITY : s.v0[i];
            r += (ni > nj) || (ni == nj && i < j);
        }
        s.perm[r] = j;
    }
    __syncthreads();
    const FastDiv dN(N);
    wg_stream(N * N, [&](int e) { const int c = dN.div(e), r = e - c * N; return Mg[(size_t)s.perm[c] * N + r]; },
              [&](int e, double v) { Mpg[e] = v; });
    __syncthreads();
    for (int k0 = 0; k0 < N; k0 += NB) {
        const int jb = min(NB, N - k0), m = N - k0;
        const FastDiv dm(m);
        SubClock pclk;
        pclk.start();
        wg_stream(m * jb, [&](int e) { const int c = dm.div(e), r = e - c * m; return Mpg[(size_t)(k0 + c) * N + k0 + r]; },
                  [&](int e, double v) { P[e] = v; });
        __syncthreads();
        qr_panel(m, jb, k0, N, M, Mp, Tg, s);
        pclk.stop(1);
        // trailing columns: A2 <- (I - V T^T V^T) A2
        const int n2 = N - k0 - jb;
        if (n2 > 0) {
            const double* V = M + (size_t)k0 * N + k0;
            double* A2 = Mp + (size_t)(k0 + jb) * N + k0;
            if (m <= 256) reflect_regs(V, N, Tg + (size_t)(k0 / NB) * NB * NB, true, A2, N, m, n2, jb);
            else wg_reflect(V, N, Tg + (size_t)(k0 / NB) * NB * NB, true, A2, N, m, n2, jb, Wsm, Wsm2, s);
        }
    }
    // D and Tl = D^-1 R P^T (D also kept in LDS for the division)
    for (int k = tid(); k < N; k += NT) {
        const double dk = fabs(Mpg[(size_t)k * N + k]);
        gptr(D)[k] = dk;
        s.v2[k] = dk;
    }
    __syncthreads();
    {
        AS_G double* Tlg = gptr(uni(Tl));
        wg_stream(N * N, [&](int e) { const int c = dN.div(e), r = e - c * N; return c >= r ? Mpg[e] / s.v2[r] : 0.0; },
                  [&](int e, double v) { const int c = dN.div(e), r = e - c * N; Tlg[(size_t)s.perm[c] * N + r] = v; });
    }
    // Q = H_0 H_1 ... : U <- (I - V T V^T) U, panels from the last
    wg_identity(U, N);
    __syncthreads();
    for (int k0 = ((N - 1) / NB) * NB; k0 >= 0; k0 -= NB) {
        const int jb = min(NB, N - k0), m = N - k0;
        const double* V = M + (size_t)k0 * N + k0;
        const double* T = Tg + (size_t)(k0 / NB) * NB * NB;
        double* U2 = U + (size_t)k0 * N + k0;
        if (m <= 256) reflect_regs(V, N, T, false, U2, N, m, m, jb);
        else wg_reflect(V, N, T, false, U2, N, m, m, jb, Wsm, Wsm2, s);
    }
}

// ---------------------------------------------------------------------------
// UDV stacks.  The left and right stacks of a spin share one array of
// n_stack slots: the right stack fills slots [0, n_right) upward, the left
// stack fills [n_right, n_stack) downward, so a pop of one and the push of
// the other always meet at the same slot.  Each slot keeps U, D and the
// cumulative T (= V^T) of the stack product up to that level.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double* slotU(Desc d, int sp, int k) {
    return d.stkU + ((size_t)sp * d.n_stack + k) * d.N * d.N;
}
__device__ __forceinline__ double* slotT(Desc d, int sp, int k) {
    return d.stkT + ((size_t)sp * d.n_stack + k) * d.N * d.N;
}
__device__ __forceinline__ double* slotD(Desc d, int sp, int k) {
    return d.stkD + ((size_t)sp * d.n_stack + k) * d.N;
}

// push A onto the stack whose current top is slot `below` (-1 if empty); result at `slot`
// stabilization scratch of one spin (each spin has its own, so the two spins of a pair launch
// stabilize side by side)
__device__ __forceinline__ double* spinW(Desc d, int sp) { return d.W + (size_t)sp * 3 * d.N * d.N; }
__device__ __forceinline__ double* spinWq(Desc d, int sp) { return d.Wq + (size_t)sp * 3 * 16 * (d.N + 16); }

__device__ __noinline__ void wg_push(Desc d, int sp, int slot, int below, const double* A, Lds&) {
    Lds& s = g_lds;
    const int N = d.N;
    double* W0 = spinW(d, sp);
    double* W1 = W0 + (size_t)N * N;
    if (below < 0) wg_copy(W0, A, N * N);
    else {
        const double* Db = slotD(d, sp, below);   // the GEMM's column scaling, staged in LDS
        for (int i = tid(); i < N; i += NT) s.v3[i] = gptr(Db)[i];
        __syncthreads();
        gemm64(N, N, N, A, N, false, slotU(d, sp, below), N, false, W0, N, nullptr, nullptr, s.v3, 0.0, s);
    }
    wg_qr_udv(W0, W0 + 2 * (size_t)N * N, N, slotU(d, sp, slot), slotD(d, sp, slot), W1, spinWq(d, sp), s);
    __syncthreads();
    if (below < 0) wg_copy(slotT(d, sp, slot), W1, N * N);
    else gemm64(N, N, N, W1, N, false, slotT(d, sp, below), N, false, slotT(d, sp, slot), N, nullptr, nullptr, nullptr, 0.0, s);
}

// dst1(r,c) = sc[r] src(c,r) (transpose, optional row scaling) and, with dst2,
// dst2(r,c) = dst1(r,c) + sa[r] add(r,c).  Column blocks of src are staged in
// LDS with coalesced reads, so both the reads and the writes run along columns.
// sc, sa live in LDS; all matrices are N x N in global memory.
__device__ __noinline__ void wg_transpose_scaled(double* dst1_, double* dst2_, const double* src_, const double* add_,
                                                 int N, const double* sc_, const double* sa_, Lds&) {
    Lds& s = g_lds;
    N = uni(N);
    AS_G double* dst1 = gptr(uni(dst1_));
    AS_G double* dst2 = gptr(uni(dst2_));
    const AS_G double* src = gptr(uni(src_));
    const AS_G double* add = gptr(uni(add_));
    // LDS views of the scalings; the null tests below use the generic pointers
    const AS_L double* sc = (const AS_L double*)uni(sc_);
    const AS_L double* sa = (const AS_L double*)uni(sa_);
    // (the block's columns sit N + 1 doubles apart in LDS, so the transposed reads of consecutive
    // threads fall on different banks)
    const int ldb = N + 1;
    const FastDiv dN(N);
    const int nb = max(1, min(N, PANEL_DOUBLES / ldb));
    for (int j0 = 0; j0 < N; j0 += nb) {
        const int nc = min(nb, N - j0);
        const FastDiv dnc(nc);
        const AS_G double* blk = src + (size_t)j0 * N;
        wg_stream(nc * N, [&](int e) { return blk[e]; }, [&](int e, double v) { s.panel[e + dN.div(e)] = v; });
        __syncthreads();
        // element (r, c) of dst with r = j0 + e % nc, c = e / nc is src(c, r) = panel[c + (r - j0) ldb]
        if (dst2_) {
            wg_stream(nc * N, [&](int e) { const int c = dnc.div(e), r = j0 + e - c * nc; return add[(size_t)c * N + r]; },
                      [&](int e, double a) {
                          const int c = dnc.div(e), rr = e - c * nc, r = j0 + rr;
                          const double v = (sc_ ? sc[r] : 1.0) * s.panel[c + rr * ldb];
                          dst1[(size_t)c * N + r] = v;
                          dst2[(size_t)c * N + r] = v + sa[r] * a;
                      });
        } else {
            for (int e = tid(); e < nc * N; e += NT) {
                const int c = dnc.div(e), rr = e - c * nc, r = j0 + rr;
                dst1[(size_t)c * N + r] = (sc_ ? sc[r] : 1.0) * s.panel[c + rr * ldb];
            }
        }
        __syncthreads();
    }
}

// G = (1 + B_left B_right)^-1 from the stack tops (lslot/rslot = -1 when that stack is empty);
// returns the sign of det of the matrix its LU factored (for one stack, H = Db U^T + Ds T)
__device__ __noinline__ double wg_fresh_green(Desc d, int sp, int lslot, int rslot, double* Gout, Lds&) {
    Lds& s = g_lds;
    const int N = d.N;
    double* W0 = spinW(d, sp);
    double* W1 = W0 + (size_t)N * N;
    if (lslot >= 0 && rslot >= 0) {
        const double* UL = slotU(d, sp, lslot); const double* TL = slotT(d, sp, lslot); const double* DL = slotD(d, sp, lslot);
        const double* UR = slotU(d, sp, rslot); const double* TR = slotT(d, sp, rslot); const double* DR = slotD(d, sp, rslot);
        // v0 = 1/dlmax, v1 = dlmin, v2 = 1/drmax, v3 = drmin  (include/utils/numerical_stable.hpp:57-71)
        for (int i = tid(); i < N; i += NT) {
            const double dl = gptr(DL)[i], dr = gptr(DR)[i];
            s.v0[i] = dl >= 1.0 ? 1.0 / dl : 1.0; s.v1[i] = dl >= 1.0 ? 1.0 : dl;
            s.v2[i] = dr >= 1.0 ? 1.0 / dr : 1.0; s.v3[i] = dr >= 1.0 ? 1.0 : dr;
        }
        __syncthreads();
        // W0 = dlmax^-1 (UL^T UR) drmax^-1 + dlmin (TL TR^T) drmin
        gemm64(N, N, N, UL, N, true, UR, N, false, W0, N, s.v0, nullptr, s.v2, 0.0, s);
        gemm64(N, N, N, TL, N, false, TR, N, true, W0, N, s.v1, nullptr, s.v3, 1.0, s);
        // W1 = dlmax^-1 UL^T
        wg_transpose_scaled(W1, nullptr, UL, nullptr, N, s.v0, nullptr, s);
        const double ds = wg_lu_solve(W0, W1, N, N, s);
        // G = UR drmax^-1 W1
        gemm64(N, N, N, UR, N, false, W1, N, false, Gout, N, nullptr, s.v2, nullptr, 0.0, s);
        return ds;
    }
    // one stack only: X = (Db U^T + Ds T)^-1 Db U^T   (include/utils/numerical_stable.hpp:132-152)
    const int k = lslot >= 0 ? lslot : rslot;
    const double* U = slotU(d, sp, k); const double* T = slotT(d, sp, k); const double* D = slotD(d, sp, k);
    for (int i = tid(); i < N; i += NT) {
        const double x = gptr(D)[i];
        s.v0[i] = x > 1.0 ? 1.0 / x : 1.0;
        s.v1[i] = x > 1.0 ? 1.0 : x;
    }
    __syncthreads();
    // W1 = Db U^T, W0 = W1 + Ds T
    wg_transpose_scaled(W1, W0, U, T, N, s.v0, s.v1, s);
    const double ds = wg_lu_solve(W0, W1, N, N, s);
    if (lslot >= 0) wg_copy(Gout, W1, N * N);
    else wg_transpose_scaled(Gout, nullptr, W1, nullptr, N, nullptr, nullptr, s);  // right stack: G = X^T
    return ds;
}

// ---------------------------------------------------------------------------
// Metropolis over the sites of one slice with delayed rank-k updates.
// G_cur = G0 + X Y^T; the diagonal of G_cur is kept in LDS; thread 0 draws
// the decisions (it can walk over rejected sites without synchronization).
// ---------------------------------------------------------------------------
// Submatrix updates (Nukala, Maier, Summers, Alvarez, Schulthess, PRB 80,
// 195111 (2009)).  The sites of a slice are taken in blocks S of k
// consecutive sites.  Within a block every accepted flip keeps
//   G_cur = G0 + G0[:,S] C (G0[S,:] - E_S),   C (k x k), C <- 0 at block start,
// so the decisions need only the k x k submatrix G0[S,S]:
//   t = C (G0[S,S] - I) e_p,  G_cur(s_p,s_p) = G0(s_p,s_p) + G0[S,S](p,:) t,
//   and an accepted flip with gamma does C <- C + gamma (e_p + t)(e_p + C^T G0[S,S](p,:)^T)^T,
// which is the reference's rank-1 update G <- G - gamma G[:,i] (e_i^T - G[i,:])
// (src/model/repulsive_hubbard.cpp:127-155) written in the block basis.
// One wavefront walks the block (lane a owns row a and column a of C in LDS,
// so the walk needs shuffles only); the block then lands in G with
// W = C (G0[S,:] - E_S) and G += G0[:,S] W on the matrix cores.
#define KB 32   // one half-wavefront per spin in the walk

// LDS layout of the Metropolis phase (KP = KB + 1).  The walk's arrays live in
// the front of the LDS object (the GEMM staging tiles and the scaling and
// scratch vectors, none of which the Metropolis phase uses; the scalings are
// recomputed after it):
//   CR(sp)  = row a of the block's C at a*KP (written at the end of the walk)
//   GR      = G0[S,S] of both spins interleaved by walk lane: element (a, b) of
//             spin sp at b*64 + 32 sp + a, zero outside the k x k block, so that
//             lane 32 sp + a reads its row a at stride 64 (one conflict-free
//             512-byte access per element) and row p of a spin is a broadcast
//   GT      = the same for G0[S,S]^T: element (b, a) at b*64 + 32 sp + a (the
//             columns of G0[S,S]), in the space of CR(0) and CR(1): it is staged
//             after block_stage has read CR and read by the walk before the
//             walk writes CR
//   UV      = the walk's u, v broadcast vectors
//   ACPS    = rows S' (the next block's sites) of G0[:,S] for the next block's
//             G0[S',S'] (element (a, j) at j*KB + a), one spin at a time
// The panel holds GSW = G0[S,:] (k4 x N, ld N + 1) of each spin of the workgroup
// side by side (when two fit), which becomes W in place; the block update reads
// its other operand, G0[:,S], from G itself (each wavefront loads the rows of a
// 16-row strip before it writes any of them).  Because the walk touches neither,
// the block updates of the workgroup's spins run on wavefronts 1..3 while
// wavefront 0 walks the next block.
constexpr int KP = KB + 1;
#define MP_BASE (reinterpret_cast<double*>(&g_lds))
#define MP_CR(sp) (MP_BASE + (sp) * KB * KP)
#define MP_GR (MP_BASE + 2 * KB * KP)
#define MP_GT MP_BASE
static_assert(64 * KB <= 2 * KB * KP, "the lane-interleaved G0[S,S] must fit in two C blocks");
#define MP_UV (MP_BASE + 4 * KB * KP)
#define MP_ACPS (MP_BASE + 4 * KB * KP + 128)
#define MP_SLACK 16   // reads of partial 16x16 blocks run up to 16 doubles past the last GSW
// leading dimension of GSW / W: odd, so that the 32 lanes storing one column of G0[S,:] (rows
// b = 0..31 of GSW) fall on different LDS banks
__device__ __forceinline__ int gsw_ld(int N) { return N + 1; }
static_assert(offsetof(Lds, As) == 0 && offsetof(Lds, perm) >= (4 * KB * (KB + 1) + 128 + KB * KB) * sizeof(double),
              "the walk's arrays must fit in front of Lds::perm");

// block size: the delay depth, at most KB, and small enough that one spin's GSW fits in the
// panel (the same in every launch mapping, so a chain's numbers do not depend on it)
__device__ __forceinline__ int block_depth(Desc d) {
    const int N = d.N;
    int kb = min(d.kdelay, KB);
    const int cap = ((PANEL_DOUBLES - MP_SLACK) / gsw_ld(N)) & ~3;
    return max(1, min(kb, cap));
}


// G0[S,S] of block (b0, k) of one spin into GSS, zero outside the k x k block
__device__ void stage_gss(const double* G_, int N, int b0, int k, int sp, Lds&) {
    const AS_G double* Gg = gptr(G_);
    double* gr = MP_GR + 32 * sp;
    double* gt = MP_GT + 32 * sp;
    for (int q = tid(); q < KB * KB; q += NT) { gr[(q >> 5) * 64 + (q & 31)] = 0.0; gt[(q >> 5) * 64 + (q & 31)] = 0.0; }
    __syncthreads();
    const FastDiv dk(k);
    wg_stream(k * k, [&](int q) { const int c = dk.div(q), a = q - c * k; return Gg[(size_t)(b0 + c) * N + b0 + a]; },
              [&](int q, double v) { const int c = dk.div(q), a = q - c * k; gr[c * 64 + a] = v; gt[a * 64 + c] = v; });
}

// W = C (G0[S,:] - E_S) over GSW in place, for a block of padded depth K4: a wavefront advances up
// to 4 of its column blocks together; all loads are unconditional (clamped indices) so the
// compiler keeps them in flight.  Rows 16..31 of C are zero when K4 <= 16 (not multiplied).
template <int K4>
__device__ __forceinline__ void stage_w_product(int N, int ldw, int b0, int k, const double* cr, double* gsw) {
    constexpr bool hi = K4 > 16;
    const int l = lane(), w = wave();
    const int nbc = (N + 15) >> 4;
    const int lr = l & 15, lk = l >> 4;
    for (int j0 = w; j0 < nbc; j0 += 16) {
        dbl4 acc0[4], acc1[4];
        int colq[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc0[u] = dbl4{0.0, 0.0, 0.0, 0.0};
            acc1[u] = dbl4{0.0, 0.0, 0.0, 0.0};
            colq[u] = min((j0 + 4 * u) * 16 + lr, N - 1);   // a column past N repeats column N-1; not stored
        }
#pragma unroll
        for (int kk = 0; kk < K4; kk += 4) {
            const int kr = kk + lk;
            const double a0 = cr[lr * KP + kr];
            const double a1 = hi ? cr[(16 + lr) * KP + kr] : 0.0;
            double bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) bv[u] = gsw[kr * ldw + colq[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc0[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bv[u], acc0[u], 0, 0, 0);
                if (hi) acc1[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bv[u], acc1[u], 0, 0, 0);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int col = (j0 + 4 * u) * 16 + lr;
            const bool colok = j0 + 4 * u < nbc && col < N;
            const bool in_s = col >= b0 && col < b0 + k;
            const int cs = min(max(col - b0, 0), KB - 1);
            double corr0[4], corr1[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int a = lk + 4 * i;
                corr0[i] = cr[a * KP + cs];
                corr1[i] = hi ? cr[(16 + a) * KP + cs] : 0.0;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int a = lk + 4 * i;
                if (colok && a < K4) gsw[a * ldw + col] = acc0[u][i] - ((in_s && a < k) ? corr0[i] : 0.0);
                if (hi && colok && 16 + a < K4) gsw[(16 + a) * ldw + col] = acc1[u][i] - ((in_s && 16 + a < k) ? corr1[i] : 0.0);
            }
        }
    }
}

// The block's update of G for one spin, on the matrix cores, and the next block's G0[S',S'].
// block_stage:  GSW = G0[S,:] to LDS slot `slot` of the panel, then W = C (G0[S,:] - E_S) over
// GSW; with a next block (b1, k1 > 0), also its G0[S',S'] = G[S',S'] + ACP[S',:] W[:,S'] formed
// from these factors before they are applied (the same update block_apply makes to those
// elements), into this spin's half of GR only (GT shares its space with the C rows that the other
// spin's staging may still read, and is filled by gss_transpose afterwards).  The loads of
// G[S',S'] and of ACP[S',:] = G0[S',S] (staged in ACPS) are issued ahead of the GSW stream, so
// all of the stage's loads are in flight together.
__device__ __noinline__ void block_stage(const double* G_, int N, int b0, int k, int b1, int k1, int sp, int slot, Lds&) {
    Lds& s = g_lds;
    N = uni(N); b0 = uni(b0); k = uni(k); b1 = uni(b1); k1 = uni(k1); sp = uni(sp); slot = uni(slot);
    const AS_G double* G = gptr(uni(G_));
    const int k4 = (k + 3) & ~3;
    const FastDiv dk4(k4);
    const int ldw = gsw_ld(N);
    double* const gsw = s.panel + (size_t)slot * k4 * ldw;
    double* const acps = MP_ACPS;
    double* const gr = MP_GR + 32 * sp;
    const double* const cr = MP_CR(sp);
    const bool nxt = k1 > 0;
    const int kn = nxt ? k1 : 1;
    const FastDiv dkn(kn);
    // the thread's (at most 4) elements of G0[S',S'] (q = c k1 + a) and of ACPS (q = j k1 + a)
    constexpr int U = KB * KB / NT;
    double gv[U], av[U];
    int ga[U], gc[U], aj[U], aa[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int q = tid() + u * NT;
        const int qg = min(q, kn * kn - 1), c = dkn.div(qg), a = qg - c * kn;   // (clamped; not stored)
        const int qa = min(q, kn * k4 - 1), j = dkn.div(qa), a2 = qa - j * kn;
        ga[u] = a; gc[u] = c; aj[u] = j; aa[u] = a2;
        gv[u] = nxt ? G[(size_t)(b1 + c) * N + b1 + a] : 0.0;
        av[u] = nxt ? G[(size_t)(b0 + min(j, k - 1)) * N + b1 + a2] : 0.0;
    }
    if (nxt)
        for (int q = tid(); q < KB * KB; q += NT) gr[(q >> 5) * 64 + (q & 31)] = 0.0;
    // GSW element e = c k4 + b: consecutive threads read down a column of G (in one round of loads
    // when the block has at most 18 elements per thread, e.g. 32 x 144)
    auto gld = [&](int e) { const int c = dk4.div(e), b = e - c * k4; return G[(size_t)c * N + b0 + min(b, k - 1)]; };
    auto gst = [&](int e, double v) { const int c = dk4.div(e), b = e - c * k4; gsw[b * ldw + c] = b < k ? v : 0.0; };
    if (N * k4 > 16 * NT && N * k4 <= 18 * NT) wg_stream<18>(N * k4, gld, gst);
    else wg_stream<16>(N * k4, gld, gst);
    if (nxt)
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (tid() + u * NT < k1 * k4) acps[aj[u] * KB + aa[u]] = aj[u] < k ? av[u] : 0.0;
    __syncthreads();
    // W = C G0[S,:] - C E_S, column block by column block (each wavefront owns whole column blocks,
    // so W may overwrite GSW in place), with the depth loop unrolled for the padded depth k4
    switch (k4) {
        case 4: stage_w_product<4>(N, ldw, b0, k, cr, gsw); break;
        case 8: stage_w_product<8>(N, ldw, b0, k, cr, gsw); break;
        case 12: stage_w_product<12>(N, ldw, b0, k, cr, gsw); break;
        case 16: stage_w_product<16>(N, ldw, b0, k, cr, gsw); break;
        case 20: stage_w_product<20>(N, ldw, b0, k, cr, gsw); break;
        case 24: stage_w_product<24>(N, ldw, b0, k, cr, gsw); break;
        case 28: stage_w_product<28>(N, ldw, b0, k, cr, gsw); break;
        default: stage_w_product<32>(N, ldw, b0, k, cr, gsw); break;
    }
    __syncthreads();
    if (nxt) {
        // G0[S',S']: the (at most 4) elements of a thread advance together, four independent sums
        for (int j = 0; j < k4; ++j) {
#pragma unroll
            for (int u = 0; u < U; ++u) gv[u] += acps[j * KB + ga[u]] * gsw[j * ldw + b1 + gc[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (tid() + u * NT < k1 * k1) gr[gc[u] * 64 + ga[u]] = gv[u];
        __syncthreads();   // the next spin's staging overwrites ACPS
    }
}

// block_apply:  G_m += ACP_m W_m (ACP_m = G0_m[:,S], S = b0..b0+k-1) for the nmat (1 or 2) spins
// of the workgroup, on wavefronts wave_lo..3 (no workgroup barrier inside, so it can run beside
// the walk of wavefront 0).  The work goes in strips of 16 rows of G (the spins' strips in
// turn), handed out in turn or by the LDS counter `next`; the wavefront that owns a strip loads
// the strip's rows of ACP from G into registers before it writes any of its elements, then runs
// along the strip by steps of CB 16-column blocks, loading each step's G under the previous
// step's products (and the next strip's ACP and first step under the strip's last).  The
// product is formed transposed (W^T ACP^T), so 16 lanes run down one column of G; W_m (K4 x N,
// ld N + 1) is in LDS.  Each block sums its K4/4 MFMAs in depth order.
template <int K4, int CB>
__device__ __forceinline__ void block_apply_k(int nmat, AS_G double* G0, AS_G double* G1, const AS_L double* W0,
                                              const AS_L double* W1, int N, int b0, int k, int wave_lo, AS_L int* next) {
    const int ldw = gsw_ld(N);
    constexpr int KQ = K4 / 4;
    const int l = lane(), w = wave();
    const int nw = 4 - wave_lo, wi = w - wave_lo;
    const int lr = l & 15, lk = l >> 4;
    const int nrb = (N + 15) >> 4, nu = nmat * nrb, ncp = (N + 16 * CB - 1) / (16 * CB);
    auto claim = [&]() {
        int g = 0;
        if (l == 0) g = __hip_atomic_fetch_add(next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return __builtin_amdgcn_readfirstlane(g);
    };
    double avn[KQ], cvn[CB][4];
    // ACP of strip u (row rbase + lr, depth 4q + lk; zero past the block's k sites)
    auto load_a = [&](int u) {
        const bool m1 = u >= nrb;
        const AS_G double* const Gm = m1 ? G1 : G0;
        const int row = min((u - (m1 ? nrb : 0)) * 16 + lr, N - 1);
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
            const int j = 4 * q + lk;
            const double v = Gm[(size_t)(b0 + min(j, k - 1)) * N + row];
            avn[q] = j < k ? v : 0.0;
        }
    };
    // G of column step cp of strip u
    auto load_c = [&](int u, int cp) {
        const bool m1 = u >= nrb;
        const AS_G double* const Gm = m1 ? G1 : G0;
        const int row = min((u - (m1 ? nrb : 0)) * 16 + lr, N - 1);
#pragma unroll
        for (int b = 0; b < CB; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) cvn[b][i] = Gm[(size_t)min((CB * cp + b) * 16 + lk + 4 * i, N - 1) * N + row];
    };
    int u = next ? claim() : wi;
    if (u < nu) {
        load_a(u);
        load_c(u, 0);
    }
    int cp = 0;
    double av[KQ];
    while (u < nu) {
        if (cp == 0)
#pragma unroll
            for (int q = 0; q < KQ; ++q) av[q] = avn[q];
        double cv[CB][4];
#pragma unroll
        for (int b = 0; b < CB; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) cv[b][i] = cvn[b][i];
        // the next step: the next column step of this strip, or the first step of the next strip
        int un = u, cpn = cp + 1;
        if (cpn == ncp) {
            un = next ? claim() : u + nw;
            cpn = 0;
            if (un < nu) load_a(un);
        }
        if (un < nu) load_c(un, cpn);
        const bool m1 = u >= nrb;
        AS_G double* const Gm = m1 ? G1 : G0;
        const AS_L double* const Wm = m1 ? W1 : W0;
        const int row = (u - (m1 ? nrb : 0)) * 16 + lr;
        int ob[CB];
#pragma unroll
        for (int b = 0; b < CB; ++b) ob[b] = min((CB * cp + b) * 16 + lr, N - 1);   // (columns past N repeat the last; not stored)
        dbl4 acc[CB];
#pragma unroll
        for (int b = 0; b < CB; ++b) acc[b] = dbl4{0.0, 0.0, 0.0, 0.0};
        // (W operands double-buffered: step q + 1's LDS reads are issued before step q's MFMAs, so
        // their latency runs under the matrix cores instead of between the steps)
        double wv[2][CB];
#pragma unroll
        for (int b = 0; b < CB; ++b) wv[0][b] = Wm[lk * ldw + ob[b]];
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
            const int cur = q & 1;
            if (q + 1 < KQ) {
                const int kw = (4 * q + 4 + lk) * ldw;
#pragma unroll
                for (int b = 0; b < CB; ++b) wv[cur ^ 1][b] = Wm[kw + ob[b]];
            }
            __builtin_amdgcn_sched_barrier(0);   // (keeps the reads ahead of the MFMAs)
#pragma unroll
            for (int b = 0; b < CB; ++b) acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(wv[cur][b], av[q], acc[b], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int b = 0; b < CB; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int col = (CB * cp + b) * 16 + lk + 4 * i;
                if (row < N && col < N) Gm[(size_t)col * N + row] = cv[b][i] + acc[b][i];
            }
        u = un;
        cp = cpn;
    }
}
template <int CB>
__device__ __forceinline__ void apply_depth(int k4, int nmat, AS_G double* G0, AS_G double* G1, const AS_L double* W0,
                                            const AS_L double* W1, int N, int b0, int k, int wave_lo, AS_L int* next) {
    switch (k4) {
        case 4: block_apply_k<4, CB>(nmat, G0, G1, W0, W1, N, b0, k, wave_lo, next); break;
        case 8: block_apply_k<8, CB>(nmat, G0, G1, W0, W1, N, b0, k, wave_lo, next); break;
        case 12: block_apply_k<12, CB>(nmat, G0, G1, W0, W1, N, b0, k, wave_lo, next); break;
        case 16: block_apply_k<16, CB>(nmat, G0, G1, W0, W1, N, b0, k, wave_lo, next); break;
        case 20: block_apply_k<20, CB>(nmat, G0, G1, W0, W1, N, b0, k, wave_lo, next); break;
        case 24: block_apply_k<24, CB>(nmat, G0, G1, W0, W1, N, b0, k, wave_lo, next); break;
        case 28: block_apply_k<28, CB>(nmat, G0, G1, W0, W1, N, b0, k, wave_lo, next); break;
        default: block_apply_k<32, CB>(nmat, G0, G1, W0, W1, N, b0, k, wave_lo, next); break;
    }
}
// the block updates (block b0..b0+k-1) of spins lo..hi-1 (their W in panel slots 0.. in order)
__device__ __noinline__ void block_apply(Desc d, int lo, int hi, int b0, int k, int wave_lo, bool dyn, Lds&) {
    Lds& s = g_lds;
    const int N = uni(d.N);
    lo = uni(lo); hi = uni(hi); b0 = uni(b0); k = uni(k); wave_lo = uni(wave_lo); dyn = uni(dyn);
    const int k4 = (k + 3) & ~3, nmat = hi - lo;
    if (!dyn && wave() < wave_lo) return;
    AS_L int* const next = dyn ? (AS_L int*)&s.ivar[3] : nullptr;
    if (dyn) wave_lo = 0;
    AS_G double* const G0 = gptr(uni(d.G) + (size_t)lo * N * N);
    AS_G double* const G1 = gptr(uni(d.G) + (size_t)(hi - 1) * N * N);
    const AS_L double* const W0 = (const AS_L double*)s.panel;
    const AS_L double* const W1 = W0 + (size_t)(nmat - 1) * k4 * gsw_ld(N);
    // steps of 4, 3 or 2 column blocks: the widest that pads the fewest columns
    const int p4 = (N + 63) / 64 * 64, p3 = (N + 47) / 48 * 48, p2 = (N + 31) / 32 * 32;
    if (p4 <= p3 && p4 <= p2) apply_depth<4>(k4, nmat, G0, G1, W0, W1, N, b0, k, wave_lo, next);
    else if (p3 <= p2) apply_depth<3>(k4, nmat, G0, G1, W0, W1, N, b0, k, wave_lo, next);
    else apply_depth<2>(k4, nmat, G0, G1, W0, W1, N, b0, k, wave_lo, next);
}

// GT half of one spin from its GR half (the whole 32 x 32, zeros included)
__device__ void gss_transpose(int sp, Lds&) {
    __syncthreads();
    const double* gr = MP_GR + 32 * sp;
    double* gt = MP_GT + 32 * sp;
    // diagonal order (lane t of group g takes a = t, c = (t + g) mod 32), so that neither the reads
    // nor the writes of 32 lanes fall on one LDS bank (both arrays have rows 64 doubles apart)
    for (int q = tid(); q < KB * KB; q += NT) {
        const int a = q & 31, c = ((q >> 5) + a) & 31;
        gt[a * 64 + c] = gr[c * 64 + a];
    }
}

// The in-block walk, run by wavefront 0 only.  Lanes 0..31 carry spin up and
// lanes 32..63 spin down; lane h*32 + a owns site s_a of the block in spin h
// and keeps row a and column a of that spin's C in registers for the whole
// block, so a flip costs LDS reads only (the loops run over the KW >= k rows
// of the block's padded depth with fixed trip counts and unroll; G0[S,S] and C
// are zero outside the k x k block, so the padding terms vanish exactly).  The
// decisions are computed by every lane on identical data, in the reference's
// site order.  On return, the first KW entries of row a of C are in CR(h) for
// the block update (its W product reads no further).
struct WalkOut {
    uint64_t rng;
    double sign;
    int accepted;
};
template <int KW>
__device__ __forceinline__ WalkOut walk_block(Desc d, int e, int b0, int k, uint64_t rng, double sign,
                                           bool record, Lds&) {
    Lds& s = g_lds;
    const int N = uni(d.N);
    const int ns = uni(d.nspin);
    b0 = uni(b0); k = uni(k);
    const double pref0 = uni(d.pref[0]), pref1 = uni(d.pref[1]);
    const double dm00 = uni(d.dm1[0][0]), dm01 = uni(d.dm1[0][1]), dm10 = uni(d.dm1[1][0]), dm11 = uni(d.dm1[1][1]);
    // (global, not generic, stores: a generic store would also count against the
    // LDS counter and make the walk's next LDS wait a global-memory round trip)
    AS_G int8_t* const field_row = (AS_G int8_t*)(uni(d.field) + (size_t)e * N);
    AS_G uint8_t* const rec_row = (AS_G uint8_t*)(uni(d.recAcc) + (size_t)e * N);
    const int h = lane() >> 5, a = lane() & 31;
    const double* const gr = MP_GR;
    const double* const gt = MP_GT;
    double* const ush = MP_UV + h * 32;        // u at [lane], v at [64 + lane]
    double* const vsh = MP_UV + 64 + h * 32;
    double crr[KW], ccr[KW];
#pragma unroll
    for (int b = 0; b < KW; ++b) { crr[b] = 0.0; ccr[b] = 0.0; }
    double dq = gr[a * 64 + lane()];
    int nacc = 0;
    // the block's k uniform draws, one per lane (site a), bit-identical to drawing them in order
    const double draw = block_canonical(rng, a < k ? a : 0);
    for (int p = 0; p < k; ++p) {
        const double dp0 = readlane_d(dq, p);
        const double dp1 = ns == 2 ? readlane_d(dq, 32 + p) : dp0;
        const int j = b0 + p;
        const int f = __builtin_amdgcn_readfirstlane(s.frow[j]);
        const bool up = f > 0;
        const double m0 = up ? dm00 : dm01, m1 = up ? dm10 : dm11;
        const double x0 = 1.0 + (1.0 - dp0) * m0;
        const double x1 = 1.0 + (1.0 - dp1) * m1;
        const double r = (up ? pref0 : pref1) * x0 * x1;
        const double pr = fmin(1.0, fabs(r));
        const int ok = __builtin_amdgcn_readfirstlane(readlane_d(draw, p) < pr ? 1 : 0);
        if (record && lane() == 0) rec_row[j] = (uint8_t)ok;
        if (!ok) continue;
        ++nacc;
        sign = (r >= 0) ? sign : -sign;
        if (lane() == 0) {
            s.frow[j] = -f;
            field_row[j] = (int8_t)(-f);
        }
        const double g = h == 0 ? m0 / x0 : m1 / x1;
        // u_a = d_ap + (C (Gss - I))(a,p),  v_a = d_ap + (Gss C)(p,a)
        // (four partial sums each, so the FMA dependency chains are 8 long)
        // (row p and column p of this spin's Gss are broadcasts at stride 64; the
        // reads of chunk c + 1 of four are issued before the products of chunk c)
        double t4[4] = {0.0, 0.0, 0.0, 0.0}, w4[4] = {0.0, 0.0, 0.0, 0.0};
        {
            const AS_L double* const colp = lds_row(gt + h * 32 + p);
            const AS_L double* const rowp = lds_row(gr + h * 32 + p);
            double xc[2][4], yr[2][4];
#pragma unroll
            for (int j = 0; j < 4; ++j) { xc[0][j] = colp[64 * j]; yr[0][j] = rowp[64 * j]; }
#pragma unroll
            for (int c = 0; c < KW / 4; ++c) {
                const int cur = c & 1, nxt = cur ^ 1;
                if (c + 1 < KW / 4) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) { xc[nxt][j] = colp[64 * (4 * c + 4 + j)]; yr[nxt][j] = rowp[64 * (4 * c + 4 + j)]; }
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int b = 4 * c + j;
                    t4[j] += crr[b] * (xc[cur][j] - (b == p ? 1.0 : 0.0));
                    w4[j] += yr[cur][j] * ccr[b];
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        const double u = (a == p ? 1.0 : 0.0) + ((t4[0] + t4[1]) + (t4[2] + t4[3]));
        const double v = (a == p ? 1.0 : 0.0) + ((w4[0] + w4[1]) + (w4[2] + w4[3]));
        ush[a] = u;
        vsh[a] = v;
        __builtin_amdgcn_wave_barrier();
        // C += g u v^T (row a and column a), and the block diagonal
        // d_a += g (Gss u)_a (v^T (Gss - I))_a, the identity's share v_a taken out of the sum
        double x4[4] = {0.0, 0.0, 0.0, 0.0}, y4[4] = {0.0, 0.0, 0.0, 0.0};
        const double gu = g * u, gv = g * v;
        {
            const AS_L double* const rowa = lds_row(gr + lane());
            const AS_L double* const cola = lds_row(gt + lane());
            const AS_L double* const uo = lds_row(ush);
            const AS_L double* const vo = lds_row(vsh);
            double ub[2][4], vb[2][4], ga[2][4], gc[2][4];
#pragma unroll
            for (int j = 0; j < 4; ++j) { ub[0][j] = uo[j]; vb[0][j] = vo[j]; ga[0][j] = rowa[64 * j]; gc[0][j] = cola[64 * j]; }
#pragma unroll
            for (int c = 0; c < KW / 4; ++c) {
                const int cur = c & 1, nxt = cur ^ 1;
                if (c + 1 < KW / 4) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int b = 4 * c + 4 + j;
                        ub[nxt][j] = uo[b]; vb[nxt][j] = vo[b]; ga[nxt][j] = rowa[64 * b]; gc[nxt][j] = cola[64 * b];
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int b = 4 * c + j;
                    crr[b] += gu * vb[cur][j];
                    ccr[b] += ub[cur][j] * gv;
                    x4[j] += ga[cur][j] * ub[cur][j];
                    y4[j] += vb[cur][j] * gc[cur][j];
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        __builtin_amdgcn_wave_barrier();
        dq += g * ((x4[0] + x4[1]) + (x4[2] + x4[3])) * (((y4[0] + y4[1]) + (y4[2] + y4[3])) - v);
    }
    double* const cro = MP_CR(h) + a * KP;
#pragma unroll
    for (int b = 0; b < KW; ++b) cro[b] = crr[b];
    rng = mulmod_m31(rng, kMinstdPow.v[2 * k]);   // the block consumed 2k draws
    return WalkOut{rng, sign, nacc};
}

// ---------------------------------------------------------------------------
// Spin pairs.  In a pair launch a repulsive chain runs on two workgroups, one
// per spin; the launch is no larger than the device holds at once and pair
// launches never overlap each other, so the partners are resident together.  The spins are
// independent except in the walk (the acceptance ratio is the product of the
// two spins' ratios), so per delayed-update block each workgroup publishes its
// spin's G0[S,S] and takes the partner's; both then walk the block with
// identical inputs and reach identical decisions, fields, sign and random state,
// and each applies the block to its own spin.  The partners also meet around a
// measurement (which reads both spins' G) and at the end of every sweep (status
// bits and wrap error merged, so both leave the sweep loop together).
// An exchange point: every wavefront's stores reach the L2 (vmcnt(0)), thread
// 0 releases the running exchange number on this spin's flag (agent scope, so
// the partner may sit on another XCD) and acquires the partner's flag.  The wait
// is bounded by the 100 MHz wall clock: a partner silent for 30 s (e.g. kept
// off the device by another process's long kernel) sets
// DQMC_STATUS_PAIR_SYNC (fatal) and later waits are skipped, so the kernel always
// ends.  The flags are zeroed by the host before each pair launch and compared
// by signed difference.  Data published at exchange q lives in buffer slot
// q & 1: a workgroup writes slot q & 1 again only at exchange q + 2, after the
// partner has arrived at q + 1, which it does only after reading slot q & 1.
// ---------------------------------------------------------------------------
#define PAIR_TIMEOUT_TICKS 3000000000ull   // 30 s of the 100 MHz wall clock

__device__ __forceinline__ int pair_slot() { return (int)((g_lds.xseq + 1u) & 1u); }   // slot of the next exchange

__device__ __noinline__ void pair_exchange(Desc d, Lds&) {
    Lds& s = g_lds;
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this wavefront's stores are in the L2
    __syncthreads();
    if (tid() == 0) {
        const unsigned q = s.xseq + 1u;
        s.xseq = q;
        unsigned* mine = d.xflag + 32 * s.me;
        unsigned* theirs = d.xflag + 32 * (1 - s.me);
        __hip_atomic_store(mine, q, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        if (!s.xbroken) {
            const unsigned long long t0 = wall_clock64();
            while ((int)(__hip_atomic_load(theirs, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) - q) < 0) {
                if (wall_clock64() - t0 > PAIR_TIMEOUT_TICKS) {
                    s.xbroken = 1;
                    s.status |= DQMC_STATUS_PAIR_SYNC;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
    }
    __syncthreads();
}

// publish this workgroup's half of the walk's G0[S,S] (k x k, from LDS) and take the partner's
// into the other half (zero outside the k x k block, as stage_gss leaves it)
__device__ __noinline__ void pair_swap_gss(Desc d, int k, Lds&) {
    Lds& s = g_lds;
    k = uni(k);
    const int me = uni(s.me), ot = 1 - me;
    const int slot = uni(pair_slot());
    const FastDiv dk(k);
    __syncthreads();   // this spin's half is complete
    AS_G double* mine = gptr(d.xbuf + (size_t)(me * 2 + slot) * KB * KB);
    const double* gr = MP_GR + 32 * me;
    for (int q = tid(); q < k * k; q += NT) {
        const int c = dk.div(q), a = q - c * k;
        mine[q] = gr[c * 64 + a];
    }
    pair_exchange(d, s);
    const AS_G double* theirs = gptr(d.xbuf + (size_t)(ot * 2 + slot) * KB * KB);
    double* gro = MP_GR + 32 * ot;
    double* gto = MP_GT + 32 * ot;
    for (int q = tid(); q < KB * KB; q += NT) { gro[(q >> 5) * 64 + (q & 31)] = 0.0; gto[(q >> 5) * 64 + (q & 31)] = 0.0; }
    __syncthreads();
    wg_stream(k * k, [&](int q) { return theirs[q]; },
              [&](int q, double v) { const int c = dk.div(q), a = q - c * k; gro[c * 64 + a] = v; gto[a * 64 + c] = v; });
}

// end of a sweep in a pair launch: status bits and wrap error of the two spins merged (thread 0's st)
__device__ __noinline__ void pair_merge_state(Desc d, ChainState& st, Lds&) {
    Lds& s = g_lds;
    const int me = s.me, ot = 1 - me;
    const int slot = pair_slot();
    AS_G double* slots = gptr(d.xbuf + 4 * KB * KB);   // [spin][slot][status, wrap error]
    if (tid() == 0) {
        slots[(me * 2 + slot) * 2] = (double)s.status;
        slots[(me * 2 + slot) * 2 + 1] = st.wrap_error;
    }
    pair_exchange(d, s);
    if (tid() == 0) {
        s.status |= (int)slots[(ot * 2 + slot) * 2];
        const double w = slots[(ot * 2 + slot) * 2 + 1];
        if (!(w <= st.wrap_error)) st.wrap_error = w;   // (a NaN from either side is kept)
    }
    __syncthreads();
}

// (the results come back by value: references to the caller's variables would be
// generic pointers into private memory, read and written with flat instructions)
__device__ __noinline__ WalkOut wg_metropolis(Desc d, int e, uint64_t rng, double sign, bool record, Lds&) {
    Lds& s = g_lds;
    const int N = __builtin_amdgcn_readfirstlane(d.N);
    const int ns = __builtin_amdgcn_readfirstlane(d.nspin);
    const bool pair = uni(s.pair) != 0;
    const int me = uni(s.me);
    // the spins this workgroup updates: both (one-spin model: spin 0) or its own in a pair launch
    const int lo = pair ? me : 0, hi = pair ? me + 1 : ns;
    int naccept = 0;
    const int kb = __builtin_amdgcn_readfirstlane(block_depth(d));
    // both spins' GSW in the panel together (then both updates run beside the next walk), or one
    // at a time (the first spin's update is applied before the second is staged)
    const bool both = hi - lo == 2 && 2 * gsw_ld(N) * ((kb + 3) & ~3) + MP_SLACK <= PANEL_DOUBLES;
    const int plo = both ? lo : hi - 1;   // the spins whose updates wait for the next walk: plo..hi-1
    auto Gs = [&](int sp) { return d.G + (size_t)sp * N * N; };
    // G0[S,S] of the first block, both halves of the walk (also in the one-spin model)
    if (pair) {
        stage_gss(Gs(me), N, 0, min(kb, N), me, s);
        pair_swap_gss(d, min(kb, N), s);
    } else {
        for (int sp = 0; sp < 2; ++sp) stage_gss(Gs(sp < ns ? sp : 0), N, 0, min(kb, N), sp, s);
    }
    bool pending = false;   // the updates of the previous block (pb0, pk) are still to be applied
    int pb0 = 0, pk = 0;
    for (int b0 = 0; b0 < N; b0 += kb) {
        const int k = min(kb, N - b0);
        if (tid() == 0) { s.ivar[2] = 0; s.ivar[3] = 0; }
        __syncthreads();
        SubClock walk_clk;
        walk_clk.start();
        // the pending updates run beside the walk on wavefronts 1..3; wavefront 0 joins them when
        // its walk is done (the updates' block groups are handed out through ivar[3])
        if (wave() == 0) {
            WalkOut o;
            switch ((k + 7) & ~7) {   // the walk's loops run over the block's depth rounded up to 8
                case 8: o = walk_block<8>(d, e, b0, k, rng, sign, record, s); break;
                case 16: o = walk_block<16>(d, e, b0, k, rng, sign, record, s); break;
                case 24: o = walk_block<24>(d, e, b0, k, rng, sign, record, s); break;
                default: o = walk_block<32>(d, e, b0, k, rng, sign, record, s); break;
            }
            rng = o.rng;
            sign = o.sign;
            if (lane() == 0) s.ivar[2] = o.accepted;
        }
        if (pending) block_apply(d, plo, hi, pb0, pk, 1, true, s);
        __syncthreads();
        walk_clk.stop(0);
        pending = false;
        const int nblk = s.ivar[2];
        naccept += nblk;
        const int b1 = b0 + kb, k1 = min(kb, N - b1);   // the next block (k1 <= 0: none)
        if (nblk > 0) {
            for (int sp = lo; sp < hi; ++sp) {
                // (every spin's next G0[S',S'] comes from the factors, before they are applied, so a
                // pair launch, where each spin's update is the pending one, computes the same numbers)
                block_stage(Gs(sp), N, b0, k, b1, k1, sp, both ? sp - lo : 0, s);
                if (sp < plo) {
                    block_apply(d, sp, sp + 1, b0, k, 0, false, s);
                    __syncthreads();
                }
            }
            if (k1 > 0) {
                // the next walk starts from the factors; the updates of spins plo.. run beside it
                pending = true;
                pb0 = b0;
                pk = k;
                for (int sp = lo; sp < hi; ++sp) gss_transpose(sp, s);
            } else {
                block_apply(d, plo, hi, b0, k, 0, false, s);
                __syncthreads();
            }
        }
        if (k1 > 0) {
            if (pair) {
                if (nblk == 0) stage_gss(Gs(me), N, b1, k1, me, s);
                pair_swap_gss(d, k1, s);
            } else {
                // (after accepted flips: only the one-spin model's second half, a copy it does not use)
                for (int sp = nblk > 0 ? ns : 0; sp < 2; ++sp) stage_gss(Gs(sp < ns ? sp : 0), N, b1, k1, sp, s);
            }
        }
    }
    __syncthreads();
    return WalkOut{rng, sign, naccept};
}

// ---------------------------------------------------------------------------
// Equal-time measurements of one slice (src/measure/measure_methods.cpp:22-255)
// ---------------------------------------------------------------------------
__device__ __noinline__ void wg_measure_from(Desc d, const double* G, double cs, Lds&) {
    Lds& s = g_lds;
    const int N = d.N, side = d.side;
    const FastDiv dN(N), dS(side);
    const AS_G double* Gu = gptr(G);
    const AS_G double* Gd = d.nspin == 2 ? Gu + (size_t)N * N : Gu;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // tr, docc, kin, loc, md, sdw, cdw, sw
    for (int a = tid(); a < N; a += NT) {
        const double du = Gu[(size_t)a * N + a], dd = Gd[(size_t)a * N + a];
        s.du[a] = du; s.dd[a] = dd;
        v[0] += du + dd;
        v[1] += (1 - du) * (1 - dd);
        v[3] += du + dd - 2 * du * dd;
        const int x = a % side, y = a / side;
        const int jx = (x + 1) % side + side * y, jy = x + side * ((y + 1) % side);
        v[2] += Gu[(size_t)jx * N + a] + Gd[(size_t)jx * N + a] + Gu[(size_t)jy * N + a] + Gd[(size_t)jy * N + a];
    }
    __syncthreads();
    const AS_G double* Gug = Gu;
    const AS_G double* Gdg = Gd;
    const AS_G double* cosq = gptr(d.cosq);
    constexpr int MU = 4;   // elements per thread with their loads in flight together
    for (int p0 = tid(); p0 < N * N; p0 += MU * NT) {
      double q_gu_ij[MU], q_gd_ij[MU], q_gu_ji[MU], q_gd_ji[MU], q_ff[MU];
#pragma unroll
      for (int u = 0; u < MU; ++u) {
        const int p = min(p0 + u * NT, N * N - 1);
        const int j = dN.div(p), i = p - j * N;
        const int yi = dS.div(i), xi = i - yi * side, yj = dS.div(j), xj = j - yj * side;
        const int dx = xj - xi + side, dy = yj - yi + side;   // in [1, 2 side - 1]
        const int dsp = (dx >= side ? dx - side : dx) + side * (dy >= side ? dy - side : dy);
        q_ff[u] = cosq[dsp];
        q_gu_ij[u] = Gug[(size_t)j * N + i]; q_gd_ij[u] = Gdg[(size_t)j * N + i];
        q_gu_ji[u] = Gug[(size_t)i * N + j]; q_gd_ji[u] = Gdg[(size_t)i * N + j];
      }
#pragma unroll
      for (int u = 0; u < MU; ++u) {
        const int p = p0 + u * NT;
        if (p >= N * N) break;
        const int j = dN.div(p), i = p - j * N;
        const double ff = q_ff[u];
        const double gu_ij = q_gu_ij[u], gd_ij = q_gd_ij[u], gu_ji = q_gu_ji[u], gd_ji = q_gd_ji[u];
        const double dij = (i == j) ? 1.0 : 0.0;
        const double guc_ij = dij - gu_ji, gdc_ij = dij - gd_ji;
        const double guc_ii = 1.0 - s.du[i], gdc_ii = 1.0 - s.dd[i];
        const double guc_jj = 1.0 - s.du[j], gdc_jj = 1.0 - s.dd[j];
        v[4] += (gu_ji + gd_ji) * ff;
        const double common = guc_ii * guc_jj + guc_ij * gu_ij + gdc_ii * gdc_jj + gdc_ij * gd_ij;
        const double cross = gdc_ii * guc_jj + guc_ii * gdc_jj;
        v[5] += ff * (common - cross);
        v[6] += ff * (common + cross);
        v[7] += guc_ij * gdc_ij;
      }
    }
    wg_sum8(v, 8, s);
    if (tid() == 0) {
        const double Nd = (double)N;
        AS_G double* acc = gptr(d.acc);
        acc[0] += cs;
        acc[1] += cs * (2 - v[0] / Nd);
        acc[2] += cs * v[1] / Nd;
        acc[3] += cs * (2 * d.hop_t) * v[2] / Nd;
        acc[4] += cs * (1 - 0.5 * v[4] / Nd);
        acc[5] += cs * v[3] / Nd;
        acc[6] += cs * v[5] / (Nd * Nd);
        acc[7] += cs * v[6] / (Nd * Nd);
        acc[8] += cs * v[7] / Nd;
    }
    __syncthreads();
}

__device__ void wg_measure(Desc d, double cs, Lds&) { Lds& s = g_lds; wg_measure_from(d, d.G, cs, s); }

__device__ void record_slot(Desc d, int slot, double sign, bool with_sign) {
    Lds& s = g_lds;
    const int N = d.N;
    const size_t n2 = (size_t)N * N;
    const int lo = uni(s.sp_lo), hi = uni(s.sp_hi);   // this workgroup's spins
    AS_G double* dst = gptr(d.recG + ((size_t)slot * d.nspin + lo) * n2);
    const AS_G double* src = gptr(d.G + lo * n2);
    wg_stream((int)(n2 * (hi - lo)), [&](int e) { return src[e]; }, [&](int e, double v) { dst[e] = v; });
    if (with_sign && tid() == 0 && lo == 0) gptr(d.recSign)[slot] = sign;
    __syncthreads();
}

// the status bits of a recompute's error |G_fresh - G_wrapped| (NaN or inf when either held
// non-finite values) and the chain's running maximum (thread 0)
__device__ __forceinline__ void note_error(Desc d, ChainState& st, double err, bool track_error) {
    Lds& s = g_lds;
    if (!(err <= 1.7976931348623157e308)) s.status |= DQMC_STATUS_NONFINITE;
    if (!track_error) return;
    if (err > d.wrap_limit) s.status |= DQMC_STATUS_WRAP_LIMIT;
    if (!(err <= st.wrap_error)) st.wrap_error = err;
}

// fresh G for both spins into Gw, wrap error against G, then G <- Gw
__device__ __noinline__ void stabilize_green(Desc d, int lslot, int rslot, bool track_error, ChainState& st, Lds&) {
    Lds& s = g_lds;
    const int N = d.N;
    const size_t n2 = (size_t)N * N;
    const int lo = uni(s.sp_lo), hi = uni(s.sp_hi);   // this workgroup's spins
    double err = 0.0;
    for (int sp = lo; sp < hi; ++sp) {
        double* Gf = d.Gw + sp * n2;
        wg_fresh_green(d, sp, lslot, rslot, Gf, s);
        const AS_G double* Gfg = gptr(Gf);
        const AS_G double* Gg = gptr(d.G + sp * n2);
        wg_stream((int)n2, [&](int e) { return fabs(Gfg[e] - Gg[e]); }, [&](int, double v) { err = nan_max(err, v); });
    }
    err = wg_max(err, s);
    if (tid() == 0) note_error(d, st, err, track_error);
    wg_copy(d.G + lo * n2, d.Gw + lo * n2, (int)(n2 * (hi - lo)));
}

// ---------------------------------------------------------------------------
// Time-displaced Green's functions (src/dqmc_walker.cpp:424-524).  With the left
// stack B(t,0) = U_L D_L T_L and the right stack B(beta,t)^T = U_R D_R T_R (U
// orthogonal, T = V^T of any conditioning) and D = Dmax Dmin split at 1:
//   1 + B(t,0) B(beta,t) = U_L D_Lmax H D_Rmax U_R^T,
//   H = D_Lmax^-1 U_L^T U_R D_Rmax^-1 + D_Lmin T_L T_R^T D_Rmin,
// one LU of H with 2N right-hand sides H [X1 X2] = [D_Lmax^-1 U_L^T, D_Lmin T_L] gives
//   G(t,t) = U_R D_Rmax^-1 X1,  G(t,0) = G(t,t) B(t,0) = U_R D_Rmax^-1 X2,
//   G(0,t) = -B(beta,t) G(t,t) = -T_R^T D_Rmin X1.
// At t = beta (right stack empty), H = Db U^T + Ds T with Db = 1/max(D,1), Ds =
// min(D,1): H [X1 X2] = [Db U^T, Ds T] gives G(beta,beta) = X1, G(beta,0) = X2
// and G(0,beta) = -X1 (include/utils/numerical_stable.hpp:159-179, :246-325 with
// V^T written as T).  The wrap errors of G(t,0) and G(0,t) are collected
// (src/dqmc_walker.cpp:486-492); the fresh G(t,t) replaces the wrapped one.
// ---------------------------------------------------------------------------
__device__ __noinline__ void dynamic_fresh(Desc d, int lslot, int rslot, ChainState& st, Lds&) {
    Lds& s = g_lds;
    const int N = d.N;
    const size_t n2 = (size_t)N * N;
    double* W0 = d.W;
    double* W1 = d.W + n2;
    double* W2 = d.W + 2 * n2;
    double err = 0.0;
    for (int sp = 0; sp < d.nspin; ++sp) {
        const double* UL = slotU(d, sp, lslot); const double* TL = slotT(d, sp, lslot); const double* DL = slotD(d, sp, lslot);
        double* G = d.G + sp * n2;
        double* Gw = d.Gw + sp * n2;
        if (rslot >= 0) {
            const double* UR = slotU(d, sp, rslot); const double* TR = slotT(d, sp, rslot); const double* DR = slotD(d, sp, rslot);
            for (int i = tid(); i < N; i += NT) {
                const double dl = gptr(DL)[i], dr = gptr(DR)[i];
                s.v0[i] = dl >= 1.0 ? 1.0 / dl : 1.0; s.v1[i] = dl >= 1.0 ? 1.0 : dl;
                s.v2[i] = dr >= 1.0 ? 1.0 / dr : 1.0; s.v3[i] = dr >= 1.0 ? 1.0 : dr;
            }
            __syncthreads();
            gemm64(N, N, N, UL, N, true, UR, N, false, W0, N, s.v0, nullptr, s.v2, 0.0, s);
            gemm64(N, N, N, TL, N, false, TR, N, true, W0, N, s.v1, nullptr, s.v3, 1.0, s);
            wg_transpose_scaled(W1, nullptr, UL, nullptr, N, s.v0, nullptr, s);
        } else {
            for (int i = tid(); i < N; i += NT) {
                const double x = gptr(DL)[i];
                s.v0[i] = x > 1.0 ? 1.0 / x : 1.0;
                s.v1[i] = x > 1.0 ? 1.0 : x;
            }
            __syncthreads();
            wg_transpose_scaled(W1, W0, UL, TL, N, s.v0, s.v1, s);
        }
        {   // W2 = D_Lmin T_L (Ds T at t = beta): the second block of right-hand sides
            const AS_G double* src = gptr(TL);
            AS_G double* dst = gptr(W2);
            const AS_L double* v1 = (const AS_L double*)s.v1;
            const FastDiv dN(N);
            wg_stream((int)n2, [&](int e) { return src[e]; }, [&](int e, double v) { dst[e] = v1[e - dN.div(e) * N] * v; });
            __syncthreads();
        }
        wg_lu_solve(W0, W1, N, 2 * N, s);
        const AS_G double* x1 = gptr(W1);
        const AS_G double* x2 = gptr(W2);
        const AS_G double* f0t = x1;   // fresh G(0,t) = -f0t
        const AS_G double* ft0 = x2;   // fresh G(t,0)
        if (rslot >= 0) {
            const double* UR = slotU(d, sp, rslot); const double* TR = slotT(d, sp, rslot);
            gemm64(N, N, N, UR, N, false, W2, N, false, Gw, N, nullptr, s.v2, nullptr, 0.0, s);
            gemm64(N, N, N, TR, N, true, W1, N, false, W0, N, nullptr, s.v3, nullptr, 0.0, s);
            gemm64(N, N, N, UR, N, false, W1, N, false, G, N, nullptr, s.v2, nullptr, 0.0, s);
            ft0 = gptr(Gw);
            f0t = gptr(W0);
        } else {
            wg_copy(G, W1, (int)n2);
        }
        AS_G double* gt0 = gptr(d.Gt0 + sp * n2);
        AS_G double* g0t = gptr(d.G0t + sp * n2);
        wg_stream2((int)n2, [&](int e) { return ft0[e]; },
                   [&](int e, double v) { err = nan_max(err, fabs(v - gt0[e])); gt0[e] = v; },
                   [&](int e) { return f0t[e]; },
                   [&](int e, double v) { err = nan_max(err, fabs(-v - g0t[e])); g0t[e] = -v; });
        __syncthreads();
    }
    err = wg_max(err, s);
    if (tid() == 0) note_error(d, st, err, true);
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Dynamic measurements (src/measure/measure_methods.cpp:264-461), accumulated
// while the time-displaced sweep runs.  Entry t >= 1 of the reference's loop
// reads slot t-1, i.e. the matrices current at slice t; entry 0 reads slot L-1
// (slice L, measured after the last slice).  Terms that need G(0,0) = G(L-1,L-1)
// (the uncorrelated parts of the spin susceptibility and of rho_s) are kept as
// per-slice vectors and closed by dyn_finalize once slot L-1 exists:
//   DSS(t) = cs/(4N) sum_i [d_t(i) d_0(i) - sum_spin G(0,t)_ii G(t,0)_ii],
//     d_t(i) = G(t,t)_dn,ii - G(t,t)_up,ii,
//   rho_s  = t^2 cs/(4N^2) [-sum_ij f(i,j) A(j) A_0(i) + sum_t C_t],
//     A(j) = sum_t sum_spin (G(t,t)(j,j+x) - G(t,t)(j+x,j)), A_0 its slice-L term,
//     C_t the correlated four-term sum of the reference, f = cos(kx rx) - cos(kx ry).
// G(k,t) and D(t) are linear in the displacement sums gr[t][r] (the host applies
// the Fourier factors of the momenta it is asked for).
// ---------------------------------------------------------------------------
// site coordinates without integer division (side <= 21, i < 448: the multiply-shift
// quotient is exact since i (ceil(2^16/side) - 2^16/side) / 2^16 < 1/side)
struct SiteMap {
    int side, mg;
    __device__ explicit SiteMap(int s) : side(s), mg((65536 + s - 1) / s) {}
    __device__ __forceinline__ int y(int i) const { return (i * mg) >> 16; }
    __device__ __forceinline__ int x(int i) const { return i - y(i) * side; }
    __device__ __forceinline__ int wrap(int v) const { return v >= side ? v - side : (v < 0 ? v + side : v); }
    __device__ __forceinline__ int plus_x(int i) const { const int yy = y(i), xx = i - yy * side; return wrap(xx + 1) + side * yy; }
    // displacement index of j seen from i (Lattice::Square::Displacement, src/lattice/square.cpp:88-105)
    __device__ __forceinline__ int disp(int i, int j) const {
        const int yi = y(i), yj = y(j);
        return wrap(j - yj * side - (i - yi * side)) + side * wrap(yj - yi);
    }
    __device__ __forceinline__ int shift(int i, int r) const {
        const int yi = y(i), yr = y(r);
        return wrap(i - yi * side + r - yr * side) + side * wrap(yi + yr);
    }
};

__device__ __noinline__ void dyn_measure_entry(Desc d, int t, double cs, Lds&) {
    Lds& s = g_lds;
    const int N = uni(d.N), L = d.L, ns = uni(d.nspin);
    const SiteMap sm(uni(d.side));
    const size_t n2 = (size_t)N * N;
    const double spin_w = ns == 2 ? 1.0 : 2.0;   // attractive chains carry one spin: G_dn = G_up
    AS_G double* acc = gptr(d.dacc);
    AS_G double* delta = gptr(d.dscr + (size_t)t * N);
    AS_G double* abar = gptr(d.dscr + (size_t)L * N);
    const AS_G double* gtt = gptr(d.G);
    const AS_G double* gt0 = gptr(d.Gt0);
    const AS_G double* g0t = gptr(d.G0t);
    double v[2] = {0.0, 0.0};   // sum_i sum_spin G(0,t)_ii G(t,0)_ii, correlated stiffness sum
    for (int i = tid(); i < N; i += NT) {
        const int ipx = sm.plus_x(i);
        const double du = gtt[(size_t)i * N + i];
        const double dd = ns == 2 ? gtt[n2 + (size_t)i * N + i] : du;
        delta[i] = dd - du;
        double a = 0.0, c = 0.0;
        for (int sp = 0; sp < ns; ++sp) {
            const size_t o = sp * n2;
            a += gtt[o + (size_t)ipx * N + i] - gtt[o + (size_t)i * N + ipx];
            c += g0t[o + (size_t)i * N + i] * gt0[o + (size_t)i * N + i];
        }
        abar[i] += spin_w * a;
        v[0] += spin_w * c;
    }
    // G(t,0) columns and G(0,t) rows of one lattice row of sites fit in LDS up to side 18
    const int side = sm.side;
    const bool staged = 2 * side * N <= PANEL_DOUBLES;
    for (int i = tid(); i < N; i += NT) s.v0[i] = gptr(d.rhoq)[i];
    __syncthreads();
    const AS_L double* rq = (const AS_L double*)s.v0;
    // p / N by multiply-high: exact for p < N^2 <= 448^2 since p (ceil(2^32/N) - 2^32/N) / 2^32 < 1/N
    const unsigned mgN = (unsigned)((0x100000000ULL + N - 1) / N);
    // displacement sums of the spin-averaged G(t,0) (G(0,0) at entry 0): thread r owns gr[t][r]
    // and walks the columns, eight of them per round with their loads in flight together
    if (!staged || t == 0) {
        const AS_G double* src = t == 0 ? gtt : gt0;
        for (int r = tid(); r < N; r += NT) {
            double g = 0.0;
            constexpr int U = 8;
            for (int i0 = 0; i0 < N; i0 += U) {
                double q[2 * U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int i = min(i0 + u, N - 1);
                    const size_t e = (size_t)i * N + sm.shift(i, r);
                    q[2 * u] = src[e];
                    q[2 * u + 1] = ns == 2 ? src[n2 + e] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (i0 + u < N) g += q[2 * u] + q[2 * u + 1];
            }
            acc[(size_t)t * N + r] += cs * (ns == 2 ? 0.5 * g : g);
        }
    }
    if (staged) {
        // per block of R lattice rows (sites i0 .. i0 + R side - 1, closed under i -> i + x):
        // A[c N + j] = G(t,0)(j, i0 + c) (contiguous columns), B[c N + b] = G(0,t)(i0 + c, b)
        // (pieces of every column), so each site pair reads both matrices from LDS; for entries
        // t >= 1 the same columns give the displacement sums
        const int rows = max(1, min(side, PANEL_DOUBLES / (2 * side * N)));
        const unsigned mgS = (unsigned)((0x100000000ULL + side - 1) / side);   // e / side for e < 2^16
        double gb[2] = {0.0, 0.0};   // bins r = tid and tid + NT of entry t (N <= 448 < 2 NT)
        for (int sp = 0; sp < ns; ++sp) {
            const AS_G double* gt = gt0 + sp * n2;
            const AS_G double* gz = g0t + sp * n2;
            for (int y0 = 0; y0 < side; y0 += rows) {
                const int i0 = y0 * side, nc = min(rows, side - y0) * side, sN = nc * N;
                AS_L double* A = (AS_L double*)s.panel;
                AS_L double* B = A + sN;
                const unsigned mgC = (unsigned)((0x100000000ULL + nc - 1) / nc);   // e / nc for e < 2^16
                wg_stream2(sN, [&](int e) { return gt[(size_t)i0 * N + e]; }, [&](int e, double x) { A[e] = x; },
                           [&](int e) {
                               const int bcol = (int)__umulhi((unsigned)e, mgC), c = e - bcol * nc;
                               return gz[(size_t)bcol * N + i0 + c];
                           },
                           [&](int e, double x) {
                               const int bcol = (int)__umulhi((unsigned)e, mgC), c = e - bcol * nc;
                               B[c * N + bcol] = x;
                           });
                __syncthreads();
                for (int e = tid(); e < sN; e += NT) {
                    const int c = (int)__umulhi((unsigned)e, mgN), j = e - c * N;
                    const int cx = c - (int)__umulhi((unsigned)c, mgS) * side;
                    const int cp = cx + 1 == side ? c + 1 - side : c + 1;
                    const int jpx = sm.plus_x(j);
                    const double f = rq[sm.disp(i0 + c, j)];
                    // -g0t(ipx,jpx) gt0(j,i) + g0t(i,jpx) gt0(j,ipx) + g0t(ipx,j) gt0(jpx,i) - g0t(i,j) gt0(jpx,ipx)
                    const double w = -B[cp * N + jpx] * A[c * N + j] + B[c * N + jpx] * A[cp * N + j]
                                     + B[cp * N + j] * A[c * N + jpx] - B[c * N + j] * A[cp * N + jpx];
                    v[1] += spin_w * f * w;
                }
                if (t > 0) {
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const int r = tid() + q * NT;
                        if (r < N)
                            for (int c = 0; c < nc; ++c) gb[q] += A[c * N + sm.shift(i0 + c, r)];
                    }
                }
                __syncthreads();
            }
        }
        if (t > 0) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int r = tid() + q * NT;
                if (r < N) acc[(size_t)t * N + r] += cs * (ns == 2 ? 0.5 * gb[q] : gb[q]);
            }
        }
    } else {
        // G(0,t)^T of each spin into W, so that the stiffness pairs read both matrices along columns
        for (int sp = 0; sp < ns; ++sp)
            wg_transpose_scaled(d.W + sp * n2, nullptr, d.G0t + sp * n2, nullptr, N, nullptr, nullptr, s);
        for (int sp = 0; sp < ns; ++sp) {
            const AS_G double* g0tT = gptr(d.W + sp * n2);   // g0t(a,b) at a*N + b
            const AS_G double* gt = gt0 + sp * n2;            // gt0(a,b) at b*N + a
            constexpr int MU = 4;
            for (int p0 = tid(); p0 < N * N; p0 += MU * NT) {
                double q[MU][8];
                double f[MU];
#pragma unroll
                for (int u = 0; u < MU; ++u) {
                    const int p = min(p0 + u * NT, N * N - 1);
                    const int i = (int)__umulhi((unsigned)p, mgN), j = p - i * N;
                    const int ipx = sm.plus_x(i), jpx = sm.plus_x(j);
                    f[u] = rq[sm.disp(i, j)];
                    q[u][0] = g0tT[(size_t)ipx * N + jpx]; q[u][1] = gt[(size_t)i * N + j];
                    q[u][2] = g0tT[(size_t)i * N + jpx];   q[u][3] = gt[(size_t)ipx * N + j];
                    q[u][4] = g0tT[(size_t)ipx * N + j];   q[u][5] = gt[(size_t)i * N + jpx];
                    q[u][6] = g0tT[(size_t)i * N + j];     q[u][7] = gt[(size_t)ipx * N + jpx];
                }
#pragma unroll
                for (int u = 0; u < MU; ++u)
                    if (p0 + u * NT < N * N)
                        v[1] += spin_w * f[u] * (-q[u][0] * q[u][1] + q[u][2] * q[u][3] + q[u][4] * q[u][5] - q[u][6] * q[u][7]);
            }
        }
    }
    wg_sum8(v, 2, s);
    if (tid() == 0) {
        const double Nd = (double)N;
        acc[(size_t)L * N + t] -= 0.25 * cs / Nd * v[0];
        acc[(size_t)L * N + L] += 0.25 * d.hop_t * d.hop_t * cs * v[1] / (Nd * Nd);
    }
    __syncthreads();
}

// the G(0,0)-dependent parts, once entry 0 (slot L-1) is measured; then sign and count
__device__ __noinline__ void dyn_finalize(Desc d, double cs, Lds&) {
    Lds& s = g_lds;
    const int N = uni(d.N), L = d.L, ns = uni(d.nspin);
    const SiteMap sm(uni(d.side));
    const size_t n2 = (size_t)N * N;
    const double spin_w = ns == 2 ? 1.0 : 2.0;
    AS_G double* acc = gptr(d.dacc);
    const AS_G double* delta = gptr(d.dscr);
    const AS_G double* gtt = gptr(d.G);
    // A_0(i) into v0, the summed current vector into v1, d_0 into v2, the stiffness factors into v3
    for (int i = tid(); i < N; i += NT) {
        const int ipx = sm.plus_x(i);
        double a = 0.0;
        for (int sp = 0; sp < ns; ++sp) a += gtt[sp * n2 + (size_t)ipx * N + i] - gtt[sp * n2 + (size_t)i * N + ipx];
        s.v0[i] = spin_w * a;
        s.v1[i] = gptr(d.dscr)[(size_t)L * N + i];
        s.v2[i] = delta[i];
        s.v3[i] = gptr(d.rhoq)[i];
    }
    __syncthreads();
    double v[1] = {0.0};
    for (int i = tid(); i < N; i += NT) {
        double f = 0.0;
        for (int j = 0; j < N; ++j) f += s.v3[sm.disp(i, j)] * s.v1[j];
        v[0] -= s.v0[i] * f;
    }
    // uncorrelated spin susceptibility, one slice entry per thread
    for (int t = tid(); t < L; t += NT) {
        double u = 0.0;
        constexpr int U = 8;
        for (int i0 = 0; i0 < N; i0 += U) {
            double q[U];
#pragma unroll
            for (int k = 0; k < U; ++k) q[k] = delta[(size_t)t * N + min(i0 + k, N - 1)];
#pragma unroll
            for (int k = 0; k < U; ++k)
                if (i0 + k < N) u += q[k] * s.v2[i0 + k];
        }
        acc[(size_t)L * N + t] += 0.25 * cs / (double)N * u;
    }
    wg_sum8(v, 1, s);
    if (tid() == 0) {
        const double Nd = (double)N;
        acc[(size_t)L * N + L] += 0.25 * d.hop_t * d.hop_t * cs * v[0] / (Nd * Nd);
        acc[(size_t)L * N + L + 1] += cs;
        acc[(size_t)L * N + L + 2] += 1.0;
    }
    __syncthreads();
}

__device__ void record_dynamic(Desc d, int slot) {
    const size_t n = (size_t)d.nspin * d.N * d.N;
    wg_copy(d.recGt0 + slot * n, d.Gt0, (int)n);
    wg_copy(d.recG0t + slot * n, d.G0t, (int)n);
}

// the equal-time measurement of the current slice; in a pair launch the partners meet while
// spin 0's workgroup reads both spins' G (and before either changes its G again)
__device__ void measure_slice(Desc d, double cs, Lds&) {
    Lds& s = g_lds;
    if (!uni(s.pair)) { wg_measure(d, cs, s); return; }
    pair_exchange(d, s);
    if (uni(s.me) == 0) wg_measure(d, cs, s);
    pair_exchange(d, s);
}

// per-phase device time (thread 0 reads the 100 MHz wall clock at barriers)
struct PhaseClock {
    bool on;
    unsigned long long t;
    __device__ explicit PhaseClock(bool enabled) : on(enabled), t(enabled ? wall_clock64() : 0) {}
    __device__ void lap(ChainState& st, int ph) {
        if (!on) return;
        const unsigned long long n = wall_clock64();
        if (tid() == 0) st.phase[ph] += n - t;
        t = n;
    }
};

// ---------------------------------------------------------------------------
// sweeps (src/dqmc_walker.cpp:246-325 and :334-412)
// ---------------------------------------------------------------------------
__device__ void sweep_up(Desc d, ChainState& st, uint64_t& rng, double& sign, int flags, int& nacc, Lds&) {
    Lds& s = g_lds;
    const int N = d.N, L = d.L;
    const size_t n2 = (size_t)N * N;
    const bool rec = flags & DQMC_FLAG_RECORD, meas = flags & DQMC_FLAG_MEASURE;
    const int lo = uni(s.sp_lo), hi = uni(s.sp_hi);   // this workgroup's spins
    PhaseClock clk(flags & DQMC_FLAG_PROFILE);
    for (int sp = lo; sp < hi; ++sp) wg_identity(d.Tmp + sp * n2, N);
    for (int t = 1; t <= L; ++t) {
        const int e = t - 1;
        load_field_row(d, e, s);
        for (int sp = lo; sp < hi; ++sp) wg_wrap(d, sp, true, d.G + sp * n2, d.Gw + sp * n2, s);
        clk.lap(st, DQMC_PH_WRAP);
        if (!(flags & DQMC_FLAG_FROZEN)) {
            const WalkOut o = wg_metropolis(d, e, rng, sign, rec, s);
            rng = o.rng;
            sign = o.sign;
            nacc += o.accepted;
        }
        clk.lap(st, DQMC_PH_METROPOLIS);
        refresh_scalings(d, s);
        for (int sp = lo; sp < hi; ++sp) wg_bmul(d, 0, sp, d.Tmp + sp * n2, d.Gw + sp * n2, s);
        clk.lap(st, DQMC_PH_PRODUCT);
        if (t % d.pace == 0 || t == L) {
            st.n_right -= 1;
            const int slot = st.n_right;
            for (int sp = lo; sp < hi; ++sp)
                wg_push(d, sp, slot, slot + 1 < d.n_stack ? slot + 1 : -1, d.Tmp + sp * n2, s);
            clk.lap(st, DQMC_PH_PUSH);
            stabilize_green(d, slot, slot > 0 ? slot - 1 : -1, true, st, s);
            for (int sp = lo; sp < hi; ++sp) wg_identity(d.Tmp + sp * n2, N);
            clk.lap(st, DQMC_PH_FRESH);
        }
        if (rec) record_slot(d, e, sign, true);
        if (meas) measure_slice(d, sign, s);
        clk.lap(st, DQMC_PH_MEASURE);
    }
    st.pos = L;
}

__device__ void sweep_down(Desc d, ChainState& st, uint64_t& rng, double& sign, int flags, int& nacc, Lds&) {
    Lds& s = g_lds;
    const int N = d.N, L = d.L;
    const size_t n2 = (size_t)N * N;
    const bool rec = flags & DQMC_FLAG_RECORD, meas = flags & DQMC_FLAG_MEASURE;
    const int lo = uni(s.sp_lo), hi = uni(s.sp_hi);   // this workgroup's spins
    PhaseClock clk(flags & DQMC_FLAG_PROFILE);
    for (int sp = lo; sp < hi; ++sp) wg_identity(d.Tmp + sp * n2, N);
    double sign_top = sign;
    for (int t = L; t >= 1; --t) {
        const int e = t - 1;
        if (t % d.pace == 0 && t != L) {
            const int slot = st.n_right;
            for (int sp = lo; sp < hi; ++sp)
                wg_push(d, sp, slot, slot > 0 ? slot - 1 : -1, d.Tmp + sp * n2, s);
            st.n_right += 1;
            clk.lap(st, DQMC_PH_PUSH);
            stabilize_green(d, slot + 1 < d.n_stack ? slot + 1 : -1, slot, true, st, s);
            for (int sp = lo; sp < hi; ++sp) wg_identity(d.Tmp + sp * n2, N);
            clk.lap(st, DQMC_PH_FRESH);
        }
        load_field_row(d, e, s);
        if (!(flags & DQMC_FLAG_FROZEN)) {
            const WalkOut o = wg_metropolis(d, e, rng, sign, rec, s);
            rng = o.rng;
            sign = o.sign;
            nacc += o.accepted;
        }
        clk.lap(st, DQMC_PH_METROPOLIS);
        if (rec) record_slot(d, e, sign, true);
        if (t == L) sign_top = sign;
        else if (meas) measure_slice(d, sign, s);
        clk.lap(st, DQMC_PH_MEASURE);
        refresh_scalings(d, s);
        for (int sp = lo; sp < hi; ++sp) wg_bmul(d, 4, sp, d.Tmp + sp * n2, d.Gw + sp * n2, s);
        clk.lap(st, DQMC_PH_PRODUCT);
        for (int sp = lo; sp < hi; ++sp) wg_wrap(d, sp, false, d.G + sp * n2, d.Gw + sp * n2, s);
        clk.lap(st, DQMC_PH_WRAP);
    }
    {
        const int slot = st.n_right;
        for (int sp = lo; sp < hi; ++sp) wg_push(d, sp, slot, slot > 0 ? slot - 1 : -1, d.Tmp + sp * n2, s);
        st.n_right += 1;
        stabilize_green(d, -1, slot, false, st, s);
    }
    if (rec) record_slot(d, L - 1, sign, false);
    if (meas) measure_slice(d, sign_top, s);
    st.pos = 0;
}

// DqmcWalker::sweep_for_dynamic_greens (src/dqmc_walker.cpp:424-524): from slice 0 (left stack
// empty) to beta with the fields fixed.  G(t,t) is wrapped as in the upward sweep, G(t,0) <- B(t)
// G(t-1,0) and G(0,t) <- G(0,t-1) B(t)^-1 from G(0,0) and G(0,0) - 1, the slice product
// accumulates for the left stack, and at every stabilization the three are recomputed from the
// stacks (dynamic_fresh).  No equal-time measurement runs in this direction (src/dqmc.cpp:54-57).
__device__ void sweep_dynamic(Desc d, ChainState& st, double sign, int flags, Lds&) {
    Lds& s = g_lds;
    const int N = d.N, L = d.L;
    const size_t n2 = (size_t)N * N;
    const bool rec = flags & DQMC_FLAG_RECORD, meas = flags & DQMC_FLAG_MEASURE;
    PhaseClock clk(flags & DQMC_FLAG_PROFILE);
    const FastDiv dN(N);
    for (int sp = 0; sp < d.nspin; ++sp) {
        const AS_G double* g = gptr(d.G + sp * n2);
        AS_G double* a = gptr(d.Gt0 + sp * n2);
        AS_G double* b = gptr(d.G0t + sp * n2);
        wg_stream((int)n2, [&](int e) { return g[e]; }, [&](int e, double v) {
            a[e] = v;
            b[e] = (e == dN.div(e) * (N + 1)) ? v - 1.0 : v;
        });
        wg_identity(d.Tmp + sp * n2, N);
    }
    if (meas)
        for (int i = tid(); i < N; i += NT) gptr(d.dscr)[(size_t)L * N + i] = 0.0;
    __syncthreads();
    for (int t = 1; t <= L; ++t) {
        const int e = t - 1;
        load_field_row(d, e, s);
        for (int sp = 0; sp < d.nspin; ++sp) wg_wrap(d, sp, true, d.G + sp * n2, d.Gw + sp * n2, s);
        clk.lap(st, DQMC_PH_WRAP);
        for (int sp = 0; sp < d.nspin; ++sp) {
            wg_bmul(d, 0, sp, d.Gt0 + sp * n2, d.Gw + sp * n2, s, true);
            wg_bmul(d, 3, sp, d.G0t + sp * n2, d.Gw + sp * n2, s, true);
            wg_bmul(d, 0, sp, d.Tmp + sp * n2, d.Gw + sp * n2, s);
        }
        clk.lap(st, DQMC_PH_PRODUCT);
        if (t % d.pace == 0 || t == L) {
            st.n_right -= 1;
            const int slot = st.n_right;
            for (int sp = 0; sp < d.nspin; ++sp)
                wg_push(d, sp, slot, slot + 1 < d.n_stack ? slot + 1 : -1, d.Tmp + sp * n2, s);
            clk.lap(st, DQMC_PH_PUSH);
            dynamic_fresh(d, slot, slot > 0 ? slot - 1 : -1, st, s);
            for (int sp = 0; sp < d.nspin; ++sp) wg_identity(d.Tmp + sp * n2, N);
            clk.lap(st, DQMC_PH_FRESH);
        }
        if (rec) {
            record_slot(d, e, sign, true);
            record_dynamic(d, e);
        }
        if (meas) dyn_measure_entry(d, t == L ? 0 : t, sign, s);
        clk.lap(st, DQMC_PH_MEASURE);
    }
    if (meas) dyn_finalize(d, sign, s);
    clk.lap(st, DQMC_PH_MEASURE);
    st.pos = L;
}

__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, 1))) k_init(const ChainDesc* __restrict__ descs) {
    Lds& s = g_lds;
    Desc d = ((const AS_C ChainDesc*)descs)[blockIdx.x];
    ChainState st = *d.st;
    const int N = d.N, L = d.L;
    const size_t n2 = (size_t)N * N;
    if (tid() == 0) { s.prof_on = 0; s.status = 0; s.pair = 0; s.me = 0; s.sp_lo = 0; s.sp_hi = d.nspin; }
    __syncthreads();
    for (int sp = 0; sp < d.nspin; ++sp) wg_identity(d.Tmp + sp * n2, N);
    st.n_right = 0;
    for (int t = L; t >= 1; --t) {   // src/dqmc_walker.cpp:118-133
        load_field_row(d, t - 1, s);
        for (int sp = 0; sp < d.nspin; ++sp) wg_bmul(d, 4, sp, d.Tmp + sp * n2, d.Gw + sp * n2, s);
        if ((t - 1) % d.pace == 0) {
            const int slot = st.n_right;
            for (int sp = 0; sp < d.nspin; ++sp) wg_push(d, sp, slot, slot > 0 ? slot - 1 : -1, d.Tmp + sp * n2, s);
            st.n_right += 1;
            for (int sp = 0; sp < d.nspin; ++sp) wg_identity(d.Tmp + sp * n2, N);
        }
    }
    // G(0) from the right stack, and the configuration sign from the same factorization: with the
    // stack product M = U D T, det(1 + M) = det(U) det(Db)^-1 det(Db U^T + Ds T), so sgn det G(0) =
    // sgn det U x sgn det H, H the matrix the recompute factors.  (The reference takes sgn det G,
    // src/dqmc_walker.cpp:162, which is rounding noise once G has eigenvalues below FP64
    // resolution; both are well-conditioned here.)
    double sg = 1.0;
    for (int sp = 0; sp < d.nspin; ++sp) {
        const double dh = wg_fresh_green(d, sp, -1, st.n_right - 1, d.G + sp * n2, s);
        wg_copy(d.Gw, slotU(d, sp, st.n_right - 1), (int)n2);
        const double du = wg_lu_solve(d.Gw, nullptr, N, 0, s);
        sg *= dh * du;
        if (d.nspin == 1) sg *= dh * du;   // the attractive model's G_dn is G_up
    }
    {   // a non-finite G(0) (e.g. from overflowing products) is a failed initialization
        double m = 0.0;
        const AS_G double* g = gptr(d.G);
        wg_stream((int)(n2 * d.nspin), [&](int e) { return fabs(g[e]); }, [&](int, double v) { m = nan_max(m, v); });
        m = wg_max(m, s);
        if (tid() == 0 && !(m <= 1.7976931348623157e308)) s.status |= DQMC_STATUS_NONFINITE;
    }
    if (tid() == 0) {
        st.sign = sg >= 0 ? 1.0 : -1.0;
        st.wrap_error = 0.0;
        st.pos = 0;
        st.status = s.status;
        st.accepted = 0;
        for (int q = 0; q < 8; ++q) st.phase[q] = 0;
        *d.st = st;
    }
}

// Joining a pair launch.  Its two workgroups are not guaranteed to be resident
// together (another stream's or process's kernels may hold the CUs, or the grid
// may exceed what the device holds), so they meet on the chain's join word first
// (0 nobody, 1 one arrived, 2 paired, 3 taken alone; zeroed by the host before
// the launch): the first to arrive waits up to PAIR_JOIN_TICKS for its partner
// and then claims the chain for itself; the late partner finds the claim and
// leaves.  A pair therefore starts only when both workgroups run, and once both
// run they stay resident to the end, so every later wait ends.  A chain alone
// computes the same numbers as a pair (the one-workgroup path), so results never
// depend on how the join went.  Returns 1 (paired), 0 (alone) or -1 (leave).
#define PAIR_JOIN_TICKS 10000ull   // 100 us of the 100 MHz wall clock
__device__ int pair_join(Desc d) {
    unsigned* w = d.xflag + 64;
    unsigned seen = 0u;
    if (__hip_atomic_compare_exchange_strong(w, &seen, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        const unsigned long long t0 = wall_clock64();   // first: wait for the partner
        while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 2u) {
            if (wall_clock64() - t0 > PAIR_JOIN_TICKS) {
                unsigned one = 1u;
                if (__hip_atomic_compare_exchange_strong(w, &one, 3u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT))
                    return 0;   // claimed: the chain runs here alone
                return 1;       // the partner arrived meanwhile
            }
            __builtin_amdgcn_s_sleep(2);
        }
        return 1;
    }
    if (seen == 1u) {   // second: join unless the first has claimed the chain
        unsigned one = 1u;
        if (__hip_atomic_compare_exchange_strong(w, &one, 2u, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return 1;
    }
    return -1;
}

// Workgroup -> (chain, spin).  One workgroup per chain, or in a pair launch
// (pair != 0, grid 16 x ceil(n_chains / 8)) chain 8 g + r and spin h at workgroup
// 16 g + 8 h + r (r < 8): the two workgroups of a chain are 8 apart, so under the
// round-robin dispatch over the 8 XCDs they share an XCD and its L2.  A chain that
// cannot pair (one spin, a time-displaced sweep) or whose join failed runs in one
// workgroup, which then writes the chain state; a paired chain's state is written
// by spin 0's workgroup.
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, 1))) k_sweep(const ChainDesc* __restrict__ descs, int n_sweeps, int flags,
                                                                                        int n_chains, int pair) {
    Lds& s = g_lds;
    int c = blockIdx.x, me = 0;
    if (pair) {
        c = (blockIdx.x >> 4) * 8 + (blockIdx.x & 7);
        me = (blockIdx.x >> 3) & 1;
        if (c >= n_chains) return;
    }
    Desc d = ((const AS_C ChainDesc*)descs)[c];
    ChainState st = *d.st;
    uint64_t rng = st.rng;
    double sign = st.sign;
    int nacc = 0;
    if (d.launch_sweeps > 0) n_sweeps = d.launch_sweeps;
    if (d.launch_flags >= 0) flags = d.launch_flags;
    if (pair) {
        // one spin or a time-displaced sweep: spin 0's workgroup runs the chain alone
        const bool can = d.nspin == 2 && !(flags & DQMC_FLAG_DYNAMIC);
        if (tid() == 0) s.join = can ? pair_join(d) : (me == 0 ? 0 : -1);
        __syncthreads();
        const int j = s.join;
        if (j < 0) return;
        pair = j;
        if (!pair) me = 0;
    }
    if (tid() == 0) {
        s.prof[0] = s.prof[1] = 0;
        s.prof_on = (flags & DQMC_FLAG_PROFILE) ? 1 : 0;
        s.status = st.status;
        s.pair = pair;
        s.me = me;
        s.sp_lo = pair ? me : 0;
        s.sp_hi = pair ? me + 1 : d.nspin;
        s.xseq = 0;
        s.xbroken = 0;
    }
    __syncthreads();
    // a chain that met a fatal numerical failure sweeps no further (every thread reads the
    // status after a barrier, so the whole workgroup leaves the loop together; the partners of
    // a pair launch merge their status bits at the end of every sweep, so they leave together)
    for (int k = 0; k < n_sweeps && !(s.status & DQMC_STATUS_FATAL); ++k) {
        if (st.pos == 0) {
            if (flags & DQMC_FLAG_DYNAMIC) sweep_dynamic(d, st, sign, flags, s);
            else sweep_up(d, st, rng, sign, flags, nacc, s);
        } else sweep_down(d, st, rng, sign, flags, nacc, s);
        __syncthreads();
        if (pair) pair_merge_state(d, st, s);
    }
    if (pair && me != 0) return;
    if (tid() == 0) {
        st.status = s.status;
        st.rng = rng;
        st.sign = sign;
        st.accepted = nacc;
        st.paired = pair;
        st.phase[DQMC_PH_WALK] += s.prof[0];
        st.phase[DQMC_PH_PANEL] += s.prof[1];
        *d.st = st;
    }
}

// MeasureHandler::equaltime_measure over the slices kept by a RECORD sweep
// (src/measure/measure_handler.cpp:136-150: every method loops over the L stored slices)
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, 1))) k_measure_recorded(const ChainDesc* __restrict__ descs) {
    Lds& s = g_lds;
    Desc d = ((const AS_C ChainDesc*)descs)[blockIdx.x];
    const size_t slot_size = (size_t)d.nspin * d.N * d.N;
    for (int slot = 0; slot < d.L; ++slot) wg_measure_from(d, d.recG + slot * slot_size, d.recSign[slot], s);
}

// one workgroup runs one GEMM variant (0 vector FP64, 1 MFMA FP64, 2 MFMA FP32, 3/4 MFMA FP64 with A / A and B
// staged in LDS, 5/6 MFMA FP64 on wide 128x128 / 160x160 tiles)
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, 1))) k_gemm(int variant, int M, int Nc, int K, const double* A, int lda, int tA,
                                             const double* B, int ldb, int tB, double* C, int ldc, const double* rs,
                                             const double* ms, const double* cs, double beta) {
    Lds& s = g_lds;
    if (variant == 0) { wg_gemm(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, s); return; }
    // the matrix-core GEMM takes its scalings from LDS
    if (ms) {
        for (int i = tid(); i < K; i += NT) s.v0[i] = ms[i];
        ms = s.v0;
    }
    if (rs) {
        for (int i = tid(); i < M; i += NT) s.v1[i] = rs[i];
        rs = s.v1;
    }
    if (cs) {
        for (int i = tid(); i < Nc; i += NT) s.v2[i] = cs[i];
        cs = s.v2;
    }
    // variants 3 and 4 stage A (and B) in LDS first, as the engine's in-LDS operands are
    const size_t na = (size_t)M * K, nb = (size_t)K * Nc;
    if (variant == 3 || variant == 4) {
        for (size_t e = tid(); e < na; e += NT) s.panel[e] = A[e];
        A = s.panel;
    }
    if (variant == 4) {
        for (size_t e = tid(); e < nb; e += NT) s.panel[na + e] = B[e];
        B = s.panel + na;
    }
    __syncthreads();
    if (variant == 1) wg_gemm_mfma<false, false, false>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, s);
    else if (variant == 2) wg_gemm_mfma<true, false, false>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, s);
    else if (variant == 3) wg_gemm_mfma<false, true, false>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, s);
    else if (variant == 4) wg_gemm_mfma<false, true, true>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, s);
    else if (variant == 5) wg_gemm_wide<4, 4, 16>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, 1.0);
    else wg_gemm_wide<5, 5, 8>(M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc, rs, ms, cs, beta, 1.0);
}

// The single-flip update of the equal-time Green's functions after the field at
// (t, i) flips (src/model/repulsive_hubbard.cpp:127-155,
// src/model/attractive_hubbard.cpp:128-156):
//   G <- G - gamma G[:,i] (e_i^T - G[i,:]),  gamma = m / (1 + (1 - G(i,i)) m),
// m = exp(-2 sigma alpha s) - 1 (d.dm1); the attractive model applies the spin-up
// gamma to both matrices.  The sweep applies the same update in blocks (the
// submatrix walk); this is its one-flip form.  The field row of t is in LDS.
__device__ void wg_rank1(Desc d, int i, double* Gu_, double* Gd_, Lds&) {
    Lds& s = g_lds;
    const int N = d.N;
    const int neg = s.frow[i] < 0 ? 1 : 0;
    const double m0 = d.dm1[0][neg];
    const double gam0 = m0 / (1.0 + (1.0 - gptr(Gu_)[(size_t)i * N + i]) * m0);
    for (int sp = 0; sp < 2; ++sp) {
        AS_G double* G = gptr(sp ? Gd_ : Gu_);
        if (!(sp ? Gd_ : Gu_)) continue;
        double gam = gam0;
        if (sp == 1 && d.model != 1) {
            const double m1 = d.dm1[1][neg];
            gam = m1 / (1.0 + (1.0 - G[(size_t)i * N + i]) * m1);
        }
        for (int a = tid(); a < N; a += NT) {
            s.v0[a] = G[(size_t)i * N + a];                        // column i
            s.v1[a] = (a == i ? 1.0 : 0.0) - G[(size_t)a * N + i];  // e_i - row i
        }
        __syncthreads();
        const int n2 = N * N;
        for (int e = tid(); e < n2; e += NT) G[e] -= gam * s.v0[e % N] * s.v1[e / N];
        __syncthreads();
    }
}

// Update ratio of flipping the field at (t, i) from the current G of both spins
// (RepulsiveHubbard::get_update_ratio, src/model/repulsive_hubbard.cpp:112-124;
// AttractiveHubbard::get_update_ratio, src/model/attractive_hubbard.cpp:112-125):
//   r = pref(s) [1 + (1 - G_up(i,i)) m_up(s)] [1 + (1 - G_dn(i,i)) m_dn(s)],
// the same expression the submatrix walk evaluates for its decisions.
__device__ double unit_ratio(Desc d, int t, int i, const double* Gu, const double* Gd) {
    const int N = d.N;
    const int f = ((const AS_G int8_t*)d.field)[(size_t)t * N + i];
    const bool up = f > 0;
    const double m0 = up ? d.dm1[0][0] : d.dm1[0][1], m1 = up ? d.dm1[1][0] : d.dm1[1][1];
    const double g0 = gptr(Gu)[(size_t)i * N + i];
    const double g1 = (d.nspin == 2 && Gd) ? gptr(Gd)[(size_t)i * N + i] : g0;
    const double x0 = 1.0 + (1.0 - g0) * m0, x1 = 1.0 + (1.0 - g1) * m1;
    return (up ? d.pref[0] : d.pref[1]) * x0 * x1;
}

// stack slot k of spin 0 <- (U, D, V^T) from caller buffers (U, V column-major N x N)
__device__ void unit_put_slot(Desc d, int k, const double* U, const double* D, const double* V, Lds&) {
    Lds& s = g_lds;
    const int N = uni(d.N);
    wg_copy(slotU(d, 0, k), U, N * N);
    for (int i = tid(); i < N; i += NT) gptr(slotD(d, 0, k))[i] = gptr(D)[i];
    wg_transpose_scaled(slotT(d, 0, k), nullptr, V, nullptr, N, nullptr, nullptr, s);
    __syncthreads();
}

// One building block of the sweep on chain 0 of a descriptor array (one workgroup):
//   DQMC_UNIT_BMUL   X <- B(a) X or its variants (op c: 0 B X, 1 X B, 2 B^-1 X, 3 X B^-1,
//                    4 B^T X; 5 B X B^-1, 6 B^-1 X B) for spin index b
//   DQMC_UNIT_RANK1  the single-flip update of X (spin up) and Y (spin down, may be null)
//                    for the field at (a, b)
//   DQMC_UNIT_PUSH   push X onto spin-0 stack slot b (-1: empty stack), result in slot a
//   DQMC_UNIT_FRESH  X = (1 + B_left B_right^T)^-1 from spin-0 stack slots a (left) and b
//                    (right); -1 marks an empty stack
//   DQMC_UNIT_RATIO  *ext.out = the update ratio of the field at (a, b) from X (up), Y (down)
//   DQMC_UNIT_FLIP   the field at (a, b) changes sign
//   DQMC_UNIT_PUSH_EXT   (ext.U, ext.D, ext.V) = push of ext.A onto the top (ext.U_in, D_in,
//                    V_in; U_in null: empty stack), through stack slots 0 and 1
//   DQMC_UNIT_FRESH_EXT  ext.G = (1 + B_L B_R^T)^-1 from the tops (ext.UL, DL, VL) and (ext.UR,
//                    DR, VR) (UL or UR null: that stack is empty), through slots 0 and 1
// The failure bits of PUSH, FRESH and their external forms go to the chain state and, for the
// external forms, to *ext.status when it is given.
struct UnitPtrs {
    const double *A, *U_in, *D_in, *V_in;
    double *U, *D, *V;
    const double *UL, *DL, *VL, *UR, *DR, *VR;
    double* out;     // the recompute's G, or the update ratio (one double)
    int* status;     // failure bits of the call (may be null)
};
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, 1))) k_unit(const ChainDesc* __restrict__ descs, int unit, int a, int b, int c,
                                             double* X, double* Y, UnitPtrs ext) {
    Lds& s = g_lds;
    Desc d = ((const AS_C ChainDesc*)descs)[0];
    if (tid() == 0) { s.prof_on = 0; s.status = 0; s.pair = 0; s.me = 0; s.sp_lo = 0; s.sp_hi = d.nspin; }
    __syncthreads();
    // a non-finite entry among the first n of M sets DQMC_STATUS_NONFINITE
    auto check_finite = [&](const double* M, int n) {
        double m = 0.0;
        const AS_G double* g = gptr(M);
        wg_stream(n, [&](int e) { return fabs(g[e]); }, [&](int, double v) { m = nan_max(m, v); });
        m = wg_max(m, s);
        if (tid() == 0 && !(m <= 1.7976931348623157e308)) s.status |= DQMC_STATUS_NONFINITE;
    };
    const int N = uni(d.N);
    switch (unit) {
        case DQMC_UNIT_BMUL:
            load_field_row(d, a, s);
            if (c <= 4) wg_bmul(d, c, b, X, d.W, s);
            else wg_wrap(d, b, c == 5, X, d.W, s);
            break;
        case DQMC_UNIT_RANK1:
            load_field_row(d, a, s);
            wg_rank1(d, b, X, Y, s);
            break;
        case DQMC_UNIT_PUSH:
            wg_push(d, 0, a, b, X, s);
            check_finite(slotU(d, 0, a), N * N);
            check_finite(slotT(d, 0, a), N * N);
            break;
        case DQMC_UNIT_FRESH:
            wg_fresh_green(d, 0, a, b, X, s);
            check_finite(X, N * N);
            break;
        case DQMC_UNIT_RATIO:
            if (tid() == 0) gptr(ext.out)[0] = unit_ratio(d, a, b, X, Y);
            break;
        case DQMC_UNIT_FLIP:
            if (tid() == 0) {
                AS_G int8_t* f = (AS_G int8_t*)d.field + (size_t)a * N + b;
                *f = (int8_t)(-*f);
            }
            break;
        case DQMC_UNIT_PUSH_EXT: {
            const bool top = ext.U_in != nullptr;
            if (top) unit_put_slot(d, 0, ext.U_in, ext.D_in, ext.V_in, s);
            wg_push(d, 0, 1, top ? 0 : -1, ext.A, s);
            __syncthreads();
            wg_copy(ext.U, slotU(d, 0, 1), N * N);
            for (int i = tid(); i < N; i += NT) gptr(ext.D)[i] = gptr(slotD(d, 0, 1))[i];
            wg_transpose_scaled(ext.V, nullptr, slotT(d, 0, 1), nullptr, N, nullptr, nullptr, s);
            __syncthreads();
            check_finite(ext.U, N * N);
            check_finite(ext.V, N * N);
            break;
        }
        case DQMC_UNIT_FRESH_EXT: {
            const bool left = ext.UL != nullptr, right = ext.UR != nullptr;
            if (left) unit_put_slot(d, 0, ext.UL, ext.DL, ext.VL, s);
            if (right) unit_put_slot(d, 1, ext.UR, ext.DR, ext.VR, s);
            wg_fresh_green(d, 0, left ? 0 : -1, right ? 1 : -1, ext.out, s);
            __syncthreads();
            check_finite(ext.out, N * N);
            break;
        }
    }
    __syncthreads();
    if (tid() == 0) {
        d.st->status = s.status;   // the building block's failure bits, for the host
        if (ext.status) ((AS_G int*)ext.status)[0] = s.status;
    }
}

// a descriptor passed by value into device memory (graph-capturable, no host copy)
__global__ void k_put_desc(ChainDesc desc, ChainDesc* dst) {
    if (threadIdx.x == 0) *dst = desc;
}

}  // namespace dqmc

extern "C" {
hipError_t dqmc_launch_measure_recorded(const ChainDesc* d_descs, int n_chains, hipStream_t stream) {
    hipLaunchKernelGGL(dqmc::k_measure_recorded, dim3(n_chains), dim3(NT), 0, stream, d_descs);
    return hipGetLastError();
}
hipError_t dqmc_launch_gemm(int variant, int M, int Nc, int K, const double* A, int lda, int tA, const double* B,
                            int ldb, int tB, double* C, int ldc, const double* rs, const double* ms, const double* cs,
                            double beta, hipStream_t stream) {
    hipLaunchKernelGGL(dqmc::k_gemm, dim3(1), dim3(NT), 0, stream, variant, M, Nc, K, A, lda, tA, B, ldb, tB, C, ldc,
                       rs, ms, cs, beta);
    return hipGetLastError();
}
hipError_t dqmc_launch_unit(const ChainDesc* d_descs, int unit, int a, int b, int c, double* X, double* Y,
                            hipStream_t stream) {
    dqmc::UnitPtrs ext{};
    hipLaunchKernelGGL(dqmc::k_unit, dim3(1), dim3(NT), 0, stream, d_descs, unit, a, b, c, X, Y, ext);
    return hipGetLastError();
}
// a building block on caller-provided device buffers: the descriptor goes to d_desc by value
// (unless desc is null), then the block runs on it (both on `stream`)
hipError_t dqmc_launch_unit_ext(const ChainDesc* desc, ChainDesc* d_desc, int unit, int a, int b, int c, double* X,
                                double* Y, const DqmcUnitPtrs* ptrs, hipStream_t stream) {
    dqmc::UnitPtrs ext{};
    if (ptrs) {
        ext.A = ptrs->A; ext.U_in = ptrs->U_in; ext.D_in = ptrs->D_in; ext.V_in = ptrs->V_in;
        ext.U = ptrs->U; ext.D = ptrs->D; ext.V = ptrs->V;
        ext.UL = ptrs->UL; ext.DL = ptrs->DL; ext.VL = ptrs->VL;
        ext.UR = ptrs->UR; ext.DR = ptrs->DR; ext.VR = ptrs->VR;
        ext.out = ptrs->out; ext.status = ptrs->status;
    }
    if (desc) {   // (null: the descriptor is already in place)
        hipLaunchKernelGGL(dqmc::k_put_desc, dim3(1), dim3(64), 0, stream, *desc, d_desc);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(dqmc::k_unit, dim3(1), dim3(NT), 0, stream, (const ChainDesc*)d_desc, unit, a, b, c, X, Y, ext);
    return hipGetLastError();
}
hipError_t dqmc_launch_init(const ChainDesc* d_descs, int n_chains, hipStream_t stream) {
    hipLaunchKernelGGL(dqmc::k_init, dim3(n_chains), dim3(NT), 0, stream, d_descs);
    return hipGetLastError();
}
hipError_t dqmc_launch_sweep(const ChainDesc* d_descs, int n_chains, int n_sweeps, int flags, hipStream_t stream) {
    int pair = 0;
    hipLaunchKernelGGL(dqmc::k_sweep, dim3(n_chains), dim3(NT), 0, stream, d_descs, n_sweeps, flags, n_chains, pair);
    return hipGetLastError();
}
// pair launch: 16 x ceil(n_chains / 8) workgroups; the chains' exchange flags and join words must
// be zero (the caller clears them on the stream)
hipError_t dqmc_launch_sweep_pair(const ChainDesc* d_descs, int n_chains, int n_sweeps, int flags, hipStream_t stream) {
    int pair = 1;
    hipLaunchKernelGGL(dqmc::k_sweep, dim3(16 * ((n_chains + 7) / 8)), dim3(NT), 0, stream, d_descs, n_sweeps, flags,
                       n_chains, pair);
    return hipGetLastError();
}
// workgroups of k_sweep that can be resident at once on a device (the pair launch's limit)
hipError_t dqmc_sweep_resident_limit(int device, int* limit) {
    int ncu = 0, per_cu = 0;
    hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dqmc::k_sweep, NT, 0);
    *limit = e == hipSuccess ? ncu * per_cu : 0;
    return e;
}
}
