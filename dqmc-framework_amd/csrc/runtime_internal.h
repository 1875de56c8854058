// Internal structures of the host runtime shared by the batch API and the
// serving scheduler (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dqmc_amd.h"
#include "chain_desc.h"

extern "C" hipError_t dqmc_launch_init(const ChainDesc* d_descs, int n_chains, hipStream_t stream);
extern "C" hipError_t dqmc_launch_measure_recorded(const ChainDesc* d_descs, int n_chains, hipStream_t stream);
extern "C" hipError_t dqmc_launch_sweep(const ChainDesc* d_descs, int n_chains, int n_sweeps, int flags,
                                        hipStream_t stream);
extern "C" hipError_t dqmc_launch_sweep_pair(const ChainDesc* d_descs, int n_chains, int n_sweeps, int flags,
                                             hipStream_t stream);
extern "C" hipError_t dqmc_sweep_resident_limit(int device, int* limit);
extern "C" hipError_t dqmc_launch_unit_ext(const ChainDesc* desc, ChainDesc* d_desc, int unit, int a, int b, int c,
                                           double* X, double* Y, const DqmcUnitPtrs* ptrs, hipStream_t stream);

struct dqmc_context {
    int device = 0;
    std::string err;
    double* noalloc_scratch = nullptr;   // the device's address of the no-allocation entry point's scratch
    int resident_limit = -1;             // sweep workgroups resident at once (0: no cooperative launch); -1 unknown
};

struct HostChain {
    dqmc_chain_params p;
    int N = 0, nspin = 2, n_stack = 0, kdelay = 1;
    double dtau = 0, alpha = 0;
    std::vector<int8_t> fields;
    std::vector<double> expK, invExpK, cosq, rhoq;
    std::vector<double> kron;     // Kronecker factors of expK and invExpK (4 x 16 x 16), empty: dense products
    std::vector<double> kx, ky;   // k-stars of the lattice (src/lattice/square.cpp:42-66)
    long long sweeps = 0;
    double counts = 0;  // measured slices (same count for every observable)
    int pos = 0;        // 0: the next sweep goes up (from slice 0), 1: down
    size_t off_desc_bytes = 0;
};

struct Layout {
    size_t G, Gw, Tmp, stkU, stkT, stkD, W, Wq, expK, invExpK, kron, cosq, field, st, acc, Gt0, G0t, dacc, dscr, rhoq,
        recG, recSign, recAcc, recGt0, recG0t, xbuf, total;
};

struct dqmc_batch {
    dqmc_context* ctx = nullptr;
    hipStream_t stream = nullptr;
    int n = 0;
    bool record = false, initialized = false;
    std::vector<HostChain> chains;
    std::vector<Layout> layouts;
    std::vector<size_t> base;       // arena offset of each chain
    std::vector<ChainDesc> descs;   // host copy
    char* arena = nullptr;
    ChainDesc* d_descs = nullptr;
    size_t arena_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // bracket the last sweep launch on the batch stream
    bool timed = false;
    bool owns_arena = false, owns_stream = false;
    double wrap_limit = DQMC_DEFAULT_WRAP_LIMIT;   // copied into every chain's descriptor
    // sweep launch replayed from a captured HIP graph (per n_sweeps/flags)
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_n = -1, graph_flags = -1;
    // chain -> workgroup mapping of the sweep launches (DQMC_LAUNCH_*), the one the last launch used,
    // and the spin-pair exchange flags (n x DQMC_PAIR_FLAG_WORDS, after the descriptor array)
    int launch_mode = DQMC_LAUNCH_AUTO, last_mode = 0;
    unsigned* d_xflags = nullptr;
};

// scratch of the building blocks on caller-provided device buffers: a one-chain batch (host
// constants and tables, descriptor slot, stack slots and GEMM/factorization scratch)
struct dqmc_workspace {
    dqmc_batch* b = nullptr;
    int time_size = 0;   // the caller's L (the batch may carry L = 2 for a one-slice chain)
};

#define HIPCHK(ctx, call)                                                           \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);         \
            return DQMC_ERR_HIP;                                                    \
        }                                                                           \
    } while (0)

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
// head of a batch arena: the descriptor array, then the spin-pair flags of every chain
inline size_t batch_head_bytes(int n_chains) {
    return align256(sizeof(ChainDesc) * (size_t)n_chains) +
           align256(sizeof(unsigned) * DQMC_PAIR_FLAG_WORDS * (size_t)n_chains);
}

// minstd_rand0 (std::default_random_engine of libstdc++) and the canonical double of
// std::bernoulli_distribution (C++ standard [rand.util.canonical], two engine draws per
// decision), as the reference's Utils::Random draws them (src/random.cpp:6-10); the device
// twin is in dqmc_kernels.hip
struct Minstd {
    uint64_t x = 1;
    void seed(long long s) {
        x = ((uint64_t)(int64_t)s) % 2147483647ULL;
        if (x == 0) x = 1;
    }
    uint64_t next() { x = (16807ULL * x) % 2147483647ULL; return x; }
};
inline double canonical(Minstd& g) {
#pragma clang fp contract(off)
    const double R = 2147483646.0;
    double sum = 0.0, tmp = 1.0;
    for (int k = 0; k < 2; ++k) {
        sum = sum + (double)(g.next() - 1) * tmp;
        tmp = tmp * R;
    }
    double r = sum / tmp;
    if (r >= 1.0) r = std::nextafter(1.0, 0.0);
    return r;
}
// the random HS field of a chain (src/model/repulsive_hubbard.cpp:86-99: t-outer, i-inner, one
// bernoulli(0.5) each); returns the engine state after the draws
inline uint64_t draw_random_fields(long long seed, std::vector<int8_t>& fields) {
    Minstd g;
    g.seed(seed);
    for (size_t e = 0; e < fields.size(); ++e) fields[e] = canonical(g) < 0.5 ? 1 : -1;
    return g.x;
}
int batch_prepare(dqmc_context* ctx, int n_chains, const dqmc_chain_params* params, bool record, dqmc_batch* b);
int batch_bind(dqmc_batch* b);

Layout chain_layout(int N, int L, int n_stack, int nspin, int kdelay, bool dense, bool record);
int validate_params(const dqmc_chain_params& p, std::string& why);
