// Internal structures of the host runtime shared by the batch API and the
// serving scheduler (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/dqmc_amd.h"
#include "chain_desc.h"

extern "C" hipError_t dqmc_launch_init(const ChainDesc* d_descs, int n_chains, hipStream_t stream);
extern "C" hipError_t dqmc_launch_measure_recorded(const ChainDesc* d_descs, int n_chains, hipStream_t stream);
extern "C" hipError_t dqmc_launch_sweep(const ChainDesc* d_descs, int n_chains, int n_sweeps, int flags,
                                        hipStream_t stream);

struct dqmc_context {
    int device = 0;
    std::string err;
};

struct HostChain {
    dqmc_chain_params p;
    int N = 0, nspin = 2, n_stack = 0, kdelay = 1;
    double dtau = 0, alpha = 0;
    std::vector<int8_t> fields;
    std::vector<double> expK, invExpK, cosq, rhoq;
    std::vector<double> kx, ky;   // k-stars of the lattice (src/lattice/square.cpp:42-66)
    long long sweeps = 0;
    double counts = 0;  // measured slices (same count for every observable)
    int pos = 0;        // 0: the next sweep goes up (from slice 0), 1: down
    size_t off_desc_bytes = 0;
};

struct Layout {
    size_t G, Gw, Tmp, stkU, stkT, stkD, W, Wq, expK, invExpK, cosq, field, st, acc, Gt0, G0t, dacc, dscr, rhoq, recG,
        recSign, recAcc, recGt0, recG0t, total;
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
    // sweep launch replayed from a captured HIP graph (per n_sweeps/flags)
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_n = -1, graph_flags = -1;
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
int batch_prepare(dqmc_context* ctx, int n_chains, const dqmc_chain_params* params, bool record, dqmc_batch* b);
int batch_bind(dqmc_batch* b);

Layout chain_layout(int N, int L, int n_stack, int nspin, int kdelay, bool dense, bool record);
int validate_params(const dqmc_chain_params& p, std::string& why);
