// Host runtime of the DQMC engine: contexts, chain batches, the device arena
// of each batch, host-side chain setup (lattice tables, exp(-dtau K),
// checkerboard and HS constants, random fields) and the C-ABI of dqmc_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "runtime_internal.h"

extern "C" hipError_t dqmc_launch_init(const ChainDesc* d_descs, int n_chains, hipStream_t stream);
extern "C" hipError_t dqmc_launch_sweep(const ChainDesc* d_descs, int n_chains, int n_sweeps, int flags,
                                        hipStream_t stream);

extern "C" hipError_t dqmc_launch_gemm(int variant, int M, int Nc, int K, const double* A, int lda, int tA,
                                       const double* B, int ldb, int tB, double* C, int ldc, const double* rs,
                                       const double* ms, const double* cs, double beta, hipStream_t stream);

extern "C" hipError_t dqmc_launch_unit(const ChainDesc* d_descs, int unit, int a, int b, int c, double* X, double* Y,
                                       hipStream_t stream);

static_assert(DQMC_SWEEP_MEASURE == DQMC_FLAG_MEASURE && DQMC_SWEEP_RECORD == DQMC_FLAG_RECORD &&
                  DQMC_SWEEP_FROZEN == DQMC_FLAG_FROZEN && DQMC_SWEEP_PROFILE == DQMC_FLAG_PROFILE &&
                  DQMC_SWEEP_DYNAMIC == DQMC_FLAG_DYNAMIC,
              "the sweep flags of the C-ABI are the device flags");

namespace {



bool is_finite(double x) { return std::isfinite(x); }

// exp(-s K) of K = -t H + mu I on the periodic square lattice, with H the
// hopping matrix of src/lattice/square.cpp:164-177 (nearest neighbours in +x
// and +y, added symmetrically).  K is diagonal in plane waves, eps_k =
// -2t(cos kx + cos ky) + mu, so exp(-s K)(i,j) depends on the displacement only:
//   f(d) = 1/N sum_k exp(-s eps_k) cos(k . r_d).
void expk_closed_form(int side, double t, double mu, double s, std::vector<double>& out) {
    const int N = side * side;
    std::vector<double> f(N, 0.0);
    std::vector<double> ck(side);
    for (int a = 0; a < side; ++a) ck[a] = std::cos(2.0 * M_PI * a / side);
    for (int dy = 0; dy < side; ++dy)
        for (int dx = 0; dx < side; ++dx) {
            long double acc = 0.0L;
            for (int ky = 0; ky < side; ++ky)
                for (int kx = 0; kx < side; ++kx) {
                    const double eps = -2.0 * t * (ck[kx] + ck[ky]) + mu;
                    const double ph = 2.0 * M_PI * ((double)(kx * dx) / side + (double)(ky * dy) / side);
                    acc += (long double)std::exp(-s * eps) * (long double)std::cos(ph);
                }
            f[dx + side * dy] = (double)(acc / N);
        }
    out.assign((size_t)N * N, 0.0);
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const int dx = ((j % side) - (i % side) + side) % side;
            const int dy = ((j / side) - (i / side) + side) % side;
            out[(size_t)j * N + i] = f[dx + side * dy];
        }
}

// Kronecker factors of a dense exp(K) on the side x side lattice (site x + side y):
// F[(x,y),(x',y')] = X[x][x'] Y[y][y'] with X the block y = y' = 0 of F and Y the entries
// x = x' = 0 divided by F(0,0), both written 16 x 16 column-major and zero padded.  They
// are accepted when X (x) Y reproduces every entry of F within 1e-13 of its largest (true
// of the periodic square lattice, whose hopping is a sum over the two directions).
bool kron_factors(int side, const std::vector<double>& F, double* X, double* Y) {
    if (side > DQMC_KRON_MAX_SIDE) return false;
    const int N = side * side;
    const double f00 = F[0];
    if (!(std::fabs(f00) > 0.0)) return false;
    std::fill(X, X + 256, 0.0);
    std::fill(Y, Y + 256, 0.0);
    for (int a = 0; a < side; ++a)
        for (int b = 0; b < side; ++b) {
            X[b * 16 + a] = F[(size_t)b * N + a];                             // F[(a,0),(b,0)]
            Y[b * 16 + a] = F[(size_t)(side * b) * N + side * a] / f00;       // F[(0,a),(0,b)] / F(0,0)
        }
    double fmax = 0.0, err = 0.0;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const double f = F[(size_t)j * N + i];
            fmax = std::max(fmax, std::fabs(f));
            const double p = X[(j % side) * 16 + i % side] * Y[(j / side) * 16 + i / side];
            err = std::max(err, std::fabs(p - f));
        }
    return err <= 1e-13 * fmax;
}

// the chain's Kronecker factors (of expK and invExpK), or none
void kron_setup(HostChain& h) {
    h.kron.clear();
    if (h.p.checkerboard) return;
    std::vector<double> k(4 * 256);
    if (kron_factors(h.p.side, h.expK, k.data(), k.data() + 256) &&
        kron_factors(h.p.side, h.invExpK, k.data() + 512, k.data() + 768))
        h.kron.swap(k);
}

// ---------------------------------------------------------------------------
// Scratch of dqmc_k_greens_recompute_noalloc: one file-scope device array (each device holds its
// own copy until the process ends).  Host code never passes `scratch` itself to a kernel: its
// address on each device comes from hipGetSymbolAddress, taken once per device with that device
// current when a context is created (noalloc_scratch_of).  Layout in doubles, sized for
// DQMC_NOALLOC_MAX_N sites: the descriptor slot, the chain state, two stack slots (U, T, D) and
// the recompute's scratch W (3 N^2).  Nothing in it needs to start at zero: every call writes
// the descriptor, the stack slots and W before reading them, and the status word is stored,
// not accumulated.
// ---------------------------------------------------------------------------
constexpr size_t na_round(size_t doubles) { return (doubles + 31) / 32 * 32; }   // 256-byte granules
constexpr size_t NA_N2 = (size_t)DQMC_NOALLOC_MAX_N * DQMC_NOALLOC_MAX_N;
constexpr size_t NA_DESC = 0;
constexpr size_t NA_ST = NA_DESC + na_round((sizeof(ChainDesc) + 7) / 8);
constexpr size_t NA_U = NA_ST + na_round((sizeof(ChainState) + 7) / 8);
constexpr size_t NA_T = NA_U + 2 * NA_N2;
constexpr size_t NA_D = NA_T + 2 * NA_N2;
constexpr size_t NA_W = NA_D + na_round(2 * DQMC_NOALLOC_MAX_N);
constexpr size_t NA_TOTAL = NA_W + 3 * NA_N2;
}  // namespace

static __device__ __align__(256) double scratch[NA_TOTAL];

namespace {
std::mutex g_noalloc_mu;
std::map<int, double*> g_noalloc_ptr;   // device -> address of that device's copy of scratch

// the device address of `scratch` on `device` (the current device), resolved once per device
hipError_t noalloc_scratch_of(int device, double** out) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    auto it = g_noalloc_ptr.find(device);
    if (it != g_noalloc_ptr.end()) { *out = it->second; return hipSuccess; }
    double* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_noalloc_ptr[device] = ptr;
    *out = ptr;
    return hipSuccess;
}
}  // namespace

Layout chain_layout(int N, int L, int n_stack, int nspin, int kdelay, bool dense, bool record) {
    Layout l{};
    size_t o = 0;
    const size_t n2 = (size_t)N * N * sizeof(double);
    auto take = [&](size_t bytes) { size_t r = o; o += align256(bytes); return r; };
    l.G = take(nspin * n2);
    l.Gw = take(nspin * n2);
    l.Tmp = take(nspin * n2);
    l.stkU = take((size_t)nspin * n_stack * n2);
    l.stkT = take((size_t)nspin * n_stack * n2);
    l.stkD = take((size_t)nspin * n_stack * N * sizeof(double));
    l.W = take((size_t)nspin * 3 * n2);   // one stabilization scratch per spin
    l.Wq = take((size_t)nspin * 3 * 16 * (N + 16) * sizeof(double));
    l.expK = dense ? take(n2) : o;
    l.invExpK = dense ? take(n2) : o;
    l.kron = dense ? take(4 * 256 * sizeof(double)) : o;
    l.cosq = take(N * sizeof(double));
    l.field = take((size_t)L * N);
    l.st = take(sizeof(ChainState));
    l.acc = take(16 * sizeof(double));
    l.Gt0 = take(nspin * n2);
    l.G0t = take(nspin * n2);
    l.dacc = take(DQMC_DYN_ACC_SIZE(L, N) * sizeof(double));
    l.dscr = take(((size_t)L * N + N) * sizeof(double));
    l.rhoq = take(N * sizeof(double));
    l.recG = record ? take((size_t)L * nspin * n2) : o;
    l.recSign = record ? take(L * sizeof(double)) : o;
    l.recAcc = record ? take((size_t)L * N) : o;
    l.recGt0 = record ? take((size_t)L * nspin * n2) : o;
    l.recG0t = record ? take((size_t)L * nspin * n2) : o;
    l.xbuf = nspin == 2 ? take(DQMC_PAIR_XBUF_DOUBLES * sizeof(double)) : o;   // spin-pair exchange blocks
    l.total = o;
    return l;
}

int validate_params(const dqmc_chain_params& p, std::string& why) {
    const int N = p.side * p.side;
    if (p.side < 2 || N > DQMC_NMAX) { why = "side must satisfy 2 <= side and side*side <= 448"; return DQMC_ERR_ARG; }
    if (p.time_size < 1) { why = "time_size must be >= 1"; return DQMC_ERR_ARG; }
    if (p.stabilization_pace < 1) { why = "stabilization_pace must be >= 1"; return DQMC_ERR_ARG; }
    if (p.model != DQMC_MODEL_REPULSIVE && p.model != DQMC_MODEL_ATTRACTIVE) { why = "unknown model"; return DQMC_ERR_ARG; }
    if (p.checkerboard && (p.side % 2 != 0)) { why = "checkerboard break-up needs an even side length"; return DQMC_ERR_ARG; }
    if (p.delay < 1 || p.delay > DQMC_KMAX) { why = "delay must be in [1, 32] (a block of the submatrix walk takes at most 32 sites)"; return DQMC_ERR_ARG; }
    if (p.precision < DQMC_PREC_FP64 || p.precision > DQMC_PREC_FP32_PRODUCT) { why = "unsupported precision mode"; return DQMC_ERR_ARG; }
    if (!is_finite(p.beta) || !is_finite(p.hopping_t) || !is_finite(p.onsite_u) || !is_finite(p.chemical_potential)) {
        why = "non-finite model parameter"; return DQMC_ERR_NUMERIC;
    }
    if (p.beta <= 0.0) { why = "beta must be > 0"; return DQMC_ERR_ARG; }
    if (p.hopping_t < 0.0) { why = "hopping_t must be >= 0"; return DQMC_ERR_ARG; }
    if (p.model == DQMC_MODEL_REPULSIVE && p.onsite_u < 0.0) { why = "repulsive model needs onsite_u >= 0"; return DQMC_ERR_ARG; }
    const int nk = (p.side / 2 + 1) * (p.side / 2 + 2) / 2;
    if (p.momentum_index < -1 || p.momentum_index >= nk) { why = "momentum_index out of range"; return DQMC_ERR_ARG; }
    return DQMC_OK;
}






extern "C" {

int dqmc_context_create(int device, dqmc_ctx_t* out) {
    if (!out) return DQMC_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return DQMC_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return DQMC_ERR_HIP;
    dqmc_context* c = new dqmc_context();
    c->device = device;
    // set-up of the no-allocation entry point: this device's address of its static scratch
    if (noalloc_scratch_of(device, &c->noalloc_scratch) != hipSuccess) { delete c; return DQMC_ERR_HIP; }
    *out = c;
    return DQMC_OK;
}

int dqmc_context_destroy(dqmc_ctx_t ctx) {
    delete ctx;
    return DQMC_OK;
}

const char* dqmc_context_error(dqmc_ctx_t ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int dqmc_batch_footprint(int n_chains, const dqmc_chain_params* params, int record, size_t* bytes) {
    if (n_chains < 1 || !params || !bytes) return DQMC_ERR_ARG;
    size_t total = batch_head_bytes(n_chains);
    for (int c = 0; c < n_chains; ++c) {
        const dqmc_chain_params& p = params[c];
        std::string why;
        const int rc = validate_params(p, why);
        if (rc != DQMC_OK) return rc;
        const int N = p.side * p.side;
        const int n_stack = (p.time_size + p.stabilization_pace - 1) / p.stabilization_pace;
        const int nspin = p.model == DQMC_MODEL_ATTRACTIVE ? 1 : 2;
        total += chain_layout(N, p.time_size, n_stack, nspin, p.delay, !p.checkerboard, record != 0).total;
    }
    *bytes = total;
    return DQMC_OK;
}

}  // extern "C"

// Host-side setup of a batch: validation, per-chain constants and tables,
// arena layout.  No device work.
int batch_prepare(dqmc_context* ctx, int n_chains, const dqmc_chain_params* params, bool record, dqmc_batch* b) {
    b->ctx = ctx;
    b->n = n_chains;
    b->record = record;
    b->chains.resize(n_chains);
    b->layouts.resize(n_chains);
    b->base.resize(n_chains);
    b->descs.resize(n_chains);
    size_t off = batch_head_bytes(n_chains);  // the descriptor array and the pair flags head the arena
    for (int c = 0; c < n_chains; ++c) {
        HostChain& h = b->chains[c];
        h.p = params[c];
        std::string why;
        const int rc = validate_params(h.p, why);
        if (rc != DQMC_OK) {
            ctx->err = "chain " + std::to_string(c) + ": " + why;
            return rc;
        }
        const dqmc_chain_params& p = h.p;
        h.N = p.side * p.side;
        h.nspin = p.model == DQMC_MODEL_ATTRACTIVE ? 1 : 2;
        h.n_stack = (p.time_size + p.stabilization_pace - 1) / p.stabilization_pace;
        h.kdelay = p.delay;
        h.dtau = p.beta / p.time_size;
        const double U = std::fabs(p.onsite_u);
        h.alpha = std::acosh(std::exp(0.5 * h.dtau * U));
        h.fields.assign((size_t)p.time_size * h.N, 1);
        if (!p.checkerboard) {
            expk_closed_form(p.side, p.hopping_t, p.chemical_potential, h.dtau, h.expK);
            expk_closed_form(p.side, p.hopping_t, p.chemical_potential, -h.dtau, h.invExpK);
        }
        kron_setup(h);
        // Re exp(-i Q r_d) of the measured momentum (src/lattice/square.cpp:42-66, :148-161)
        const int side = p.side;
        std::vector<double> kx, ky;
        for (int i = (int)std::ceil(side / 2.0); i <= side; ++i)
            for (int j = (int)std::ceil(side / 2.0); j <= i; ++j) {
                kx.push_back((double)i / side * 2 * M_PI - M_PI);
                ky.push_back((double)j / side * 2 * M_PI - M_PI);
            }
        const int q = p.momentum_index < 0 ? (int)kx.size() - 1 : p.momentum_index;
        h.cosq.resize(h.N);
        for (int d = 0; d < h.N; ++d) h.cosq[d] = std::cos((-(d % side)) * kx[q] - (d / side) * ky[q]);
        // FourierFactor(rx, 1) - FourierFactor(ry, 1) of the stiffness (src/measure/measure_methods.cpp:371-389):
        // sites (rx, 0) and (ry, 0) against the k-star of index 1
        h.rhoq.resize(h.N);
        for (int d = 0; d < h.N; ++d) {
            const int rx = d % side, ry = d / side;
            h.rhoq[d] = std::cos((-rx) * kx[1] - 0 * ky[1]) - std::cos((-ry) * kx[1] - 0 * ky[1]);
        }
        h.kx = kx;
        h.ky = ky;
        b->layouts[c] = chain_layout(h.N, p.time_size, h.n_stack, h.nspin, h.kdelay, !p.checkerboard, b->record);
        b->base[c] = off;
        off += b->layouts[c].total;
    }
    b->arena_bytes = off;
    return DQMC_OK;
}

// Bind a prepared batch to its arena (b->arena, b->stream set): zero it, build
// the descriptors and upload constants and descriptors, all stream-ordered.
int batch_bind(dqmc_batch* b) {
    dqmc_context* ctx = b->ctx;
    b->d_descs = (ChainDesc*)b->arena;
    b->d_xflags = (unsigned*)(b->arena + align256(sizeof(ChainDesc) * (size_t)b->n));
    HIPCHK(ctx, hipMemsetAsync(b->arena, 0, b->arena_bytes, b->stream));
    for (int c = 0; c < b->n; ++c) {
        HostChain& h = b->chains[c];
        const Layout& l = b->layouts[c];
        char* base = b->arena + b->base[c];
        ChainDesc& d = b->descs[c];
        std::memset(&d, 0, sizeof(d));
        const dqmc_chain_params& p = h.p;
        d.launch_sweeps = 0;
        d.launch_flags = -1;
        d.N = h.N; d.L = p.time_size; d.side = p.side; d.pace = p.stabilization_pace; d.n_stack = h.n_stack;
        d.model = p.model; d.cb = p.checkerboard; d.kdelay = h.kdelay; d.nspin = h.nspin; d.precision = p.precision;
        d.hop_t = p.hopping_t;
        d.wrap_limit = b->wrap_limit;
        // exp(+spin alpha s) of src/model/repulsive_hubbard.cpp:173 for spin*s = +1 / -1
        d.ev[0] = std::exp(+1.0 * h.alpha);
        d.ev[1] = std::exp(-1.0 * h.alpha);
        // exp(-2 sigma alpha s) - 1 (ratio and gamma, src/model/repulsive_hubbard.cpp:120-147)
        const double em = std::exp(-2 * h.alpha * 1.0) - 1, ep = std::exp(-2 * h.alpha * -1.0) - 1;
        if (p.model == DQMC_MODEL_REPULSIVE) {
            d.dm1[0][0] = em; d.dm1[0][1] = ep;   // up: exp(-2 alpha s) - 1
            d.dm1[1][0] = std::exp(+2 * h.alpha * 1.0) - 1;
            d.dm1[1][1] = std::exp(+2 * h.alpha * -1.0) - 1;  // down: exp(+2 alpha s) - 1
            d.pref[0] = d.pref[1] = 1.0;
        } else {  // src/model/attractive_hubbard.cpp:120-124
            d.dm1[0][0] = d.dm1[1][0] = em;
            d.dm1[0][1] = d.dm1[1][1] = ep;
            d.pref[0] = std::exp(2 * h.alpha * 1.0);
            d.pref[1] = std::exp(2 * h.alpha * -1.0);
        }
        // plaquettes (src/checkerboard/square.cpp:34-49)
        {
            double ch = std::cosh(h.dtau * p.hopping_t), sh = std::sinh(h.dtau * p.hopping_t);
            double f = std::exp(0.5 * h.dtau * p.chemical_potential);
            d.P[0] = f * (ch * ch); d.P[1] = f * (ch * sh); d.P[2] = f * (sh * sh);
            ch = std::cosh(-h.dtau * p.hopping_t); sh = std::sinh(-h.dtau * p.hopping_t);
            f = std::exp(-0.5 * h.dtau * p.chemical_potential);
            d.Pi[0] = f * (ch * ch); d.Pi[1] = f * (ch * sh); d.Pi[2] = f * (sh * sh);
        }
        d.G = (double*)(base + l.G); d.Gw = (double*)(base + l.Gw); d.Tmp = (double*)(base + l.Tmp);
        d.stkU = (double*)(base + l.stkU); d.stkT = (double*)(base + l.stkT); d.stkD = (double*)(base + l.stkD);
        d.W = (double*)(base + l.W); d.Wq = (double*)(base + l.Wq);
        d.expK = p.checkerboard ? nullptr : (const double*)(base + l.expK);
        d.invExpK = p.checkerboard ? nullptr : (const double*)(base + l.invExpK);
        d.kron = h.kron.empty() ? nullptr : (const double*)(base + l.kron);
        d.cosq = (const double*)(base + l.cosq);
        d.field = (int8_t*)(base + l.field);
        d.st = (ChainState*)(base + l.st);
        d.acc = (double*)(base + l.acc);
        d.recG = b->record ? (double*)(base + l.recG) : nullptr;
        d.recSign = b->record ? (double*)(base + l.recSign) : nullptr;
        d.recAcc = b->record ? (uint8_t*)(base + l.recAcc) : nullptr;
        d.Gt0 = (double*)(base + l.Gt0);
        d.G0t = (double*)(base + l.G0t);
        d.dacc = (double*)(base + l.dacc);
        d.dscr = (double*)(base + l.dscr);
        d.rhoq = (const double*)(base + l.rhoq);
        d.recGt0 = b->record ? (double*)(base + l.recGt0) : nullptr;
        d.recG0t = b->record ? (double*)(base + l.recG0t) : nullptr;
        d.xbuf = h.nspin == 2 ? (double*)(base + l.xbuf) : nullptr;
        d.xflag = h.nspin == 2 ? b->d_xflags + (size_t)c * DQMC_PAIR_FLAG_WORDS : nullptr;
        if (!p.checkerboard) {
            HIPCHK(ctx, hipMemcpyAsync((void*)d.expK, h.expK.data(), h.expK.size() * 8, hipMemcpyHostToDevice, b->stream));
            HIPCHK(ctx, hipMemcpyAsync((void*)d.invExpK, h.invExpK.data(), h.invExpK.size() * 8, hipMemcpyHostToDevice, b->stream));
            if (d.kron) HIPCHK(ctx, hipMemcpyAsync((void*)d.kron, h.kron.data(), h.kron.size() * 8, hipMemcpyHostToDevice, b->stream));
        }
        HIPCHK(ctx, hipMemcpyAsync((void*)d.cosq, h.cosq.data(), h.cosq.size() * 8, hipMemcpyHostToDevice, b->stream));
        HIPCHK(ctx, hipMemcpyAsync((void*)d.rhoq, h.rhoq.data(), h.rhoq.size() * 8, hipMemcpyHostToDevice, b->stream));
    }
    HIPCHK(ctx, hipMemcpyAsync(b->d_descs, b->descs.data(), sizeof(ChainDesc) * b->n, hipMemcpyHostToDevice, b->stream));
    return DQMC_OK;
}

extern "C" {

int dqmc_batch_create(dqmc_ctx_t ctx, int n_chains, const dqmc_chain_params* params, int record, dqmc_batch_t* out) {
    if (!ctx || !params || !out || n_chains < 1) return DQMC_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    dqmc_batch* b = new dqmc_batch();
    int rc = batch_prepare(ctx, n_chains, params, record != 0, b);
    if (rc != DQMC_OK) { delete b; return rc; }
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { ctx->err = hipGetErrorString(e); delete b; return DQMC_ERR_HIP; }
    b->owns_stream = true;
    e = hipMalloc((void**)&b->arena, b->arena_bytes);
    if (e != hipSuccess) {
        ctx->err = std::string("arena of ") + std::to_string(b->arena_bytes) + " bytes: " + hipGetErrorString(e);
        hipStreamDestroy(b->stream);
        delete b;
        return DQMC_ERR_MEMORY;
    }
    b->owns_arena = true;
    rc = batch_bind(b);
    if (rc == DQMC_OK) {
        e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) { ctx->err = hipGetErrorString(e); rc = DQMC_ERR_HIP; }
    }
    if (rc != DQMC_OK) { dqmc_batch_destroy(b); return rc; }
    *out = b;
    return DQMC_OK;
}

int dqmc_batch_destroy(dqmc_batch_t b) {
    if (!b) return DQMC_ERR_ARG;
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->arena && b->owns_arena) hipFree(b->arena);
    if (b->stream && b->owns_stream) hipStreamDestroy(b->stream);
    delete b;
    return DQMC_OK;
}

int dqmc_batch_size(dqmc_batch_t b) { return b ? b->n : DQMC_ERR_ARG; }

int dqmc_batch_use_graph(dqmc_batch_t b, int enable) {
    if (!b) return DQMC_ERR_ARG;
    b->use_graph = enable != 0;
    return DQMC_OK;
}

static int check_chain(dqmc_batch_t b, int c) {
    if (!b) return DQMC_ERR_ARG;
    if (c < 0 || c >= b->n) { b->ctx->err = "chain index out of range"; return DQMC_ERR_ARG; }
    return DQMC_OK;
}

static int upload_state(dqmc_batch_t b, int c, uint64_t rng) {
    ChainState st{};
    st.rng = rng;
    st.sign = 1.0;
    HIPCHK(b->ctx, hipMemcpyAsync(b->descs[c].st, &st, sizeof(st), hipMemcpyHostToDevice, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_fields_random(dqmc_batch_t b) {
    if (!b) return DQMC_ERR_ARG;
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    for (int c = 0; c < b->n; ++c) {
        HostChain& h = b->chains[c];
        const uint64_t x = draw_random_fields(h.p.seed, h.fields);
        HIPCHK(b->ctx, hipMemcpyAsync(b->descs[c].field, h.fields.data(), h.fields.size(), hipMemcpyHostToDevice, b->stream));
        const int rc = upload_state(b, c, x);
        if (rc) return rc;
    }
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    b->initialized = false;
    return DQMC_OK;
}

int dqmc_set_fields(dqmc_batch_t b, int c, const int8_t* f) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    if (!f) return DQMC_ERR_ARG;
    HostChain& h = b->chains[c];
    for (size_t e = 0; e < h.fields.size(); ++e)
        if (f[e] != 1 && f[e] != -1) { b->ctx->err = "fields must be +1 or -1"; return DQMC_ERR_ARG; }
    std::memcpy(h.fields.data(), f, h.fields.size());
    HIPCHK(b->ctx, hipMemcpyAsync(b->descs[c].field, h.fields.data(), h.fields.size(), hipMemcpyHostToDevice, b->stream));
    // keep the engine state if the chain already has one; otherwise seed it
    ChainState st{};
    HIPCHK(b->ctx, hipMemcpyAsync(&st, b->descs[c].st, sizeof(st), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    if (st.rng == 0) {
        Minstd g; g.seed(h.p.seed);
        rc = upload_state(b, c, g.x);
        if (rc) return rc;
    }
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    b->initialized = false;
    return DQMC_OK;
}

int dqmc_set_rng_state(dqmc_batch_t b, int c, unsigned long long state) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    if (state == 0 || state >= 2147483647ULL) { b->ctx->err = "engine state must be in [1, 2^31-2]"; return DQMC_ERR_ARG; }
    HIPCHK(b->ctx, hipMemcpyAsync(&b->descs[c].st->rng, &state, sizeof(uint64_t), hipMemcpyHostToDevice, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_get_fields(dqmc_batch_t b, int c, int8_t* f) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    HostChain& h = b->chains[c];
    HIPCHK(b->ctx, hipMemcpyAsync(f, b->descs[c].field, h.fields.size(), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_set_expk(dqmc_batch_t b, int c, const double* expK, const double* invExpK) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    HostChain& h = b->chains[c];
    if (h.p.checkerboard) { b->ctx->err = "chain uses the checkerboard break-up"; return DQMC_ERR_STATE; }
    const size_t n2 = (size_t)h.N * h.N;
    for (size_t e = 0; e < n2; ++e)
        if (!is_finite(expK[e]) || !is_finite(invExpK[e])) { b->ctx->err = "non-finite exp(K)"; return DQMC_ERR_NUMERIC; }
    h.expK.assign(expK, expK + n2);
    h.invExpK.assign(invExpK, invExpK + n2);
    kron_setup(h);
    ChainDesc& d = b->descs[c];
    d.kron = h.kron.empty() ? nullptr : (const double*)(b->arena + b->base[c] + b->layouts[c].kron);
    HIPCHK(b->ctx, hipMemcpyAsync((void*)d.expK, expK, n2 * 8, hipMemcpyHostToDevice, b->stream));
    HIPCHK(b->ctx, hipMemcpyAsync((void*)d.invExpK, invExpK, n2 * 8, hipMemcpyHostToDevice, b->stream));
    if (d.kron) HIPCHK(b->ctx, hipMemcpyAsync((void*)d.kron, h.kron.data(), h.kron.size() * 8, hipMemcpyHostToDevice, b->stream));
    HIPCHK(b->ctx, hipMemcpyAsync(b->d_descs + c, &d, sizeof(ChainDesc), hipMemcpyHostToDevice, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    b->initialized = false;
    return DQMC_OK;
}

int dqmc_get_expk(dqmc_batch_t b, int c, double* expK, double* invExpK) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    HostChain& h = b->chains[c];
    if (h.p.checkerboard) { b->ctx->err = "chain uses the checkerboard break-up"; return DQMC_ERR_STATE; }
    std::memcpy(expK, h.expK.data(), h.expK.size() * 8);
    std::memcpy(invExpK, h.invExpK.data(), h.invExpK.size() * 8);
    return DQMC_OK;
}

int dqmc_init(dqmc_batch_t b) {
    if (!b) return DQMC_ERR_ARG;
    HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
    for (int c = 0; c < b->n; ++c) {
        HIPCHK(b->ctx, hipMemsetAsync(b->descs[c].acc, 0, 16 * sizeof(double), b->stream));
        HIPCHK(b->ctx, hipMemsetAsync(b->descs[c].dacc, 0, DQMC_DYN_ACC_SIZE(b->chains[c].p.time_size, b->chains[c].N) * 8,
                                      b->stream));
        b->chains[c].counts = 0;
        b->chains[c].sweeps = 0;
        b->chains[c].pos = 0;
    }
    HIPCHK(b->ctx, dqmc_launch_init(b->d_descs, b->n, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    b->initialized = true;
    return DQMC_OK;
}

// a pair launch pays: some chain is repulsive, and 16 x ceil(n / 8) sweep workgroups fit on the
// device at once (a chain whose workgroups do not meet runs alone, so this is about speed only)
static bool pair_qualifies(dqmc_batch* b) {
    bool any = false;
    for (const auto& h : b->chains) any = any || h.nspin == 2;
    if (!any) return false;
    dqmc_context* ctx = b->ctx;
    if (ctx->resident_limit < 0) {
        int lim = 0;
        if (dqmc_sweep_resident_limit(ctx->device, &lim) != hipSuccess) {
            lim = 0;
            (void)hipGetLastError();
        }
        ctx->resident_limit = lim;
    }
    return 16 * (((size_t)b->n + 7) / 8) <= (size_t)ctx->resident_limit;
}

int dqmc_batch_set_launch_mode(dqmc_batch_t b, int mode) {
    if (b && mode == DQMC_LAUNCH_QUAD) {
        b->ctx->err = "quad launch is not available: its helper workgroups did not finish on the device";
        return DQMC_ERR_ARG;
    }
    if (!b || mode < DQMC_LAUNCH_AUTO || mode > DQMC_LAUNCH_PAIR) return DQMC_ERR_ARG;
    HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
    if (mode == DQMC_LAUNCH_PAIR && !pair_qualifies(b)) {
        b->ctx->err = "pair launch needs a repulsive chain and 16 x ceil(chains / 8) resident workgroups";
        return DQMC_ERR_ARG;
    }
    b->launch_mode = mode;
    return DQMC_OK;
}

int dqmc_batch_last_launch_mode(dqmc_batch_t b, int* mode) {
    if (!b || !mode) return DQMC_ERR_ARG;
    *mode = b->last_mode;
    return DQMC_OK;
}

int dqmc_sweep(dqmc_batch_t b, int n_sweeps, int flags) {
    if (!b || n_sweeps < 0) return DQMC_ERR_ARG;
    if (!b->initialized) { b->ctx->err = "dqmc_sweep before dqmc_init"; return DQMC_ERR_STATE; }
    if ((flags & DQMC_SWEEP_RECORD) && !b->record) { b->ctx->err = "batch created without record buffers"; return DQMC_ERR_STATE; }
    if (n_sweeps == 0) return DQMC_OK;
    HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
    if (!b->ev0) {
        HIPCHK(b->ctx, hipEventCreate(&b->ev0));
        HIPCHK(b->ctx, hipEventCreate(&b->ev1));
    }
    // the time-displaced sweep runs one workgroup per chain; a pair launch is not replayed from a graph
    const bool pair = b->launch_mode != DQMC_LAUNCH_SINGLE && !(flags & DQMC_SWEEP_DYNAMIC) && pair_qualifies(b);
    if (pair) {
        HIPCHK(b->ctx, hipMemsetAsync(b->d_xflags, 0, sizeof(unsigned) * DQMC_PAIR_FLAG_WORDS * (size_t)b->n, b->stream));
        HIPCHK(b->ctx, hipEventRecord(b->ev0, b->stream));
        HIPCHK(b->ctx, dqmc_launch_sweep_pair(b->d_descs, b->n, n_sweeps, flags, b->stream));
        HIPCHK(b->ctx, hipEventRecord(b->ev1, b->stream));
    }
    if (!pair && b->use_graph && (b->graph_n != n_sweeps || b->graph_flags != flags)) {
        if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
        if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
        HIPCHK(b->ctx, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t le = dqmc_launch_sweep(b->d_descs, b->n, n_sweeps, flags, b->stream);
        HIPCHK(b->ctx, hipStreamEndCapture(b->stream, &b->graph));
        HIPCHK(b->ctx, le);
        HIPCHK(b->ctx, hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
        b->graph_n = n_sweeps;
        b->graph_flags = flags;
    }
    if (!pair) {
        HIPCHK(b->ctx, hipEventRecord(b->ev0, b->stream));
        if (b->use_graph) HIPCHK(b->ctx, hipGraphLaunch(b->graph_exec, b->stream));
        else HIPCHK(b->ctx, dqmc_launch_sweep(b->d_descs, b->n, n_sweeps, flags, b->stream));
        HIPCHK(b->ctx, hipEventRecord(b->ev1, b->stream));
    }
    b->last_mode = pair ? DQMC_LAUNCH_PAIR : DQMC_LAUNCH_SINGLE;
    b->timed = true;
    for (auto& h : b->chains) {
        h.sweeps += n_sweeps;
        // with DQMC_SWEEP_DYNAMIC the upward sweeps are the time-displaced ones, which take no
        // equal-time measurement
        const int n_up = h.pos == 0 ? (n_sweeps + 1) / 2 : n_sweeps / 2;
        const int n_eq = (flags & DQMC_SWEEP_DYNAMIC) ? n_sweeps - n_up : n_sweeps;
        if (flags & DQMC_SWEEP_MEASURE) h.counts += (double)n_eq * h.p.time_size;
        h.pos = (h.pos + n_sweeps) & 1;
    }
    return DQMC_OK;
}

int dqmc_set_wrap_error_limit(dqmc_batch_t b, double limit) {
    if (!b || !(limit > 0.0)) return DQMC_ERR_ARG;
    b->wrap_limit = limit;
    for (auto& d : b->descs) d.wrap_limit = limit;
    HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
    HIPCHK(b->ctx, hipMemcpyAsync(b->d_descs, b->descs.data(), sizeof(ChainDesc) * b->n, hipMemcpyHostToDevice, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_last_sweep_ms(dqmc_batch_t b, float* ms) {
    if (!b || !ms) return DQMC_ERR_ARG;
    if (!b->timed) { b->ctx->err = "no sweep launched yet"; return DQMC_ERR_STATE; }
    HIPCHK(b->ctx, hipEventSynchronize(b->ev1));
    HIPCHK(b->ctx, hipEventElapsedTime(ms, b->ev0, b->ev1));
    return DQMC_OK;
}

int dqmc_synchronize(dqmc_batch_t b) {
    if (!b) return DQMC_ERR_ARG;
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_get_green(dqmc_batch_t b, int c, int spin, double* out) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    if (spin != 1 && spin != -1) return DQMC_ERR_ARG;
    const HostChain& h = b->chains[c];
    const int sp = (spin == 1 || h.nspin == 1) ? 0 : 1;
    const size_t n2 = (size_t)h.N * h.N;
    HIPCHK(b->ctx, hipMemcpyAsync(out, b->descs[c].G + sp * n2, n2 * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_get_slice_green(dqmc_batch_t b, int c, int slot, int spin, double* out) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    const HostChain& h = b->chains[c];
    if (!b->record) { b->ctx->err = "batch created without record buffers"; return DQMC_ERR_STATE; }
    if (slot < 0 || slot >= h.p.time_size || (spin != 1 && spin != -1)) return DQMC_ERR_ARG;
    const int sp = (spin == 1 || h.nspin == 1) ? 0 : 1;
    const size_t n2 = (size_t)h.N * h.N;
    HIPCHK(b->ctx, hipMemcpyAsync(out, b->descs[c].recG + ((size_t)slot * h.nspin + sp) * n2, n2 * 8,
                                  hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_get_slice_signs(dqmc_batch_t b, int c, double* out) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    if (!b->record) { b->ctx->err = "batch created without record buffers"; return DQMC_ERR_STATE; }
    HIPCHK(b->ctx, hipMemcpyAsync(out, b->descs[c].recSign, b->chains[c].p.time_size * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_get_accepts(dqmc_batch_t b, int c, uint8_t* out) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    if (!b->record) { b->ctx->err = "batch created without record buffers"; return DQMC_ERR_STATE; }
    const HostChain& h = b->chains[c];
    HIPCHK(b->ctx, hipMemcpyAsync(out, b->descs[c].recAcc, (size_t)h.p.time_size * h.N, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_get_info(dqmc_batch_t b, int c, dqmc_chain_info* out) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    ChainState st{};
    HIPCHK(b->ctx, hipMemcpyAsync(&st, b->descs[c].st, sizeof(st), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    out->sign = st.sign;
    out->wrap_error = st.wrap_error;
    out->rng = st.rng;
    out->pos = st.pos;
    out->status = st.status;
    out->accepted = st.accepted;
    out->paired = st.paired;
    out->kron = b->descs[c].kron != nullptr;
    out->sweeps = b->chains[c].sweeps;
    for (int q = 0; q < DQMC_PH_COUNT; ++q) out->phase_seconds[q] = (double)st.phase[q] * 1e-8;  // 100 MHz clock
    return DQMC_OK;
}

int dqmc_get_measure(dqmc_batch_t b, int c, double* sums, double* counts) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    double acc[16];
    HIPCHK(b->ctx, hipMemcpyAsync(acc, b->descs[c].acc, sizeof(acc), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    for (int o = 0; o < DQMC_NUM_OBSERVABLES; ++o) {
        sums[o] = acc[o];
        counts[o] = b->chains[c].counts;
    }
    return DQMC_OK;
}

int dqmc_measure_recorded(dqmc_batch_t b) {
    if (!b) return DQMC_ERR_ARG;
    if (!b->record) { b->ctx->err = "batch created without record buffers"; return DQMC_ERR_STATE; }
    if (!b->initialized) { b->ctx->err = "measurement before dqmc_init"; return DQMC_ERR_STATE; }
    HIPCHK(b->ctx, dqmc_launch_measure_recorded(b->d_descs, b->n, b->stream));
    for (auto& h : b->chains) h.counts += (double)h.p.time_size;
    return DQMC_OK;
}

int dqmc_clear_measure(dqmc_batch_t b) {
    if (!b) return DQMC_ERR_ARG;
    for (int c = 0; c < b->n; ++c) {
        HIPCHK(b->ctx, hipMemsetAsync(b->descs[c].acc, 0, 16 * sizeof(double), b->stream));
        HIPCHK(b->ctx, hipMemsetAsync(b->descs[c].dacc, 0, DQMC_DYN_ACC_SIZE(b->chains[c].p.time_size, b->chains[c].N) * 8,
                                      b->stream));
        b->chains[c].counts = 0;
    }
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_get_dynamic_green(dqmc_batch_t b, int c, int slot, int spin, int which, double* out) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    const HostChain& h = b->chains[c];
    if (!b->record) { b->ctx->err = "batch created without record buffers"; return DQMC_ERR_STATE; }
    if (!out || slot < 0 || slot >= h.p.time_size || (spin != 1 && spin != -1) || (which != 0 && which != 1))
        return DQMC_ERR_ARG;
    const int sp = (spin == 1 || h.nspin == 1) ? 0 : 1;
    const size_t n2 = (size_t)h.N * h.N;
    const double* src = which == 0 ? b->descs[c].recGt0 : b->descs[c].recG0t;
    HIPCHK(b->ctx, hipMemcpyAsync(out, src + ((size_t)slot * h.nspin + sp) * n2, n2 * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_get_dynamic_measure(dqmc_batch_t b, int c, int n_k, const int* k_index, double* greens_k, double* dos,
                             double* dss, double* scalars) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    const HostChain& h = b->chains[c];
    const int N = h.N, L = h.p.time_size, nk = (int)h.kx.size();
    if (n_k < 0 || (n_k > 0 && (!k_index || !greens_k))) return DQMC_ERR_ARG;
    for (int q = 0; q < n_k; ++q)
        if (k_index[q] < 0 || k_index[q] >= nk) { b->ctx->err = "momentum index out of range"; return DQMC_ERR_ARG; }
    std::vector<double> acc(DQMC_DYN_ACC_SIZE(L, N));
    HIPCHK(b->ctx, hipMemcpyAsync(acc.data(), b->descs[c].dacc, acc.size() * 8, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    const double* gr = acc.data();
    const int side = h.p.side;
    // G(k,t) = 1/N sum_r FourierFactor(r, k) gr[t][r]  (src/measure/measure_methods.cpp:288-306),
    // element (k, t) at k + n_k t; D(t) = gr[t][0] / N (the trace, :313-325)
    std::vector<double> ff((size_t)n_k * N);
    for (int q = 0; q < n_k; ++q)
        for (int r = 0; r < N; ++r)
            ff[(size_t)q * N + r] = std::cos((-(r % side)) * h.kx[k_index[q]] - (r / side) * h.ky[k_index[q]]);
    for (int t = 0; t < L; ++t) {
        for (int q = 0; q < n_k; ++q) {
            double g = 0.0;
            for (int r = 0; r < N; ++r) g += ff[(size_t)q * N + r] * gr[(size_t)t * N + r];
            greens_k[(size_t)t * n_k + q] = g / N;
        }
        if (dos) dos[t] = gr[(size_t)t * N] / N;
        if (dss) dss[t] = acc[(size_t)L * N + t];
    }
    if (scalars) {
        scalars[0] = acc[(size_t)L * N + L];       // rho_s sum
        scalars[1] = acc[(size_t)L * N + L + 1];   // sign sum
        scalars[2] = acc[(size_t)L * N + L + 2];   // measurement count
    }
    return DQMC_OK;
}

int dqmc_num_kstars(dqmc_batch_t b, int c, double* kxy) {
    int rc = check_chain(b, c);
    if (rc) return rc;
    const HostChain& h = b->chains[c];
    if (kxy)
        for (size_t k = 0; k < h.kx.size(); ++k) { kxy[2 * k] = h.kx[k]; kxy[2 * k + 1] = h.ky[k]; }
    return (int)h.kx.size();
}

int dqmc_k_gemm(dqmc_ctx_t ctx, int variant, int M, int Nc, int K, int tA, int tB, const double* A, const double* B,
                double* C, double beta, const double* rs, const double* ms, const double* cs) {
    if (!ctx || !A || !B || !C || M < 1 || Nc < 1 || K < 1 || variant < 0 || variant > 6) return DQMC_ERR_ARG;
    // the scalings are staged in LDS
    if (variant >= 1 && ((ms && K > DQMC_NMAX) || (rs && M > DQMC_NMAX) || (cs && Nc > DQMC_NMAX))) return DQMC_ERR_ARG;
    if (variant == 3 && (size_t)M * K > DQMC_PANEL_DOUBLES) return DQMC_ERR_ARG;
    if (variant == 4 && (size_t)M * K + (size_t)K * Nc > DQMC_PANEL_DOUBLES) return DQMC_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t na = (size_t)M * K, nb = (size_t)K * Nc, nc = (size_t)M * Nc;
    double *dA = nullptr, *dB = nullptr, *dC = nullptr, *dv = nullptr;
    const size_t nv = (size_t)M + K + Nc;
    hipError_t e = hipMalloc((void**)&dA, na * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&dB, nb * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&dC, nc * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&dv, nv * 8);
    if (e == hipSuccess) e = hipMemcpy(dA, A, na * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB, B, nb * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dC, C, nc * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && rs) e = hipMemcpy(dv, rs, (size_t)M * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && ms) e = hipMemcpy(dv + M, ms, (size_t)K * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && cs) e = hipMemcpy(dv + M + K, cs, (size_t)Nc * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = dqmc_launch_gemm(variant, M, Nc, K, dA, tA ? K : M, tA, dB, tB ? Nc : K, tB, dC, M, rs ? dv : nullptr,
                             ms ? dv + M : nullptr, cs ? dv + M + K : nullptr, beta, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(C, dC, nc * 8, hipMemcpyDeviceToHost);
    if (dA) hipFree(dA);
    if (dB) hipFree(dB);
    if (dC) hipFree(dC);
    if (dv) hipFree(dv);
    if (e != hipSuccess) { ctx->err = std::string("dqmc_k_gemm: ") + hipGetErrorString(e); return DQMC_ERR_HIP; }
    return DQMC_OK;
}

// ---------------------------------------------------------------------------
// Single building blocks of the sweep, each run alone on a one-chain batch
// (host buffers in and out, synchronous).
// ---------------------------------------------------------------------------
namespace {

// a one-chain batch whose descriptor, stacks and scratch carry one building block
int unit_batch(dqmc_ctx_t ctx, const dqmc_chain_params& p, const int8_t* fields, const double* expK,
               const double* invExpK, dqmc_batch_t* out) {
    dqmc_batch_t b = nullptr;
    int rc = dqmc_batch_create(ctx, 1, &p, 0, &b);
    if (rc != DQMC_OK) return rc;
    if (fields) rc = dqmc_set_fields(b, 0, fields);
    if (rc == DQMC_OK && expK) {
        if (!invExpK) { ctx->err = "exp(K) needs its inverse"; rc = DQMC_ERR_ARG; }
        else rc = dqmc_set_expk(b, 0, expK, invExpK);
    }
    if (rc != DQMC_OK) { dqmc_batch_destroy(b); return rc; }
    *out = b;
    return DQMC_OK;
}

// stack-only building blocks take matrices of the engine's sizes: N = side^2, 2 <= side, N <= 448
int unit_stack_params(dqmc_ctx_t ctx, int N, dqmc_chain_params& p) {
    int side = 2;
    while (side * side < N) ++side;
    if (N < 4 || side * side != N || N > DQMC_NMAX) {
        ctx->err = "N must be side*side with 2 <= side and N <= 448";
        return DQMC_ERR_ARG;
    }
    std::memset(&p, 0, sizeof(p));
    p.side = side;
    p.time_size = 2;   // two stack levels: slot 0 and slot 1 of the spin-0 stack
    p.stabilization_pace = 1;
    p.beta = 1.0;
    p.hopping_t = 1.0;
    p.model = DQMC_MODEL_REPULSIVE;
    p.delay = 1;
    p.momentum_index = -1;
    p.precision = DQMC_PREC_FP64;
    return DQMC_OK;
}

hipError_t h2d(dqmc_batch_t b, void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream);
}
hipError_t d2h(dqmc_batch_t b, void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, b->stream);
}

void transpose(int N, const double* src, double* dst) {
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) dst[(size_t)j * N + i] = src[(size_t)i * N + j];
}

// stack slot k of spin 0 <- (U, D, V^T)
hipError_t put_slot(dqmc_batch_t b, int k, const double* U, const double* D, const double* V) {
    const ChainDesc& d = b->descs[0];
    const size_t n2 = (size_t)d.N * d.N;
    std::vector<double> T(n2);
    transpose(d.N, V, T.data());
    hipError_t e = h2d(b, d.stkU + k * n2, U, n2 * 8);
    if (e == hipSuccess) e = h2d(b, d.stkD + (size_t)k * d.N, D, (size_t)d.N * 8);
    if (e == hipSuccess) e = h2d(b, d.stkT + k * n2, T.data(), n2 * 8);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e;
}

// synchronizes, then reports a HIP error or the numerical failure bits the building block left
// in the chain state (DQMC_ERR_NUMERIC)
int unit_finish(dqmc_batch_t b, const char* who, hipError_t e) {
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    ChainState st{};
    if (e == hipSuccess) e = hipMemcpy(&st, b->descs[0].st, sizeof(st), hipMemcpyDeviceToHost);
    dqmc_ctx_t ctx = b->ctx;
    dqmc_batch_destroy(b);
    if (e != hipSuccess) { ctx->err = std::string(who) + ": " + hipGetErrorString(e); return DQMC_ERR_HIP; }
    if (st.status != 0) {
        ctx->err = std::string(who) + ": numerical failure (status bits " + std::to_string(st.status) + ")";
        return DQMC_ERR_NUMERIC;
    }
    return DQMC_OK;
}

}  // namespace

// One slice's B-matrix product on an N x N matrix M, in place
// (RepulsiveHubbard::mult_B_from_left .. mult_transB_from_left, src/model/repulsive_hubbard.cpp:158-245,
// and the wraps of DqmcWalker::wrap_from_0_to_beta / wrap_from_beta_to_0, src/dqmc_walker.cpp:208-245)
int dqmc_k_wrap(dqmc_ctx_t ctx, const dqmc_chain_params* p, const int8_t* fields, const double* expK,
                const double* invExpK, int t, int spin, int op, double* M) {
    if (!ctx || !p || !fields || !M || op < 0 || op > 6 || (spin != 1 && spin != -1)) return DQMC_ERR_ARG;
    if (t < 0 || t > p->time_size) { ctx->err = "time index out of range"; return DQMC_ERR_ARG; }
    // the reference's time index t acts with field row t - 1; t = 0 is the slice at beta
    const int row = t == 0 ? p->time_size - 1 : t - 1;
    dqmc_batch_t b = nullptr;
    int rc = unit_batch(ctx, *p, fields, expK, invExpK, &b);
    if (rc != DQMC_OK) return rc;
    const ChainDesc& d = b->descs[0];
    const size_t n2 = (size_t)d.N * d.N;
    const int sp = (spin == 1 || d.nspin == 1) ? 0 : 1;
    hipError_t e = h2d(b, d.G, M, n2 * 8);
    if (e == hipSuccess) e = dqmc_launch_unit(b->d_descs, DQMC_UNIT_BMUL, row, sp, op, d.G, nullptr, b->stream);
    if (e == hipSuccess) e = d2h(b, M, d.G, n2 * 8);
    return unit_finish(b, "dqmc_k_wrap", e);
}

// The single-flip update of the equal-time Green's functions for the field at (t, i)
// (RepulsiveHubbard::update_greens_function, src/model/repulsive_hubbard.cpp:127-155;
// AttractiveHubbard::update_greens_function, src/model/attractive_hubbard.cpp:128-156)
int dqmc_k_rank1(dqmc_ctx_t ctx, const dqmc_chain_params* p, const int8_t* fields, int t, int i, double* G_up,
                 double* G_dn) {
    if (!ctx || !p || !fields || !G_up) return DQMC_ERR_ARG;
    if (t < 0 || t >= p->time_size || i < 0 || i >= p->side * p->side) {
        ctx->err = "field index out of range";
        return DQMC_ERR_ARG;
    }
    dqmc_batch_t b = nullptr;
    int rc = unit_batch(ctx, *p, fields, nullptr, nullptr, &b);
    if (rc != DQMC_OK) return rc;
    const ChainDesc& d = b->descs[0];
    const size_t n2 = (size_t)d.N * d.N;
    // spin up in G, spin down in the product workspace (both at least N x N)
    hipError_t e = h2d(b, d.G, G_up, n2 * 8);
    if (e == hipSuccess && G_dn) e = h2d(b, d.Gw, G_dn, n2 * 8);
    if (e == hipSuccess)
        e = dqmc_launch_unit(b->d_descs, DQMC_UNIT_RANK1, t, i, 0, d.G, G_dn ? d.Gw : nullptr, b->stream);
    if (e == hipSuccess) e = d2h(b, G_up, d.G, n2 * 8);
    if (e == hipSuccess && G_dn) e = d2h(b, G_dn, d.Gw, n2 * 8);
    return unit_finish(b, "dqmc_k_rank1", e);
}

// SvdStack::push (src/svd_stack.cpp:34-61) in the engine's QR form: U diag(D) V^T =
// A (U_in diag(D_in) V_in^T), or of A alone when U_in is NULL (empty stack); U orthogonal,
// D > 0, V not orthogonal (it carries the triangular factors)
int dqmc_k_udv_push(dqmc_ctx_t ctx, int N, const double* A, const double* U_in, const double* D_in,
                    const double* V_in, double* U, double* D, double* V) {
    if (!ctx || !A || !U || !D || !V) return DQMC_ERR_ARG;
    if (U_in && (!D_in || !V_in)) { ctx->err = "a stack top needs U, D and V"; return DQMC_ERR_ARG; }
    dqmc_chain_params p;
    int rc = unit_stack_params(ctx, N, p);
    if (rc != DQMC_OK) return rc;
    dqmc_batch_t b = nullptr;
    rc = unit_batch(ctx, p, nullptr, nullptr, nullptr, &b);
    if (rc != DQMC_OK) return rc;
    const ChainDesc& d = b->descs[0];
    const size_t n2 = (size_t)N * N;
    const int slot = U_in ? 1 : 0;
    hipError_t e = U_in ? put_slot(b, 0, U_in, D_in, V_in) : hipSuccess;
    if (e == hipSuccess) e = h2d(b, d.G, A, n2 * 8);
    if (e == hipSuccess) e = dqmc_launch_unit(b->d_descs, DQMC_UNIT_PUSH, slot, slot - 1, 0, d.G, nullptr, b->stream);
    std::vector<double> T(n2);
    if (e == hipSuccess) e = d2h(b, U, d.stkU + slot * n2, n2 * 8);
    if (e == hipSuccess) e = d2h(b, D, d.stkD + (size_t)slot * N, (size_t)N * 8);
    if (e == hipSuccess) e = d2h(b, T.data(), d.stkT + slot * n2, n2 * 8);
    rc = unit_finish(b, "dqmc_k_udv_push", e);
    if (rc == DQMC_OK) transpose(N, T.data(), V);
    return rc;
}

// Equal-time Green's function from the two stack tops (NumericalStable::compute_equaltime_greens,
// include/utils/numerical_stable.hpp:186-238): G = (1 + B_L B_R^T)^-1 with B_L = U_L diag(D_L) V_L^T
// and B_R = U_R diag(D_R) V_R^T; a side given as NULL is an empty stack (B = 1)
int dqmc_k_greens_recompute(dqmc_ctx_t ctx, int N, const double* UL, const double* DL, const double* VL,
                            const double* UR, const double* DR, const double* VR, double* G) {
    if (!ctx || !G || (!UL && !UR)) return DQMC_ERR_ARG;
    if ((UL && (!DL || !VL)) || (UR && (!DR || !VR))) { ctx->err = "a stack top needs U, D and V"; return DQMC_ERR_ARG; }
    dqmc_chain_params p;
    int rc = unit_stack_params(ctx, N, p);
    if (rc != DQMC_OK) return rc;
    dqmc_batch_t b = nullptr;
    rc = unit_batch(ctx, p, nullptr, nullptr, nullptr, &b);
    if (rc != DQMC_OK) return rc;
    const ChainDesc& d = b->descs[0];
    const size_t n2 = (size_t)N * N;
    hipError_t e = UL ? put_slot(b, 0, UL, DL, VL) : hipSuccess;
    if (e == hipSuccess && UR) e = put_slot(b, 1, UR, DR, VR);
    if (e == hipSuccess)
        e = dqmc_launch_unit(b->d_descs, DQMC_UNIT_FRESH, UL ? 0 : -1, UR ? 1 : -1, 0, d.G, nullptr, b->stream);
    if (e == hipSuccess) e = d2h(b, G, d.G, n2 * 8);
    return unit_finish(b, "dqmc_k_greens_recompute", e);
}

// ---------------------------------------------------------------------------
// Building blocks on caller-provided device buffers.  A workspace is a one-chain batch
// whose arena carries the descriptor slot, two stack slots and the scratch; each call
// patches a host copy of its descriptor with the caller's pointers, and the device
// receives it by value (k_put_desc) on the caller's stream before the block runs.
// ---------------------------------------------------------------------------
namespace {
int ws_check(dqmc_workspace_t w) { return (w && w->b) ? DQMC_OK : DQMC_ERR_ARG; }
int ws_launch(dqmc_workspace_t w, const ChainDesc& d, int unit, int a, int bb, int c, double* X, double* Y,
              const DqmcUnitPtrs* ptrs, void* stream, const char* who) {
    dqmc_context* ctx = w->b->ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = dqmc_launch_unit_ext(&d, w->b->d_descs, unit, a, bb, c, X, Y, ptrs, (hipStream_t)stream);
    if (e != hipSuccess) { ctx->err = std::string(who) + ": " + hipGetErrorString(e); return DQMC_ERR_HIP; }
    return DQMC_OK;
}
}  // namespace

int dqmc_workspace_create(dqmc_ctx_t ctx, const dqmc_chain_params* p, dqmc_workspace_t* out) {
    if (!ctx || !p || !out) return DQMC_ERR_ARG;
    // two stack slots serve the push and the recompute: L' = L (2 for a one-slice chain, with
    // beta doubled so that dtau and every constant stay the same) at pace ceil(L'/2)
    dqmc_chain_params q = *p;
    if (q.time_size == 1) { q.time_size = 2; q.beta *= 2.0; }
    if (q.time_size >= 2) q.stabilization_pace = (q.time_size + 1) / 2;
    dqmc_batch_t b = nullptr;
    const int rc = dqmc_batch_create(ctx, 1, &q, 0, &b);
    if (rc != DQMC_OK) return rc;
    dqmc_workspace* w = new dqmc_workspace();
    w->b = b;
    w->time_size = p->time_size;
    *out = w;
    return DQMC_OK;
}

int dqmc_workspace_destroy(dqmc_workspace_t w) {
    if (!w) return DQMC_ERR_ARG;
    if (w->b) dqmc_batch_destroy(w->b);
    delete w;
    return DQMC_OK;
}

int dqmc_workspace_expk(dqmc_workspace_t w, double* expK, double* invExpK) {
    if (ws_check(w) || !expK || !invExpK) return DQMC_ERR_ARG;
    return dqmc_get_expk(w->b, 0, expK, invExpK);
}

int dqmc_kd_wrap(dqmc_workspace_t w, const int8_t* fields, const double* expK, const double* invExpK, int t,
                 int spin, int op, double* M, void* stream) {
    if (ws_check(w) || !fields || !M || op < 0 || op > 6 || (spin != 1 && spin != -1)) return DQMC_ERR_ARG;
    if (t < 0 || t > w->time_size) { w->b->ctx->err = "time index out of range"; return DQMC_ERR_ARG; }
    ChainDesc d = w->b->descs[0];
    if ((expK != nullptr) != (invExpK != nullptr)) { w->b->ctx->err = "exp(K) needs its inverse"; return DQMC_ERR_ARG; }
    if (expK && !d.cb) { d.expK = expK; d.invExpK = invExpK; d.kron = nullptr; }   // (caller's matrices: dense products)
    d.field = const_cast<int8_t*>(fields);
    const int row = t == 0 ? w->time_size - 1 : t - 1;   // time index t acts with field row t - 1
    const int sp = (spin == 1 || d.nspin == 1) ? 0 : 1;
    return ws_launch(w, d, DQMC_UNIT_BMUL, row, sp, op, M, nullptr, nullptr, stream, "dqmc_kd_wrap");
}

int dqmc_kd_update_ratio(dqmc_workspace_t w, const int8_t* fields, int t, int i, const double* G_up,
                         const double* G_dn, double* ratio, void* stream) {
    if (ws_check(w) || !fields || !G_up || !ratio) return DQMC_ERR_ARG;
    ChainDesc d = w->b->descs[0];
    if (d.nspin == 2 && !G_dn) return DQMC_ERR_ARG;
    if (t < 0 || t >= w->time_size || i < 0 || i >= d.N) { w->b->ctx->err = "field index out of range"; return DQMC_ERR_ARG; }
    d.field = const_cast<int8_t*>(fields);
    DqmcUnitPtrs ptr{};
    ptr.out = ratio;
    return ws_launch(w, d, DQMC_UNIT_RATIO, t, i, 0, const_cast<double*>(G_up), const_cast<double*>(G_dn), &ptr,
                     stream, "dqmc_kd_update_ratio");
}

int dqmc_kd_rank1(dqmc_workspace_t w, const int8_t* fields, int t, int i, double* G_up, double* G_dn, void* stream) {
    if (ws_check(w) || !fields || !G_up) return DQMC_ERR_ARG;
    ChainDesc d = w->b->descs[0];
    if (t < 0 || t >= w->time_size || i < 0 || i >= d.N) { w->b->ctx->err = "field index out of range"; return DQMC_ERR_ARG; }
    d.field = const_cast<int8_t*>(fields);
    return ws_launch(w, d, DQMC_UNIT_RANK1, t, i, 0, G_up, G_dn, nullptr, stream, "dqmc_kd_rank1");
}

int dqmc_kd_flip_field(dqmc_workspace_t w, int8_t* fields, int t, int i, void* stream) {
    if (ws_check(w) || !fields) return DQMC_ERR_ARG;
    ChainDesc d = w->b->descs[0];
    if (t < 0 || t >= w->time_size || i < 0 || i >= d.N) { w->b->ctx->err = "field index out of range"; return DQMC_ERR_ARG; }
    d.field = fields;
    return ws_launch(w, d, DQMC_UNIT_FLIP, t, i, 0, nullptr, nullptr, nullptr, stream, "dqmc_kd_flip_field");
}

int dqmc_kd_udv_push(dqmc_workspace_t w, const double* A, const double* U_in, const double* D_in, const double* V_in,
                     double* U, double* D, double* V, int* status, void* stream) {
    if (ws_check(w) || !A || !U || !D || !V) return DQMC_ERR_ARG;
    if (U_in && (!D_in || !V_in)) { w->b->ctx->err = "a stack top needs U, D and V"; return DQMC_ERR_ARG; }
    DqmcUnitPtrs ptr{};
    ptr.A = A; ptr.U_in = U_in; ptr.D_in = D_in; ptr.V_in = V_in;
    ptr.U = U; ptr.D = D; ptr.V = V; ptr.status = status;
    return ws_launch(w, w->b->descs[0], DQMC_UNIT_PUSH_EXT, 0, 0, 0, nullptr, nullptr, &ptr, stream, "dqmc_kd_udv_push");
}

int dqmc_kd_greens_recompute(dqmc_workspace_t w, const double* UL, const double* DL, const double* VL,
                             const double* UR, const double* DR, const double* VR, double* G, int* status,
                             void* stream) {
    if (ws_check(w) || !G || (!UL && !UR)) return DQMC_ERR_ARG;
    if ((UL && (!DL || !VL)) || (UR && (!DR || !VR))) { w->b->ctx->err = "a stack top needs U, D and V"; return DQMC_ERR_ARG; }
    DqmcUnitPtrs ptr{};
    ptr.UL = UL; ptr.DL = DL; ptr.VL = VL; ptr.UR = UR; ptr.DR = DR; ptr.VR = VR;
    ptr.out = G; ptr.status = status;
    return ws_launch(w, w->b->descs[0], DQMC_UNIT_FRESH_EXT, 0, 0, 0, nullptr, nullptr, &ptr, stream,
                     "dqmc_kd_greens_recompute");
}

// Per-site plugin calls on a batch chain's resident G and fields (synchronous on the batch
// stream); the chain's own descriptor is already on the device
namespace {
int chain_unit(dqmc_batch_t b, int c, int unit, int a, int bb, int op, double* X, double* Y, const DqmcUnitPtrs* ptrs,
               const char* who) {
    HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
    hipError_t e = dqmc_launch_unit_ext(nullptr, b->d_descs + c, unit, a, bb, op, X, Y, ptrs, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) { b->ctx->err = std::string(who) + ": " + hipGetErrorString(e); return DQMC_ERR_HIP; }
    return DQMC_OK;
}
int site_check(dqmc_batch_t b, int c, int t, int i) {
    const int rc = check_chain(b, c);
    if (rc) return rc;
    if (!b->initialized) { b->ctx->err = "chain not initialized (dqmc_init)"; return DQMC_ERR_STATE; }
    if (t < 0 || t >= b->chains[c].p.time_size || i < 0 || i >= b->chains[c].N) {
        b->ctx->err = "field index out of range";
        return DQMC_ERR_ARG;
    }
    return DQMC_OK;
}
}  // namespace

int dqmc_chain_update_ratio(dqmc_batch_t b, int c, int t, int i, double* ratio) {
    int rc = site_check(b, c, t, i);
    if (rc) return rc;
    if (!ratio) return DQMC_ERR_ARG;
    const ChainDesc& d = b->descs[c];
    DqmcUnitPtrs ptr{};
    ptr.out = d.acc + DQMC_ACC_SCRATCH;   // a spare accumulator slot carries the ratio to the host
    rc = chain_unit(b, c, DQMC_UNIT_RATIO, t, i, 0, d.G, d.nspin == 2 ? d.G + (size_t)d.N * d.N : d.G, &ptr,
                    "dqmc_chain_update_ratio");
    if (rc) return rc;
    HIPCHK(b->ctx, hipMemcpy(ratio, ptr.out, sizeof(double), hipMemcpyDeviceToHost));
    return DQMC_OK;
}

int dqmc_chain_update_green(dqmc_batch_t b, int c, int t, int i) {
    const int rc = site_check(b, c, t, i);
    if (rc) return rc;
    const ChainDesc& d = b->descs[c];
    return chain_unit(b, c, DQMC_UNIT_RANK1, t, i, 0, d.G, d.nspin == 2 ? d.G + (size_t)d.N * d.N : nullptr, nullptr,
                      "dqmc_chain_update_green");
}

int dqmc_chain_flip_field(dqmc_batch_t b, int c, int t, int i) {
    const int rc = site_check(b, c, t, i);
    if (rc) return rc;
    return chain_unit(b, c, DQMC_UNIT_FLIP, t, i, 0, nullptr, nullptr, nullptr, "dqmc_chain_flip_field");
}

int dqmc_chain_mult_green(dqmc_batch_t b, int c, int t, int spin, int op) {
    const int rc = check_chain(b, c);
    if (rc) return rc;
    if (!b->initialized) { b->ctx->err = "chain not initialized (dqmc_init)"; return DQMC_ERR_STATE; }
    const HostChain& h = b->chains[c];
    if (op < 0 || op > 6 || (spin != 1 && spin != -1)) return DQMC_ERR_ARG;
    if (t < 0 || t > h.p.time_size) { b->ctx->err = "time index out of range"; return DQMC_ERR_ARG; }
    const ChainDesc& d = b->descs[c];
    const int sp = (spin == 1 || d.nspin == 1) ? 0 : 1;
    const int row = t == 0 ? h.p.time_size - 1 : t - 1;
    return chain_unit(b, c, DQMC_UNIT_BMUL, row, sp, op, d.G + (size_t)sp * d.N * d.N, nullptr, nullptr,
                      "dqmc_chain_mult_green");
}

int dqmc_set_green(dqmc_batch_t b, int c, int spin, const double* G) {
    const int rc = check_chain(b, c);
    if (rc) return rc;
    if ((spin != 1 && spin != -1) || !G) return DQMC_ERR_ARG;
    const HostChain& h = b->chains[c];
    const size_t n2 = (size_t)h.N * h.N;
    for (size_t e = 0; e < n2; ++e)
        if (!std::isfinite(G[e])) { b->ctx->err = "non-finite G"; return DQMC_ERR_NUMERIC; }
    const int sp = (spin == 1 || h.nspin == 1) ? 0 : 1;
    HIPCHK(b->ctx, hipMemcpyAsync(b->descs[c].G + sp * n2, G, n2 * 8, hipMemcpyHostToDevice, b->stream));
    HIPCHK(b->ctx, hipStreamSynchronize(b->stream));
    return DQMC_OK;
}

int dqmc_chain_device_buffers(dqmc_batch_t b, int c, double** G_up, double** G_dn, int8_t** fields, void** stream) {
    const int rc = check_chain(b, c);
    if (rc) return rc;
    const ChainDesc& d = b->descs[c];
    const size_t n2 = (size_t)d.N * d.N;
    if (G_up) *G_up = d.G;
    if (G_dn) *G_dn = d.nspin == 2 ? d.G + n2 : d.G;
    if (fields) *fields = d.field;
    if (stream) *stream = (void*)b->stream;
    return DQMC_OK;
}

// dqmc_k_greens_recompute on caller device buffers with the scratch of the static array: no
// hipMalloc, no hipFree, no host-to-device copy (the descriptor reaches the device as a kernel
// argument).  Calls that overlap on one device share the array and race.
int dqmc_k_greens_recompute_noalloc(dqmc_ctx_t ctx, int N, const double* UL, const double* DL, const double* VL,
                                    const double* UR, const double* DR, const double* VR, double* G, int* status,
                                    void* stream) {
    if (!ctx || !G || (!UL && !UR)) return DQMC_ERR_ARG;
    if ((UL && (!DL || !VL)) || (UR && (!DR || !VR))) { ctx->err = "a stack top needs U, D and V"; return DQMC_ERR_ARG; }
    int side = 2;
    while (side * side < N) ++side;
    if (N < 4 || side * side != N) { ctx->err = "N must be side*side with 2 <= side"; return DQMC_ERR_ARG; }
    if (N > DQMC_NOALLOC_MAX_N) {
        ctx->err = "dqmc_k_greens_recompute_noalloc: N = " + std::to_string(N) + " exceeds its static scratch (N <= " +
                   std::to_string(DQMC_NOALLOC_MAX_N) + ")";
        return DQMC_ERR_ARG;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    double* base = ctx->noalloc_scratch;   // this device's address of `scratch`, never the symbol itself
    if (!base) { ctx->err = "no-allocation scratch not set up"; return DQMC_ERR_STATE; }
    ChainDesc d{};
    d.N = N; d.side = side; d.L = 2; d.pace = 1; d.n_stack = 2; d.model = DQMC_MODEL_REPULSIVE; d.cb = 0;
    d.kdelay = 1; d.nspin = 1; d.precision = DQMC_PREC_FP64; d.launch_flags = -1;
    d.wrap_limit = DQMC_DEFAULT_WRAP_LIMIT;
    d.st = (ChainState*)(base + NA_ST);
    d.stkU = base + NA_U;
    d.stkT = base + NA_T;
    d.stkD = base + NA_D;
    d.W = base + NA_W;
    DqmcUnitPtrs ptr{};
    ptr.UL = UL; ptr.DL = DL; ptr.VL = VL; ptr.UR = UR; ptr.DR = DR; ptr.VR = VR;
    ptr.out = G; ptr.status = status;
    HIPCHK(ctx, dqmc_launch_unit_ext(&d, (ChainDesc*)(base + NA_DESC), DQMC_UNIT_FRESH_EXT, 0, 0, 0, nullptr, nullptr,
                                     &ptr, (hipStream_t)stream));
    return DQMC_OK;
}

// Delayed-update depth autotuner: times `sweeps` single-direction sweeps of a
// batch of `n_chains` chains of the given shape for each candidate depth (after
// one untimed sweep) with HIP events, and returns the fastest.
int dqmc_autotune_delay(dqmc_ctx_t ctx, const dqmc_chain_params* shape, int n_chains, const int* candidates,
                        int n_candidates, int sweeps, int* best, float* ms_out) {
    if (!ctx || !shape || !candidates || !best || n_chains < 1 || n_candidates < 1 || sweeps < 1) return DQMC_ERR_ARG;
    float best_ms = -1.0f;
    for (int c = 0; c < n_candidates; ++c) {
        std::vector<dqmc_chain_params> ps(n_chains, *shape);
        for (int k = 0; k < n_chains; ++k) { ps[k].delay = candidates[c]; ps[k].seed = shape->seed + k; }
        dqmc_batch_t b = nullptr;
        int rc = dqmc_batch_create(ctx, n_chains, ps.data(), 0, &b);
        if (rc != DQMC_OK) return rc;
        float ms = 0.0f;
        rc = dqmc_fields_random(b);
        if (rc == DQMC_OK) rc = dqmc_init(b);
        if (rc == DQMC_OK) rc = dqmc_sweep(b, 1, 0);
        if (rc == DQMC_OK) rc = dqmc_sweep(b, sweeps, 0);
        if (rc == DQMC_OK) rc = dqmc_last_sweep_ms(b, &ms);
        dqmc_batch_destroy(b);
        if (rc != DQMC_OK) return rc;
        if (ms_out) ms_out[c] = ms;
        if (best_ms < 0.0f || ms < best_ms) { best_ms = ms; *best = candidates[c]; }
    }
    return DQMC_OK;
}

int dqmc_autotune_launch(dqmc_ctx_t ctx, const dqmc_chain_params* shape, int n_chains, int sweeps, int* best,
                         float* ms_out) {
    if (!ctx || !shape || !best || n_chains < 1 || sweeps < 1) return DQMC_ERR_ARG;
    const int modes[2] = {DQMC_LAUNCH_SINGLE, DQMC_LAUNCH_PAIR};
    float best_ms = -1.0f;
    for (int m = 0; m < 2; ++m) {
        std::vector<dqmc_chain_params> ps(n_chains, *shape);
        for (int k = 0; k < n_chains; ++k) ps[k].seed = shape->seed + k;
        dqmc_batch_t b = nullptr;
        int rc = dqmc_batch_create(ctx, n_chains, ps.data(), 0, &b);
        if (rc != DQMC_OK) return rc;
        float ms = -1.0f;
        rc = dqmc_batch_set_launch_mode(b, modes[m]);
        if (rc == DQMC_ERR_ARG && m == 1) {   // the batch does not qualify for pairs
            dqmc_batch_destroy(b);
            break;
        }
        if (rc == DQMC_OK) rc = dqmc_fields_random(b);
        if (rc == DQMC_OK) rc = dqmc_init(b);
        if (rc == DQMC_OK) rc = dqmc_sweep(b, 1, 0);
        if (rc == DQMC_OK) rc = dqmc_sweep(b, sweeps, 0);
        if (rc == DQMC_OK) rc = dqmc_last_sweep_ms(b, &ms);
        dqmc_batch_destroy(b);
        if (rc != DQMC_OK) return rc;
        if (ms_out) ms_out[m] = ms;
        if (best_ms < 0.0f || ms < best_ms) { best_ms = ms; *best = modes[m]; }
    }
    return DQMC_OK;
}

}  // extern "C"
