// Continuous-batching job server of the DQMC engine.
//
// Jobs (one Markov chain each: lattice, beta, U, mu, sweeps, seed) arrive at
// any time through dqmc_server_submit.  A scheduler thread admits them by
// device-memory footprint into a pool allocated once (so the serving loop
// never calls hipMalloc/hipFree), keeps every admitted chain resident (G,
// UDV stacks, fields, random state) and composes sweep launches over the
// resident chains.  Chains are grouped into lanes by lattice size; each lane
// has its own stream, so launches of small and large chains overlap on the
// device.  Within a launch every chain receives as many sweeps as fit a
// target launch time (from an online cost model), cut at its warm-up/bin
// boundaries, so chains of very different cost progress together and a
// newly admitted chain joins at the next launch of its lane.
//
// The per-chain computation is the one of the batch API (same kernels, same
// state between launches), so a job's result does not depend on what else
// is resident.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "runtime_internal.h"

namespace {


// first-fit allocator over one device pool (offsets, 256-byte granules)
struct Pool {
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    void init(size_t n) { size = n; free_.clear(); free_[0] = n; }
    bool alloc(size_t n, size_t& off) {
        n = align256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                const size_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = align256(n);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
        }
    }
};

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct Job {
    long long id = 0;
    dqmc_job spec{};
    size_t bytes = 0, off = 0;
    int lane = 0;
    std::unique_ptr<dqmc_batch> b;
    ChainState st0{};
    bool needs_init = true;
    std::atomic<bool> cancel{false};   // set by dqmc_server_cancel (under s->mu), read by the scheduler
    long long done = 0, total = 0;
    int in_bin = 0, launched = 0, launched_flags = 0;
    int stage = -1;                  // slot of the lane's pinned staging the running launch copies into
    bool stage_acc = false, stage_st = false;   // which of acc / state that copy holds
    double cost = 1.0;  // FLOP per chain-sweep (model)
    double t_submit = 0, t_start = 0;
    std::vector<std::vector<double>> bins;
    // DQMC_JOB_DYNAMIC: equal-time measured sweeps of the running bin (the backward ones) and
    // the dynamic bins [|sign|, rho_s, G(k,t) (n_k x L), D(t), DSS(t)]
    int eq_in_bin = 0;
    std::vector<std::vector<double>> dyn_bins;
};

struct DynResult {
    int n_k = 0, time_size = 0;
    std::vector<double> mean, err;
};

// per job of a launch: its 16 accumulators and its chain state, copied to pinned host memory
// behind the launch on the lane stream (the scheduler never waits on a copy)
constexpr int STAGE_DOUBLES = 16 + (int)((sizeof(ChainState) + 7) / 8);

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    double* h_stage = nullptr;   // pinned, max_active x STAGE_DOUBLES
    ChainDesc* d_descs = nullptr;
    ChainDesc* d_init = nullptr;
    unsigned* d_xflags = nullptr;   // spin-pair flags of a launch's chains (max_active x DQMC_PAIR_FLAG_WORDS)
    std::vector<ChainDesc> h_descs, h_init;
    std::vector<Job*> in_flight;
    std::deque<Job*> queue;
    bool busy = false;
    double launch_flop = 0;
    std::chrono::steady_clock::time_point t_launch;
};

struct dqmc_server {
    dqmc_context* ctx = nullptr;
    size_t budget = 0, used = 0, peak = 0;
    int max_active = 0;
    char* pool_base = nullptr;
    Pool pool;
    std::vector<Lane> lanes;
    std::mutex mu;
    std::condition_variable cv;
    std::vector<Job*> active;
    std::vector<dqmc_job_result> results;
    std::map<long long, DynResult> dyn_results;
    std::thread worker;
    std::atomic<bool> stop{false};
    long long next_id = 1, submitted = 0, completed = 0, chain_sweeps = 0, launches = 0;
    std::chrono::steady_clock::time_point t0;
    double rate = 2.0e10;        // FLOP/s one chain achieves (updated online)
    double target_s = 0.02;      // target launch duration
    int status = DQMC_OK;
    std::string err;
};

namespace {

int lane_of(const dqmc_server* s, int N) {
    const int cls = N <= 64 ? 0 : (N <= 144 ? 1 : (N <= 256 ? 2 : 3));
    return std::min(cls, (int)s->lanes.size() - 1);
}

double chain_sweep_flop(const dqmc_chain_params& p) {
    const double N = (double)p.side * p.side, L = p.time_size;
    const double ns = std::ceil(L / p.stabilization_pace);
    const double spins = p.model == DQMC_MODEL_ATTRACTIVE ? 1.0 : 2.0;
    const double prop = p.checkerboard ? 96.0 * L * N * N : 12.0 * L * N * N * N;
    const double upd = 4.0 * (0.3 * N * L) * N * N;
    return spins / 2.0 * (prop + upd + 29.3 * ns * N * N * N);
}

void finish_job(dqmc_server* s, Job* j, int status, hipStream_t stream, const ChainState* staged = nullptr) {
    dqmc_job_result r{};
    r.id = j->id;
    r.status = status;
    r.t_submit = j->t_submit;
    r.t_start = j->t_start;
    r.t_finish = now_s(s->t0);
    r.sweeps = j->done;
    r.n_bins = (int)j->bins.size();
    r.wrap_error = 0.0;
    if (j->b) {
        ChainState st{};
        bool ok = false;
        if (staged) { st = *staged; ok = true; }
        else ok = hipMemcpyAsync(&st, j->b->descs[0].st, sizeof(st), hipMemcpyDeviceToHost, stream) == hipSuccess &&
                  hipStreamSynchronize(stream) == hipSuccess;
        if (ok) {
            r.wrap_error = st.wrap_error;
            r.chain_status = st.status;
            // a chain that failed numerically stopped sweeping: its observables are not results
            if ((st.status & DQMC_STATUS_FATAL) && status == DQMC_OK) r.status = DQMC_ERR_NUMERIC;
        }
    }
    const size_t nb = j->bins.size();
    for (int o = 0; o < DQMC_NUM_OBSERVABLES; ++o) {
        double m = 0.0, m2 = 0.0;
        for (auto& v : j->bins) { m += v[o]; m2 += v[o] * v[o]; }
        if (nb > 0) { m /= nb; m2 /= nb; }
        r.mean[o] = m;
        r.error[o] = nb > 1 ? std::sqrt(std::max(0.0, m2 - m * m)) / std::sqrt((double)nb - 1) : 0.0;
    }
    DynResult dr;
    const bool dyn = (j->spec.options & DQMC_JOB_DYNAMIC) && !j->dyn_bins.empty();
    if (dyn) {
        const size_t nd = j->dyn_bins.size(), nv = j->dyn_bins[0].size();
        dr.time_size = j->spec.chain.time_size;
        dr.n_k = (int)((nv - 2 - 2 * (size_t)dr.time_size) / dr.time_size);
        dr.mean.assign(nv, 0.0);
        dr.err.assign(nv, 0.0);
        for (size_t q = 0; q < nv; ++q) {   // Observable::analyse (include/measure/observable.h:169-193)
            double m = 0.0, m2 = 0.0;
            for (auto& v : j->dyn_bins) { m += v[q]; m2 += v[q] * v[q]; }
            m /= nd; m2 /= nd;
            dr.mean[q] = m;
            dr.err[q] = nd > 1 ? std::sqrt(std::max(0.0, m2 - m * m)) / std::sqrt((double)nd - 1) : 0.0;
        }
    }
    if (j->b) {
        s->pool.release(j->off, j->bytes);
        s->used -= j->bytes;
        j->b.reset();
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (dyn) s->dyn_results[j->id] = std::move(dr);
    s->results.push_back(r);
    s->completed += 1;
    s->active.erase(std::remove(s->active.begin(), s->active.end(), j), s->active.end());
    delete j;
    s->cv.notify_all();
}

// place a queued job into the pool and stage its state on the lane stream
bool admit(dqmc_server* s, Lane& ln, Job* j) {
    size_t off = 0;
    if ((int)s->active.size() >= s->max_active) return false;
    if (s->used + j->bytes > s->budget || !s->pool.alloc(j->bytes, off)) return false;
    j->off = off;
    j->b.reset(new dqmc_batch());
    dqmc_batch* b = j->b.get();
    if (batch_prepare(s->ctx, 1, &j->spec.chain, false, b) != DQMC_OK) {
        s->pool.release(off, j->bytes);
        return false;
    }
    b->arena = s->pool_base + off;
    b->stream = ln.stream;
    if (batch_bind(b) != DQMC_OK) { s->pool.release(off, j->bytes); j->b.reset(); return false; }
    HostChain& h = b->chains[0];
    const uint64_t x = draw_random_fields(h.p.seed, h.fields);
    j->st0 = ChainState{};
    j->st0.rng = x;
    j->st0.sign = 1.0;
    hipMemcpyAsync(b->descs[0].field, h.fields.data(), h.fields.size(), hipMemcpyHostToDevice, ln.stream);
    hipMemcpyAsync(b->descs[0].st, &j->st0, sizeof(ChainState), hipMemcpyHostToDevice, ln.stream);
    s->used += j->bytes;
    s->peak = std::max(s->peak, s->used);
    j->needs_init = true;
    j->t_start = now_s(s->t0);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->active.push_back(j);
    }
    return true;
}

void complete_lane(dqmc_server* s, Lane& ln) {
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - ln.t_launch).count();
    if (dt > 0 && ln.launch_flop > 0) {
        // the slowest chain bounds a launch; track the per-chain rate it implies
        double worst = 0.0;
        for (Job* j : ln.in_flight) worst = std::max(worst, j->launched * j->cost);
        const double r = worst / dt;
        s->rate = 0.8 * s->rate + 0.2 * r;
    }
    for (Job* j : ln.in_flight) {
        const bool dyn = j->spec.options & DQMC_JOB_DYNAMIC;
        if (j->launched_flags & DQMC_FLAG_MEASURE) {
            // with the dynamic option the sweeps that start at slice 0 (even sweep count so far)
            // are time-displaced and take no equal-time measurement
            const int n = j->launched;
            const int n_up = (j->done % 2 == 0) ? (n + 1) / 2 : n / 2;
            j->eq_in_bin += dyn ? n - n_up : n;
        }
        j->done += j->launched;
        s->chain_sweeps += j->launched;
        if (j->launched_flags & DQMC_FLAG_MEASURE) {
            j->in_bin += j->launched;
            if (j->in_bin >= j->spec.bin_size) {
                // the launch copied the bin's accumulators to the lane's pinned staging and
                // cleared them on the device, stream-ordered behind the sweeps
                const double* acc = ln.h_stage + (size_t)j->stage * STAGE_DOUBLES;
                if (dyn) {
                    const HostChain& h = j->b->chains[0];
                    const int N = h.N, L = h.p.time_size, side = h.p.side, nk = (int)h.kx.size();
                    std::vector<double> da(DQMC_DYN_ACC_SIZE(L, N));
                    hipMemcpyAsync(da.data(), j->b->descs[0].dacc, da.size() * 8, hipMemcpyDeviceToHost, ln.stream);
                    hipStreamSynchronize(ln.stream);
                    hipMemsetAsync(j->b->descs[0].dacc, 0, da.size() * 8, ln.stream);
                    // MeasureHandler::normalize_stats, dynamic branch (src/measure/measure_handler.cpp:191-206):
                    // sums / (count <sign>) = sums / sign sum; G(k,t) and D(t) from the displacement sums
                    const double ssum = da[(size_t)L * N + L + 1], cnt = da[(size_t)L * N + L + 2];
                    std::vector<double> v;
                    v.reserve(2 + (size_t)nk * L + 2 * (size_t)L);
                    v.push_back(std::fabs(ssum / cnt));
                    v.push_back(da[(size_t)L * N + L] / ssum);
                    std::vector<double> ff((size_t)nk * N);
                    for (int k = 0; k < nk; ++k)
                        for (int r = 0; r < N; ++r)
                            ff[(size_t)k * N + r] = std::cos((-(r % side)) * h.kx[k] - (r / side) * h.ky[k]);
                    for (int t = 0; t < L; ++t)
                        for (int k = 0; k < nk; ++k) {
                            double g = 0.0;
                            for (int r = 0; r < N; ++r) g += ff[(size_t)k * N + r] * da[(size_t)t * N + r];
                            v.push_back(g / N / ssum);
                        }
                    for (int t = 0; t < L; ++t) v.push_back(da[(size_t)t * N] / N / ssum);
                    for (int t = 0; t < L; ++t) v.push_back(da[(size_t)L * N + t] / ssum);
                    j->dyn_bins.push_back(std::move(v));
                }
                const double cnt = (double)(dyn ? j->eq_in_bin : j->in_bin) * j->spec.chain.time_size;
                std::vector<double> v(DQMC_NUM_OBSERVABLES);
                const double sg = acc[0] / cnt;  // MeasureHandler::normalize_stats
                for (int o = 1; o < DQMC_NUM_OBSERVABLES; ++o) v[o] = acc[o] / (cnt * sg);
                v[0] = std::fabs(sg);
                j->bins.push_back(v);
                j->in_bin = 0;
                j->eq_in_bin = 0;
            }
        }
        j->launched = 0;
    }
    std::vector<Job*> fin;
    for (Job* j : ln.in_flight)
        if (j->done >= j->total || j->cancel) fin.push_back(j);
    ln.in_flight.clear();
    for (Job* j : fin) {
        ChainState st{};
        const bool staged = j->stage_st;
        if (staged) std::memcpy(&st, ln.h_stage + (size_t)j->stage * STAGE_DOUBLES + 16, sizeof(ChainState));
        finish_job(s, j, j->cancel && j->done < j->total ? 1 : DQMC_OK, ln.stream, staged ? &st : nullptr);
    }
}

bool launch_lane(dqmc_server* s, Lane& ln, int lane_index) {
    std::vector<Job*> jobs;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (Job* j : s->active)
            if (j->lane == lane_index) jobs.push_back(j);
    }
    std::vector<Job*> cancelled;
    for (Job* j : jobs)
        if (j->cancel) cancelled.push_back(j);
    for (Job* j : cancelled) {
        jobs.erase(std::remove(jobs.begin(), jobs.end(), j), jobs.end());
        finish_job(s, j, 1, ln.stream);
    }
    if (jobs.empty()) return false;
    // initialize newly admitted chains first (stream-ordered before the sweeps)
    ln.h_init.clear();
    for (Job* j : jobs)
        if (j->needs_init) { ln.h_init.push_back(j->b->descs[0]); j->needs_init = false; }
    if (!ln.h_init.empty()) {
        hipMemcpyAsync(ln.d_init, ln.h_init.data(), sizeof(ChainDesc) * ln.h_init.size(), hipMemcpyHostToDevice, ln.stream);
        if (dqmc_launch_init(ln.d_init, (int)ln.h_init.size(), ln.stream) != hipSuccess) s->status = DQMC_ERR_HIP;
    }
    ln.h_descs.clear();
    ln.launch_flop = 0;
    for (Job* j : jobs) {
        const dqmc_job& sp = j->spec;
        int phase_left, flags;
        if (j->done < sp.warmup_sweeps) { phase_left = (int)(sp.warmup_sweeps - j->done); flags = 0; }
        else {
            phase_left = std::min(sp.bin_size - j->in_bin, (int)(j->total - j->done));
            flags = DQMC_FLAG_MEASURE | ((sp.options & DQMC_JOB_DYNAMIC) ? DQMC_FLAG_DYNAMIC : 0);
        }
        const double est = j->cost / s->rate;  // seconds per sweep
        int n = (int)std::floor(s->target_s / est);
        n = std::max(1, std::min(n, phase_left));
        j->launched = n;
        j->launched_flags = flags;
        ChainDesc d = j->b->descs[0];
        d.launch_sweeps = n;
        d.launch_flags = flags;
        // this launch's flags for the chain (its exchange blocks stay in the job's own memory)
        if (d.xflag) d.xflag = ln.d_xflags + ln.h_descs.size() * DQMC_PAIR_FLAG_WORDS;
        ln.h_descs.push_back(d);
        ln.launch_flop += n * j->cost;
        ln.in_flight.push_back(j);
    }
    hipMemcpyAsync(ln.d_descs, ln.h_descs.data(), sizeof(ChainDesc) * ln.h_descs.size(), hipMemcpyHostToDevice, ln.stream);
    // spin pairs: each repulsive chain's spins on two workgroups when both get a CU (a chain whose
    // partner is kept off the device by the other lanes runs alone, with the same result)
    const int nl = (int)ln.h_descs.size();
    hipError_t le = hipMemsetAsync(ln.d_xflags, 0, sizeof(unsigned) * DQMC_PAIR_FLAG_WORDS * nl, ln.stream);
    if (le == hipSuccess) le = dqmc_launch_sweep_pair(ln.d_descs, nl, 1, 0, ln.stream);
    if (le != hipSuccess) s->status = DQMC_ERR_HIP;
    // the readbacks this launch ends in, behind it on the stream: a bin's accumulators (then
    // cleared) and the state of a job that finishes
    for (size_t q = 0; q < ln.in_flight.size(); ++q) {
        Job* j = ln.in_flight[q];
        j->stage = (int)q;
        double* slot = ln.h_stage + q * STAGE_DOUBLES;
        j->stage_acc = (j->launched_flags & DQMC_FLAG_MEASURE) && j->in_bin + j->launched >= j->spec.bin_size;
        j->stage_st = j->done + j->launched >= j->total || j->cancel;
        if (j->stage_acc) {
            hipMemcpyAsync(slot, j->b->descs[0].acc, 16 * sizeof(double), hipMemcpyDeviceToHost, ln.stream);
            hipMemsetAsync(j->b->descs[0].acc, 0, 16 * sizeof(double), ln.stream);
        }
        if (j->stage_st) hipMemcpyAsync(slot + 16, j->b->descs[0].st, sizeof(ChainState), hipMemcpyDeviceToHost, ln.stream);
    }
    hipEventRecord(ln.ev, ln.stream);
    ln.t_launch = std::chrono::steady_clock::now();
    ln.busy = true;
    s->launches += 1;
    return true;
}

void scheduler(dqmc_server* s) {
    hipSetDevice(s->ctx->device);
    while (true) {
        bool progressed = false;
        for (size_t li = 0; li < s->lanes.size(); ++li) {
            Lane& ln = s->lanes[li];
            if (ln.busy) {
                if (hipEventQuery(ln.ev) != hipSuccess) continue;
                ln.busy = false;
                complete_lane(s, ln);
                progressed = true;
            }
            // admit queued jobs of this lane while they fit (FIFO within the lane)
            while (true) {
                Job* j = nullptr;
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    if (ln.queue.empty()) break;
                    j = ln.queue.front();
                }
                if (j->cancel) {
                    { std::lock_guard<std::mutex> lk(s->mu); ln.queue.pop_front(); }
                    finish_job(s, j, 1, ln.stream);
                    continue;
                }
                if (!admit(s, ln, j)) break;
                std::lock_guard<std::mutex> lk(s->mu);
                ln.queue.pop_front();
            }
            if (launch_lane(s, ln, (int)li)) progressed = true;
        }
        if (!progressed) {
            std::unique_lock<std::mutex> lk(s->mu);
            bool idle = s->active.empty();
            for (auto& ln : s->lanes) idle = idle && ln.queue.empty() && !ln.busy;
            if (s->stop && idle) break;
            s->cv.wait_for(lk, std::chrono::microseconds(idle ? 2000 : 100));
        }
    }
}

}  // namespace

extern "C" {

int dqmc_server_create(dqmc_ctx_t ctx, size_t hbm_budget, int max_active, int n_lanes, dqmc_server_t* out) {
    if (!ctx || !out || hbm_budget < 1024 || max_active < 1 || n_lanes < 1 || n_lanes > 4) return DQMC_ERR_ARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    dqmc_server* s = new dqmc_server();
    s->ctx = ctx;
    s->budget = align256(hbm_budget);
    s->max_active = max_active;
    hipError_t e = hipMalloc((void**)&s->pool_base, s->budget);
    if (e != hipSuccess) { ctx->err = std::string("server pool: ") + hipGetErrorString(e); delete s; return DQMC_ERR_MEMORY; }
    s->pool.init(s->budget);
    s->lanes.resize(n_lanes);
    for (auto& ln : s->lanes) {
        e = hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ln.ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc((void**)&ln.d_descs, sizeof(ChainDesc) * max_active);
        if (e == hipSuccess) e = hipMalloc((void**)&ln.d_init, sizeof(ChainDesc) * max_active);
        if (e == hipSuccess) e = hipMalloc((void**)&ln.d_xflags, sizeof(unsigned) * DQMC_PAIR_FLAG_WORDS * max_active);
        if (e == hipSuccess) e = hipHostMalloc((void**)&ln.h_stage, sizeof(double) * STAGE_DOUBLES * max_active, 0);
        if (e != hipSuccess) { ctx->err = hipGetErrorString(e); dqmc_server_destroy(s); return DQMC_ERR_HIP; }
    }
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread(scheduler, s);
    *out = s;
    return DQMC_OK;
}

int dqmc_server_submit(dqmc_server_t s, const dqmc_job* job, long long* id) {
    if (!s || !job) return DQMC_ERR_ARG;
    std::string why;
    int rc = validate_params(job->chain, why);
    if (rc == DQMC_OK && (job->warmup_sweeps < 0 || job->measure_sweeps < 0 || job->bin_size < 1 ||
                          (job->measure_sweeps % job->bin_size) != 0)) {
        why = "sweep counts: warmup >= 0, measure >= 0 and a multiple of bin_size >= 1";
        rc = DQMC_ERR_ARG;
    }
    if (rc == DQMC_OK && (job->options & ~DQMC_JOB_DYNAMIC)) { why = "unknown job option"; rc = DQMC_ERR_ARG; }
    if (rc == DQMC_OK && (job->options & DQMC_JOB_DYNAMIC) && (job->bin_size % 2 != 0 || job->warmup_sweeps % 2 != 0)) {
        why = "dynamic jobs measure whole forth-and-back iterations: even warmup and bin_size";
        rc = DQMC_ERR_ARG;
    }
    size_t bytes = 0;
    if (rc == DQMC_OK) rc = dqmc_batch_footprint(1, &job->chain, 0, &bytes);
    if (rc == DQMC_OK && bytes > s->budget) { why = "job footprint exceeds the server's device budget"; rc = DQMC_ERR_MEMORY; }
    if (rc != DQMC_OK) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = why;
        return rc;
    }
    Job* j = new Job();
    j->spec = *job;
    j->bytes = bytes;
    j->total = (long long)job->warmup_sweeps + job->measure_sweeps;
    j->cost = chain_sweep_flop(job->chain);
    j->lane = lane_of(s, job->chain.side * job->chain.side);
    std::lock_guard<std::mutex> lk(s->mu);
    j->id = s->next_id++;
    j->t_submit = now_s(s->t0);
    s->lanes[j->lane].queue.push_back(j);
    s->submitted += 1;
    if (id) *id = j->id;
    s->cv.notify_all();
    return DQMC_OK;
}

int dqmc_server_cancel(dqmc_server_t s, long long id) {
    if (!s) return DQMC_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    for (Job* j : s->active)
        if (j->id == id) { j->cancel = true; return DQMC_OK; }
    for (auto& ln : s->lanes)
        for (Job* j : ln.queue)
            if (j->id == id) { j->cancel = true; return DQMC_OK; }
    return DQMC_ERR_ARG;
}

int dqmc_server_poll(dqmc_server_t s, dqmc_job_result* out, int max_results, int* n_out) {
    if (!s || !n_out || (max_results > 0 && !out)) return DQMC_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    const int n = std::min((int)s->results.size(), std::max(0, max_results));
    for (int k = 0; k < n; ++k) out[k] = s->results[k];
    s->results.erase(s->results.begin(), s->results.begin() + n);
    *n_out = n;
    return s->status;
}

int dqmc_server_wait_idle(dqmc_server_t s, double timeout_s) {
    if (!s) return DQMC_ERR_ARG;
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool idle = s->active.empty();
        for (auto& ln : s->lanes) idle = idle && ln.queue.empty();
        if (idle) return DQMC_OK;
        if (std::chrono::steady_clock::now() >= until) return DQMC_ERR_STATE;
        s->cv.wait_for(lk, std::chrono::milliseconds(5));
    }
}

int dqmc_server_get_stats(dqmc_server_t s, dqmc_server_stats* out) {
    if (!s || !out) return DQMC_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->active = (long long)s->active.size();
    long long q = 0;
    for (auto& ln : s->lanes) q += (long long)ln.queue.size();
    out->queued = q;
    out->chain_sweeps = s->chain_sweeps;
    out->launches = s->launches;
    out->bytes_in_use = s->used;
    out->bytes_peak = s->peak;
    out->budget = s->budget;
    out->elapsed_s = now_s(s->t0);
    return DQMC_OK;
}

int dqmc_server_get_dynamic(dqmc_server_t s, long long id, int* n_k, int* time_size, double* mean, double* err) {
    if (!s) return DQMC_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->dyn_results.find(id);
    if (it == s->dyn_results.end()) { s->err = "no dynamic results for this job"; return DQMC_ERR_ARG; }
    if (n_k) *n_k = it->second.n_k;
    if (time_size) *time_size = it->second.time_size;
    if (mean) std::memcpy(mean, it->second.mean.data(), it->second.mean.size() * sizeof(double));
    if (err) std::memcpy(err, it->second.err.data(), it->second.err.size() * sizeof(double));
    return DQMC_OK;
}

const char* dqmc_server_error(dqmc_server_t s) { return s ? s->err.c_str() : "null server"; }

int dqmc_server_destroy(dqmc_server_t s) {
    if (!s) return DQMC_ERR_ARG;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Job* j : s->active) j->cancel = true;
            for (auto& ln : s->lanes)
                for (Job* j : ln.queue) j->cancel = true;
            s->stop = true;
        }
        s->cv.notify_all();
        s->worker.join();
    }
    for (auto& ln : s->lanes) {
        if (ln.stream) hipStreamSynchronize(ln.stream);
        if (ln.d_descs) hipFree(ln.d_descs);
        if (ln.d_init) hipFree(ln.d_init);
        if (ln.d_xflags) hipFree(ln.d_xflags);
        if (ln.h_stage) hipHostFree(ln.h_stage);
        if (ln.ev) hipEventDestroy(ln.ev);
        if (ln.stream) hipStreamDestroy(ln.stream);
    }
    if (s->pool_base) hipFree(s->pool_base);
    delete s;
    return DQMC_OK;
}

}  // extern "C"
