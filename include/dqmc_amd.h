/*
 * dqmc_amd.h — C-ABI of the MI355X DQMC engine (plain pointers, sizes and
 * int status codes; no C++ or torch types).
 *
 * The boundary replaces the reference's equal-time sweep plugin calls of
 * JefferyWangSH/dqmc-framework:
 *   chain setup      Model::ModelBase::set_model_params / initial
 *                      (include/model/model_base.h:111-130),
 *                    Lattice::Square (include/lattice/square.h, src/lattice/square.cpp:16-196),
 *                    CheckerBoard::Square::initial (src/checkerboard/square.cpp:28-50),
 *                    DqmcWalker::set_physical_params / set_stabilization_pace
 *                      (src/dqmc_walker.cpp:25-38)
 *   random fields    Model::*Hubbard::set_bosonic_fields_to_random (src/model/repulsive_hubbard.cpp:86-99)
 *                    with Utils::Random::set_seed (src/random.cpp:8-10)
 *   init             DqmcInitializer::initial_dqmc → DqmcWalker::initial_svd_stacks,
 *                    initial_greens_functions, initial_config_sign (src/dqmc_walker.cpp:109-163)
 *   sweeps           DqmcWalker::sweep_from_0_to_beta / sweep_from_beta_to_0
 *                      (src/dqmc_walker.cpp:246-412), driven as in Dqmc::thermalize / measure
 *                      (src/dqmc.cpp:48-151)
 *   measurement      MeasureHandler::equaltime_measure (src/measure/measure_handler.cpp:136-150)
 *                    with the methods of src/measure/measure_methods.cpp:22-255
 *   accessors        DqmcWalker::GreenttUp/Dn, GreenttUp/Dn(t), ConfigSign(t), WrapError
 *                      (include/dqmc_walker.h:136-164)
 *
 * Conventions: matrices are column-major N x N doubles (the element layout
 * of Eigen::MatrixXd); fields are int8 (+1/-1) with element (t, i) at t*N+i;
 * spin is +1 (up) or -1 (down).  Calls on one batch are issued on that
 * batch's HIP stream; getters synchronize it.  One context per thread.
 */
#ifndef DQMC_AMD_H
#define DQMC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define DQMC_OK 0
#define DQMC_ERR_ARG -1      /* invalid argument (sizes, indices, unsupported combination) */
#define DQMC_ERR_HIP -2      /* a HIP runtime call failed (message in dqmc_context_error) */
#define DQMC_ERR_MEMORY -3   /* device allocation failed */
#define DQMC_ERR_STATE -4    /* call out of order (e.g. sweep before init) */
#define DQMC_ERR_NUMERIC -5  /* non-finite input or a numerical failure on the device */
#define DQMC_ERR_IO -6       /* a file could not be opened, written or parsed */

/* dqmc_chain_info.status / dqmc_job_result.chain_status: numerical failures a chain met on
 * the device (bit flags; the reference stops the run on a failed decomposition,
 * include/utils/linear_algebra.hpp:70-74).  A chain with a fatal bit set does no further
 * sweeps (its state is kept as it was when the failure was seen); the wrap-limit bit is a
 * warning and the chain goes on. */
#define DQMC_STATUS_NONFINITE 1   /* a recomputed or wrapped G held inf or NaN (fatal) */
#define DQMC_STATUS_SINGULAR 2    /* an LU pivot of the recompute was zero, subnormal or non-finite (fatal) */
#define DQMC_STATUS_WRAP_LIMIT 4  /* the wrap error exceeded the batch's limit (dqmc_set_wrap_error_limit) */
#define DQMC_STATUS_PAIR_SYNC 8   /* a paired chain's partner workgroup did not answer within 30 s (fatal) */
#define DQMC_STATUS_FATAL (DQMC_STATUS_NONFINITE | DQMC_STATUS_SINGULAR | DQMC_STATUS_PAIR_SYNC)
#define DQMC_DEFAULT_WRAP_LIMIT 1e-3

/* models (Model::RepulsiveHubbard / Model::AttractiveHubbard) */
#define DQMC_MODEL_REPULSIVE 0
#define DQMC_MODEL_ATTRACTIVE 1

/* arithmetic of the dense B-matrix products (exp(-dtau K) as a matrix; the checkerboard
 * plaquette passes are FP64 in every mode): FP64 throughout, or FP32 matrix cores with FP32
 * accumulation for some of them while the UDV stacks, the (1 + UDV)^-1 recompute, the delayed
 * updates and the Metropolis ratio stay FP64:
 *   DQMC_PREC_MIXED          the slice products accumulated into the stacks and the wraps
 *   DQMC_PREC_FP32_WRAP      the wraps (and the G(t,0), G(0,t) propagation) only
 *   DQMC_PREC_FP32_PRODUCT   the slice products only
 * The FP32 modes do not hold G to 1e-6 of the FP64 chain on live chains: FP32 rounding in
 * the products carries into the fresh G, in the wraps into the wrapped G, and the updates
 * between stabilizations amplify either (the wrap error of live 16x16 chains at beta = 10,
 * U = 6, pace 8 reaches O(0.1)).  dqmc_chain_info.wrap_error reports it per chain and the
 * DQMC_STATUS_WRAP_LIMIT bit flags it. */
#define DQMC_PREC_FP64 0
#define DQMC_PREC_MIXED 1
#define DQMC_PREC_FP32_WRAP 2
#define DQMC_PREC_FP32_PRODUCT 3

/* sweep flags */
#define DQMC_SWEEP_MEASURE 1 /* equal-time measurement after every sweep (Dqmc::measure) */
#define DQMC_SWEEP_RECORD 2  /* keep per-slice G, sign and accept bits of the last sweep */
#define DQMC_SWEEP_FROZEN 4  /* no field updates: propagate, stabilize and measure a fixed
                                configuration (as the reference's sweep_for_dynamic_greens does) */
#define DQMC_SWEEP_PROFILE 8 /* accumulate device time per sweep phase (dqmc_chain_info.phase_seconds) */
#define DQMC_SWEEP_DYNAMIC 16 /* every 0->beta sweep is the time-displaced one of
                                 DqmcWalker::sweep_for_dynamic_greens (src/dqmc_walker.cpp:424-524):
                                 fields fixed, G(t,0) and G(0,t) propagated and recomputed with G(t,t);
                                 with DQMC_SWEEP_MEASURE it accumulates the dynamic observables
                                 (MeasureHandler::dynamic_measure) instead of the equal-time ones, as
                                 Dqmc::sweep_forth_and_back does (src/dqmc.cpp:48-70); the beta->0
                                 sweeps stay updating sweeps with equal-time measurements */

/* equal-time observables, index order of dqmc_get_measure */
#define DQMC_OBS_SIGN 0
#define DQMC_OBS_FILLING 1
#define DQMC_OBS_DOUBLE_OCCUPANCY 2
#define DQMC_OBS_KINETIC_ENERGY 3
#define DQMC_OBS_MOMENTUM_DISTRIBUTION 4
#define DQMC_OBS_LOCAL_SPIN_CORR 5
#define DQMC_OBS_SDW 6
#define DQMC_OBS_CDW 7
#define DQMC_OBS_SWAVE 8
#define DQMC_NUM_OBSERVABLES 9

typedef struct dqmc_context* dqmc_ctx_t;
typedef struct dqmc_batch* dqmc_batch_t;

/* One Markov chain.  Mirrors the TOML keys of src/dqmc_initializer.cpp:27-338:
 * Model.type/params, Lattice.cell (square, side x side), CheckerBoard.whether_or_not,
 * MonteCarlo.beta/time_size/stabilization_pace, Lattice.momentum. */
typedef struct {
    int model;               /* DQMC_MODEL_* ; attractive takes |U| (src/model/attractive_hubbard.cpp:37) */
    int side;                /* square lattice side length, N = side*side, 2 <= side, N <= 448 */
    int time_size;           /* L (dtau = beta / L, src/dqmc_walker.cpp:30) */
    int stabilization_pace;  /* p */
    int checkerboard;        /* 0 dense exp(-dtau K), 1 checkerboard break-up (even side only) */
    int delay;               /* delayed-update depth k, 1..32 (a block of the submatrix walk takes at most 32
                                sites: one per lane of a half-wavefront and spin); larger k is DQMC_ERR_ARG;
                                1 applies every flip at once.  Blocks take fewer sites where one spin's
                                k x N block of G does not fit in the workgroup's LDS (28 at 21 x 21); the
                                cap depends on N only, never on the launch mapping */
    int momentum_index;      /* k-star index for momentum_distribution/SDW/CDW; -1 = M point (pi,pi) */
    int precision;           /* DQMC_PREC_* */
    double beta;
    double hopping_t;
    double onsite_u;
    double chemical_potential;
    long long seed;          /* Utils::Random::set_seed(seed) before the random fields */
} dqmc_chain_params;

typedef struct {
    double sign;             /* DqmcWalker::ConfigSign() */
    double wrap_error;       /* DqmcWalker::WrapError() */
    unsigned long long rng;  /* engine state (minstd_rand0) */
    int pos;                 /* 0: at slice 0, time_size: at slice beta */
    int status;              /* 0 ok, else DQMC_STATUS_* bits */
    long long accepted;      /* accepted flips in the last sweep launch */
    long long sweeps;        /* single-direction sweeps done since init */
    double phase_seconds[8]; /* wrap, Metropolis, slice product, stack push, fresh G, measurement,
                                and two parts of those: the Metropolis in-block walk, the QR/LU
                                panel factorizations (accumulated over DQMC_SWEEP_PROFILE sweeps
                                since init) */
    int paired;              /* 1 if the last sweep launch ran the chain's spins on two workgroups */
    int kron;                /* 1 if the chain's dense B products run in Kronecker form: exp(-dtau K) =
                                X (x) Y over the two lattice directions (side <= 16, checked against
                                the chain's exp(K) within 1e-13 of its largest entry); 0 for dense
                                N x N products or the checkerboard break-up */
} dqmc_chain_info;

/* context: device selection, error string */
int dqmc_context_create(int device, dqmc_ctx_t* out);
int dqmc_context_destroy(dqmc_ctx_t ctx);
const char* dqmc_context_error(dqmc_ctx_t ctx);

/* device bytes a batch of these chains needs (admission by footprint) */
int dqmc_batch_footprint(int n_chains, const dqmc_chain_params* params, int record, size_t* bytes);

/* create a batch of independent chains resident on the context's device */
int dqmc_batch_create(dqmc_ctx_t ctx, int n_chains, const dqmc_chain_params* params, int record, dqmc_batch_t* out);
int dqmc_batch_destroy(dqmc_batch_t b);
int dqmc_batch_size(dqmc_batch_t b);
/* replay the persistent sweep launch from a captured HIP graph (re-captured
 * when n_sweeps or flags change); no counterpart in the reference */
int dqmc_batch_use_graph(dqmc_batch_t b, int enable);

/* How a sweep launch maps chains onto workgroups; no counterpart in the reference (its
 * walker runs one chain per process, src/dqmc_walker.cpp:212-216 runs the two spins one
 * after the other).
 *   DQMC_LAUNCH_SINGLE  one workgroup per chain (both spins in it)
 *   DQMC_LAUNCH_PAIR    two co-resident workgroups per repulsive chain, one per spin: wraps,
 *                       slice products, stack pushes and recomputes of the two spins run side
 *                       by side; the Metropolis walk runs in both, exchanging the k x k block of
 *                       each spin per delayed-update block.  The two workgroups of a chain
 *                       meet at the start; one that finds no partner within 100 us runs the
 *                       chain alone, so a pair launch never waits on workgroups that are not
 *                       running.  Attractive chains run alone in one workgroup (the other
 *                       leaves); time-displaced sweeps always run SINGLE.  Needs a repulsive
 *                       chain and 2 x (chains rounded up to 8) workgroups resident at once.  Not
 *                       replayed from a graph.  Results equal SINGLE's bit for bit however the
 *                       chains paired.
 *   DQMC_LAUNCH_QUAD    reserved: PAIR plus a helper workgroup per spin for half of every
 *                       dense product; not available (DQMC_ERR_ARG), its launches did not finish
 *   DQMC_LAUNCH_AUTO    PAIR when the batch qualifies, else SINGLE (the default) */
#define DQMC_LAUNCH_AUTO 0
#define DQMC_LAUNCH_SINGLE 1
#define DQMC_LAUNCH_PAIR 2
#define DQMC_LAUNCH_QUAD 3
/* DQMC_ERR_ARG when PAIR is asked of a batch that does not qualify, and for QUAD */
int dqmc_batch_set_launch_mode(dqmc_batch_t b, int mode);
/* the mapping the last sweep launch used (DQMC_LAUNCH_SINGLE or DQMC_LAUNCH_PAIR; 0 before any) */
int dqmc_batch_last_launch_mode(dqmc_batch_t b, int* mode);

/* fields: random (seed, reference draw order), or explicit */
int dqmc_fields_random(dqmc_batch_t b);
int dqmc_set_fields(dqmc_batch_t b, int chain, const int8_t* fields);
int dqmc_get_fields(dqmc_batch_t b, int chain, int8_t* fields);
/* engine state of a chain's random stream (continues after the field draws) */
int dqmc_set_rng_state(dqmc_batch_t b, int chain, unsigned long long state);

/* dense exp(-dtau K) / exp(+dtau K) of a chain (override / inspect) */
int dqmc_set_expk(dqmc_batch_t b, int chain, const double* expK, const double* invExpK);
int dqmc_get_expk(dqmc_batch_t b, int chain, double* expK, double* invExpK);

/* stacks, G(0) and sign for the current fields */
int dqmc_init(dqmc_batch_t b);

/* wrap error above which a chain's status gets DQMC_STATUS_WRAP_LIMIT (default
 * DQMC_DEFAULT_WRAP_LIMIT; the reference only reports it, DqmcWalker::WrapError) */
int dqmc_set_wrap_error_limit(dqmc_batch_t b, double limit);

/* n single-direction sweeps per chain, alternating 0->beta and beta->0 */
int dqmc_sweep(dqmc_batch_t b, int n_sweeps, int flags);
int dqmc_synchronize(dqmc_batch_t b);
/* device time of the last dqmc_sweep launch (HIP events on the batch stream) */
int dqmc_last_sweep_ms(dqmc_batch_t b, float* ms);

/* accessors */
int dqmc_get_green(dqmc_batch_t b, int chain, int spin, double* out);
int dqmc_get_slice_green(dqmc_batch_t b, int chain, int slot, int spin, double* out);
int dqmc_get_slice_signs(dqmc_batch_t b, int chain, double* out);
int dqmc_get_accepts(dqmc_batch_t b, int chain, uint8_t* out);
int dqmc_get_info(dqmc_batch_t b, int chain, dqmc_chain_info* out);

/* equal-time accumulators: sums and counts per observable (normalize per bin as
 * MeasureHandler::normalize_stats, src/measure/measure_handler.cpp:170-189) */
int dqmc_get_measure(dqmc_batch_t b, int chain, double* sums, double* counts);
/* MeasureHandler::equaltime_measure (src/measure/measure_handler.cpp:136-150) on the
 * slices kept by the last DQMC_SWEEP_RECORD sweep (batch created with record = 1) */
int dqmc_measure_recorded(dqmc_batch_t b);
/* clears the equal-time and the dynamic accumulators */
int dqmc_clear_measure(dqmc_batch_t b);

/* Time-displaced path.  G(t,0) (which = 0) or G(0,t) (which = 1) of slot t-1 kept by the
 * last DQMC_SWEEP_DYNAMIC | DQMC_SWEEP_RECORD sweep, column-major N x N
 * (DqmcWalker::Greent0Up/Dn(t), Green0tUp/Dn(t), include/dqmc_walker.h; the vectors filled at
 * src/dqmc_walker.cpp:455-463, :500-505). */
int dqmc_get_dynamic_green(dqmc_batch_t b, int chain, int slot, int spin, int which, double* out);
/* Dynamic accumulators (sums over DQMC_SWEEP_DYNAMIC | DQMC_SWEEP_MEASURE sweeps, each term
 * weighted by the configuration sign; normalize per bin as the equal-time ones):
 *   greens_k[k + n_k t]  G(k,t) for the k-stars k_index[0..n_k) (measure_greens_functions,
 *                        src/measure/measure_methods.cpp:280-306; the reference's MomentumList)
 *   dos[t]               density of states D(t) (:313-325)
 *   dss[t]               dynamic spin susceptibility (:413-461)
 *   scalars[0..2]        superfluid stiffness rho_s (:335-405), sign sum (:264-272), count
 * for t = 0..L-1.  Any output pointer may be NULL (greens_k only when n_k = 0). */
int dqmc_get_dynamic_measure(dqmc_batch_t b, int chain, int n_k, const int* k_index, double* greens_k, double* dos,
                             double* dss, double* scalars);
/* number of k-stars of the chain's lattice; kxy (may be NULL) receives (kx, ky) per k-star
 * (Lattice::Square::kStars, src/lattice/square.cpp:42-66) */
int dqmc_num_kstars(dqmc_batch_t b, int chain, double* kxy);

/* ------------------------------------------------------------------------
 * On-disk formats of the reference (QuantumMonteCarlo::DqmcIO, include/dqmc_io.h).
 * Host-only (no device work); byte-compatible with the reference's files.
 * ------------------------------------------------------------------------ */
/* HS-field file (DqmcIO::output_bosonic_fields, include/dqmc_io.h:478-534): "L N" then one
 * "t i s" line per field; fields[t * space_size + i] as in dqmc_set_fields */
int dqmc_write_fields_file(const char* path, int time_size, int space_size, const int8_t* fields);
/* DqmcIO::read_bosonic_fields_from_file (include/dqmc_io.h:537-623): DQMC_ERR_ARG when the
 * header's sizes differ from the chain's, DQMC_ERR_IO for unreadable files or entries */
int dqmc_read_fields_file(const char* path, int time_size, int space_size, int8_t* fields);
/* DqmcIO::output_k_stars (include/dqmc_io.h:432-455) and output_imaginary_time_grids (:458-475) */
int dqmc_write_kstars_file(const char* path, int side);
int dqmc_write_tgrids_file(const char* path, int time_size, double beta);
/* DqmcIO::output_observable, file branch (include/dqmc_io.h:305-355): rank 0 scalar, 1 vector
 * (rows), 2 matrix (rows x cols, element (i, j) at i + rows j); mean, error, relative error */
int dqmc_write_observable_file(const char* path, int rank, int rows, int cols, const double* mean, const double* err);
/* DqmcIO::output_observable_in_bins (include/dqmc_io.h:365-429): bins[b * size + element] */
int dqmc_write_observable_bins_file(const char* path, int rank, int n_bins, int rows, int cols, const double* bins);

/* Delayed-update depth autotuner (no counterpart in the reference, which
 * applies every flip at once): times `sweeps` sweeps of `n_chains` chains of
 * this shape for each candidate depth and returns the fastest in *best
 * (ms_out[c] = device time of candidate c, may be NULL). */
int dqmc_autotune_delay(dqmc_ctx_t ctx, const dqmc_chain_params* shape, int n_chains, const int* candidates,
                        int n_candidates, int sweeps, int* best, float* ms_out);
/* the faster chain -> workgroup mapping for n_chains chains of this shape: times `sweeps` sweeps
 * with DQMC_LAUNCH_SINGLE and, when the batch qualifies, DQMC_LAUNCH_PAIR; *best receives the
 * faster, ms_out[0..1] (may be NULL) the launch times (ms_out[1] = -1 when pairs do not apply).
 * No counterpart in the reference (one chain per process). */
int dqmc_autotune_launch(dqmc_ctx_t ctx, const dqmc_chain_params* shape, int n_chains, int sweeps, int* best,
                         float* ms_out);

/* ------------------------------------------------------------------------
 * Serving: continuous batching of DQMC jobs.  No counterpart in the
 * reference (which runs one fixed chain per MPI rank, src/dqmc_main.cpp);
 * each job runs the reference's schedule (Dqmc::thermalize then
 * Dqmc::measure, src/dqmc.cpp:73-151) on one chain.
 * ------------------------------------------------------------------------ */
typedef struct dqmc_server* dqmc_server_t;

#define DQMC_JOB_DYNAMIC 1   /* measure the dynamic observables too: every measured 0->beta sweep is the
                                time-displaced one (Dqmc::sweep_forth_and_back, src/dqmc.cpp:48-70) */

typedef struct {
    dqmc_chain_params chain;
    int warmup_sweeps;    /* unmeasured single-direction sweeps */
    int measure_sweeps;   /* measured sweeps, a multiple of bin_size */
    int bin_size;         /* measured sweeps per bin (even with DQMC_JOB_DYNAMIC) */
    int options;          /* DQMC_JOB_* */
} dqmc_job;

typedef struct {
    long long id;
    int status;           /* 0 done, 1 cancelled, < 0 error code (DQMC_ERR_NUMERIC: a fatal chain_status) */
    int chain_status;     /* DQMC_STATUS_* bits of the job's chain */
    int n_bins;
    double t_submit, t_start, t_finish;  /* seconds since the server started */
    long long sweeps;
    double wrap_error;
    double mean[DQMC_NUM_OBSERVABLES];   /* Observable::analyse over the bins */
    double error[DQMC_NUM_OBSERVABLES];
} dqmc_job_result;

typedef struct {
    long long submitted, completed, active, queued, chain_sweeps, launches;
    size_t bytes_in_use, bytes_peak, budget;
    double elapsed_s;
} dqmc_server_stats;

/* hbm_budget: device bytes for resident chains (one pool, allocated once);
 * max_active: resident chains at most; n_lanes: 1..4 size-class streams */
int dqmc_server_create(dqmc_ctx_t ctx, size_t hbm_budget, int max_active, int n_lanes, dqmc_server_t* out);
int dqmc_server_submit(dqmc_server_t s, const dqmc_job* job, long long* id);
int dqmc_server_cancel(dqmc_server_t s, long long id);
int dqmc_server_poll(dqmc_server_t s, dqmc_job_result* out, int max_results, int* n_out);
int dqmc_server_wait_idle(dqmc_server_t s, double timeout_s);
/* Dynamic results of a finished DQMC_JOB_DYNAMIC job (kept until the server is destroyed):
 * mean and error over the bins of [|dynamic sign|, rho_s, G(k,t) over every k-star
 * (n_k x L, k fastest), D(t) (L), DSS(t) (L)], 2 + n_k L + 2 L values each.  With mean = NULL
 * only *n_k and *time_size are returned.  DQMC_ERR_ARG if the job has no dynamic results. */
int dqmc_server_get_dynamic(dqmc_server_t s, long long id, int* n_k, int* time_size, double* mean, double* err);
int dqmc_server_get_stats(dqmc_server_t s, dqmc_server_stats* out);
const char* dqmc_server_error(dqmc_server_t s);
int dqmc_server_destroy(dqmc_server_t s);

/* ------------------------------------------------------------------------
 * Single-kernel entry points (host buffers in, host buffers out).
 * ------------------------------------------------------------------------ */

/* Workgroup GEMM of the B-matrix products (ModelBase::mult_expK_*,
 * src/model/model_base.cpp:7-35, and the products of
 * include/utils/numerical_stable.hpp:222-238):
 *   C = diag(rs) op(A) diag(ms) op(B) diag(cs) + beta C,  column-major,
 *   op(A) M x K, op(B) K x Nc; rs/ms/cs may be NULL.
 * variant 0: vector FP64, 1: MFMA FP64, 2: MFMA FP32 (inputs rounded to FP32),
 * 3: MFMA FP64 with A staged in LDS (M*K <= 13824), 4: A and B staged in LDS
 * (M*K + K*Nc <= 13824), 5 and 6: MFMA FP64 on wide tiles (128 x 128 and
 * 160 x 160 per workgroup, the form the engine uses for its N x N x N products;
 * bit-identical to variant 1).  The scalings rs, ms, cs are staged in LDS too,
 * so M, K, Nc <= 448 where they are given (variants 1-6). */
int dqmc_k_gemm(dqmc_ctx_t ctx, int variant, int M, int Nc, int K, int transA, int transB, const double* A,
                const double* B, double* C, double beta, const double* rs, const double* ms, const double* cs);

/* One slice's B-matrix product on the N x N matrix M, in place, for the chain `p`
 * with fields (time_size x N, +-1) and, for the dense kinetic term, exp(-dtau K)
 * and its inverse (NULL: the closed form the batch would use).  Replaces
 * RepulsiveHubbard / AttractiveHubbard::mult_B_from_left, mult_B_from_right,
 * mult_invB_from_left, mult_invB_from_right, mult_transB_from_left
 * (src/model/repulsive_hubbard.cpp:158-245, src/model/attractive_hubbard.cpp:159-245)
 * and the wraps of DqmcWalker::wrap_from_0_to_beta / wrap_from_beta_to_0
 * (src/dqmc_walker.cpp:208-245).  t is the reference's time_index, 0 <= t <= time_size:
 * B(t) takes field row t - 1, and t = 0 the last row (the slice at beta); spin +1 / -1;
 * op 0 B M, 1 M B, 2 B^-1 M, 3 M B^-1, 4 B^T M, 5 B M B^-1, 6 B^-1 M B. */
int dqmc_k_wrap(dqmc_ctx_t ctx, const dqmc_chain_params* p, const int8_t* fields, const double* expK,
                const double* invExpK, int t, int spin, int op, double* M);

/* The single-flip update of the equal-time Green's functions (column-major N x N,
 * in place) for the field at (t, i) of `fields`, before that field is flipped
 * (RepulsiveHubbard::update_greens_function, src/model/repulsive_hubbard.cpp:127-155;
 * AttractiveHubbard::update_greens_function, src/model/attractive_hubbard.cpp:128-156:
 * the attractive model updates both matrices with the spin-up factor).  G_dn may be
 * NULL. */
int dqmc_k_rank1(dqmc_ctx_t ctx, const dqmc_chain_params* p, const int8_t* fields, int t, int i, double* G_up,
                 double* G_dn);

/* Utils::SvdStack::push (src/svd_stack.cpp:34-61) in the engine's QR form:
 * U diag(D) V^T = A (U_in diag(D_in) V_in^T), or of A alone when U_in is NULL (an
 * empty stack).  U is orthogonal and D > 0 graded as singular values are; V is not
 * orthogonal (it carries the triangular factors), so U, D, V equal the reference's
 * SVD factors only through their product and the scales of D.  N = side * side with
 * 2 <= side and N <= 448 (the engine's sizes). */
int dqmc_k_udv_push(dqmc_ctx_t ctx, int N, const double* A, const double* U_in, const double* D_in,
                    const double* V_in, double* U, double* D, double* V);

/* NumericalStable::compute_equaltime_greens (include/utils/numerical_stable.hpp:186-238):
 * G = (1 + B_L B_R^T)^-1 with B_L = U_L diag(D_L) V_L^T (left stack top) and
 * B_R = U_R diag(D_R) V_R^T (right stack top); pass NULL for U, D and V of an empty
 * stack (not both).  U_L, U_R orthogonal, D > 0.  Same sizes as dqmc_k_udv_push. */
int dqmc_k_greens_recompute(dqmc_ctx_t ctx, int N, const double* UL, const double* DL, const double* VL,
                            const double* UR, const double* DR, const double* VR, double* G);

/* ------------------------------------------------------------------------
 * Single-kernel entry points on caller-provided device buffers (the form the
 * reference's per-site plugin calls take once the walker's G lives on the device).
 * A workspace holds the scratch of one chain shape (descriptor, two stack slots, GEMM and
 * factorization scratch) and is allocated once; the calls themselves allocate nothing and
 * copy nothing from the host: each is enqueued on `stream` (a hipStream_t; NULL is the
 * default stream) and returns at once.  Calls on one workspace must not overlap (one stream
 * at a time).  Pointers are device pointers; matrices column-major N x N, fields int8
 * (t, i) at t*N+i as in dqmc_set_fields.
 * ------------------------------------------------------------------------ */
typedef struct dqmc_workspace* dqmc_workspace_t;

/* scratch and constants for chains of this shape (Model::ModelBase::initial,
 * initial_params, initial_KV_matrices: include/model/model_base.h:120-122) */
int dqmc_workspace_create(dqmc_ctx_t ctx, const dqmc_chain_params* p, dqmc_workspace_t* out);
int dqmc_workspace_destroy(dqmc_workspace_t w);
/* host copies of exp(-dtau K), exp(+dtau K) the workspace's dense products use
 * (ModelBase::initial_KV_matrices, src/model/repulsive_hubbard.cpp:53-72) */
int dqmc_workspace_expk(dqmc_workspace_t w, double* expK, double* invExpK);

/* dqmc_k_wrap on device buffers: op 0..6 on M in place with the field row of time index t
 * (0 <= t <= time_size, the reference's time_index); expK/invExpK may be NULL (the workspace's) */
int dqmc_kd_wrap(dqmc_workspace_t w, const int8_t* fields, const double* expK, const double* invExpK, int t,
                 int spin, int op, double* M, void* stream);
/* Model::ModelBase::get_update_ratio (include/model/model_base.h:146; src/model/repulsive_hubbard.cpp:112-124,
 * src/model/attractive_hubbard.cpp:112-125): *ratio (one device double) = the Metropolis ratio of
 * flipping the field at (t, i), 0 <= t < time_size, from G_up(i,i) and G_dn(i,i) (G_dn ignored,
 * may be NULL, for the attractive model) */
int dqmc_kd_update_ratio(dqmc_workspace_t w, const int8_t* fields, int t, int i, const double* G_up,
                         const double* G_dn, double* ratio, void* stream);
/* Model::ModelBase::update_greens_function (include/model/model_base.h:148) = dqmc_k_rank1 on
 * device buffers, before the field at (t, i) is flipped */
int dqmc_kd_rank1(dqmc_workspace_t w, const int8_t* fields, int t, int i, double* G_up, double* G_dn, void* stream);
/* Model::ModelBase::update_bosonic_field (include/model/model_base.h:150; src/model/repulsive_hubbard.cpp:102-109) */
int dqmc_kd_flip_field(dqmc_workspace_t w, int8_t* fields, int t, int i, void* stream);
/* dqmc_k_udv_push and dqmc_k_greens_recompute on device buffers (U_in / UL / UR NULL: empty
 * stack); *status (a device int, may be NULL) receives the DQMC_STATUS_* bits of the call */
int dqmc_kd_udv_push(dqmc_workspace_t w, const double* A, const double* U_in, const double* D_in, const double* V_in,
                     double* U, double* D, double* V, int* status, void* stream);
int dqmc_kd_greens_recompute(dqmc_workspace_t w, const double* UL, const double* DL, const double* VL,
                             const double* UR, const double* DR, const double* VR, double* G, int* status,
                             void* stream);

/* The per-site plugin calls of the reference's Metropolis loop (DqmcWalker::metropolis_update,
 * src/dqmc_walker.cpp:173-197) on an initialized batch chain's resident G and fields,
 * synchronous on the batch stream: Model::ModelBase::get_update_ratio (*ratio on the host),
 * update_greens_function, update_bosonic_field (include/model/model_base.h:142-150) for the field
 * at (t, i), 0 <= t < time_size; and the B-matrix products of the wraps on the chain's own G
 * (op 0..6 as dqmc_k_wrap, time index 0 <= t <= time_size; DqmcWalker::wrap_from_0_to_beta /
 * wrap_from_beta_to_0, src/dqmc_walker.cpp:208-237).  A caller that changes the chain this way
 * keeps its own stacks (as the reference's walker does); the engine's sweeps need dqmc_init
 * again afterwards. */
int dqmc_chain_update_ratio(dqmc_batch_t b, int chain, int t, int i, double* ratio);
int dqmc_chain_update_green(dqmc_batch_t b, int chain, int t, int i);
int dqmc_chain_flip_field(dqmc_batch_t b, int chain, int t, int i);
int dqmc_chain_mult_green(dqmc_batch_t b, int chain, int t, int spin, int op);
/* *DqmcWalker::GreenttUp/Dn() = G (the reference assigns the fresh G this way,
 * src/dqmc_walker.cpp:302-303): host matrix into the chain's resident G */
int dqmc_set_green(dqmc_batch_t b, int chain, int spin, const double* G);

/* Device buffers of a batch chain's state (DqmcWalker::GreenttUp/Dn() by reference,
 * include/dqmc_walker.h:136-140): G of each spin (the same matrix twice for the attractive
 * model) and the fields, plus the batch's stream; valid until the batch is destroyed.  Work
 * enqueued on them must be ordered with the batch's own calls through that stream. */
int dqmc_chain_device_buffers(dqmc_batch_t b, int chain, double** G_up, double** G_dn, int8_t** fields,
                              void** stream);

/* dqmc_k_greens_recompute on caller device buffers with no allocation at all.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its scratch is a static
 * device array sized for DQMC_NOALLOC_MAX_N sites (about 3.7 MB, held on each device the
 * library runs on until the process ends), so N > DQMC_NOALLOC_MAX_N is DQMC_ERR_ARG.  Every call
 * on one device shares that array: calls that overlap on one device (two streams, two threads)
 * race.  Enqueued on `stream` (a hipStream_t, NULL the default stream); *status (a device int,
 * may be NULL) receives the call's DQMC_STATUS_* bits.  Same arguments and result as
 * dqmc_kd_greens_recompute. */
#define DQMC_NOALLOC_MAX_N 256
int dqmc_k_greens_recompute_noalloc(dqmc_ctx_t ctx, int N, const double* UL, const double* DL, const double* VL,
                                    const double* UR, const double* DR, const double* VR, double* G, int* status,
                                    void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DQMC_AMD_H */
