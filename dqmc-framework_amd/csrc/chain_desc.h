// Per-chain descriptor and mutable state shared by the host runtime and the
// HIP kernels.  Plain C layout: one array of ChainDesc lives in device memory
// and every workgroup of a sweep launch reads the descriptor of its chain.
#pragma once
#include <stdint.h>

#include "../../include/dqmc_amd.h"   // DQMC_STATUS_* bits of ChainState::status

#define DQMC_NT 256          // threads per workgroup (4 wavefronts of 64)
#define DQMC_NMAX 448        // largest lattice (sites) a chain may have (side <= 21)
#define DQMC_KMAX 32         // largest delayed-update depth (a Metropolis block holds its sites in the 32
                             // lanes of a half-wavefront per spin)
#define DQMC_PANEL_DOUBLES 13824   // LDS panel of a workgroup (108 KB)
// spin-pair exchange area of a repulsive chain: the k x k walk blocks [spin][parity][KMAX^2], then
// the per-sweep state slots [spin][parity][2] (status bits, wrap error)
#define DQMC_PAIR_XBUF_DOUBLES (4 * DQMC_KMAX * DQMC_KMAX + 8)
#define DQMC_PAIR_FLAG_WORDS 96    // per chain: the two spins' sequence flags and the join word, 128 bytes apart

enum {
    DQMC_FLAG_MEASURE = 1,   // accumulate equal-time observables after each sweep
    DQMC_FLAG_RECORD = 2,    // keep G, sign per slice and accept bits of the last sweep
    DQMC_FLAG_FROZEN = 4,    // propagate and measure without field updates (fixed configuration)
    DQMC_FLAG_PROFILE = 8,   // accumulate per-phase device time into ChainState::phase
    DQMC_FLAG_DYNAMIC = 16,  // an upward sweep is the time-displaced one: no field updates, G(t,0) and
                             // G(0,t) propagated and recomputed with G(t,t), dynamic observables
};

// building blocks run alone by the single-kernel entry points (one workgroup, chain 0)
enum { DQMC_UNIT_BMUL = 0, DQMC_UNIT_RANK1 = 1, DQMC_UNIT_PUSH = 2, DQMC_UNIT_FRESH = 3, DQMC_UNIT_RATIO = 4,
       DQMC_UNIT_FLIP = 5, DQMC_UNIT_PUSH_EXT = 6, DQMC_UNIT_FRESH_EXT = 7 };
// caller buffers of the building blocks that run on device memory the caller owns
struct DqmcUnitPtrs {
    const double *A, *U_in, *D_in, *V_in;          // PUSH_EXT: pushed matrix and the stack top
    double *U, *D, *V;                             // PUSH_EXT: result
    const double *UL, *DL, *VL, *UR, *DR, *VR;     // FRESH_EXT: the two stack tops
    double* out;                                   // FRESH_EXT: G; RATIO: the ratio
    int* status;                                   // failure bits of the call (may be null)
};
// phases timed with DQMC_FLAG_PROFILE (100 MHz device wall clock ticks)
enum { DQMC_PH_WRAP = 0, DQMC_PH_METROPOLIS, DQMC_PH_PRODUCT, DQMC_PH_PUSH, DQMC_PH_FRESH, DQMC_PH_MEASURE,
       DQMC_PH_WALK,     // part of METROPOLIS: the in-block decisions and C updates (one wavefront)
       DQMC_PH_PANEL,    // part of PUSH and FRESH: QR and LU panel factorizations in LDS
       DQMC_PH_COUNT };

// number of accumulated equal-time quantities, in this order:
// sign, filling, double occupancy, kinetic energy, momentum distribution,
// local spin correlation, SDW structure factor, CDW structure factor, s-wave pairing
#define DQMC_NUM_OBS 9
#define DQMC_ACC_SCRATCH 12   // accumulator slot (of 16) that carries a per-site update ratio to the host

// dynamic accumulators of a chain (L slices, N sites), in this order:
// gr[L][N]  sign x sum over site pairs (i, j = i + r) of the spin-averaged G(t,0)(j,i), per
//           displacement r (slice entry 0 holds G(0,0)); G(k,t) and D(t) are linear in it
// dss[L]    dynamic spin susceptibility sums, rho (superfluid stiffness), sign, count
#define DQMC_DYN_ACC_SIZE(L, N) ((size_t)(L) * (N) + (L) + 3)

struct ChainState {
    uint64_t rng;        // minstd_rand0 state
    double sign;         // current configuration sign
    double wrap_error;   // max |G_fresh - G_wrapped| seen so far
    double sign_top;     // sign recorded at slice L of the running backward sweep
    int pos;             // 0: at slice 0 (next sweep goes up), L: at slice L (next goes down)
    int n_right;         // length of the right UDV stack (left length = n_stack - n_right)
    int status;          // DQMC_STATUS_* bits (0 ok)
    int accepted;        // accepted flips in the last launch (for reporting)
    unsigned long long phase[8];  // device ticks per phase (DQMC_FLAG_PROFILE)
    int paired;          // the last sweep launch ran the two spins on two workgroups
};

struct ChainDesc {
    int N, L, side, pace, n_stack, model, cb, kdelay, nspin, precision;
    int launch_sweeps;   // > 0: sweeps of this chain in the current launch (overrides the launch argument)
    int launch_flags;    // >= 0: flags of this chain in the current launch (overrides the launch argument)
    double hop_t;
    double wrap_limit;   // wrap error above which the chain is flagged DQMC_STATUS_WRAP_LIMIT
    double ev[2];        // exp(+alpha), exp(-alpha): exp(alpha*s) for s=+1, s=-1
    double dm1[2][2];    // [spin][s<0]: exp(-2 sigma alpha s) - 1 entering ratio and gamma
    double pref[2];      // attractive ratio prefactor exp(2 alpha s) for s=+1, s=-1 (1 for repulsive)
    double P[3];         // checkerboard plaquette entries e^{dtau mu/2}{ch*ch, ch*sh, sh*sh}
    double Pi[3];        // inverse plaquette entries e^{-dtau mu/2}{ch*ch, -ch*sh, sh*sh}
    // device pointers (all inside the batch arena)
    double* G;           // nspin x N*N   current equal-time Green's function
    double* Gw;          // nspin x N*N   product / fresh-G workspace
    double* Tmp;         // nspin x N*N   slice-product accumulated between stabilizations
    double* stkU;        // nspin x n_stack x N*N
    double* stkT;        // nspin x n_stack x N*N   (cumulative V^T of each level)
    double* stkD;        // nspin x n_stack x N
    double* W;           // nspin x 3 x N*N scratch for the stabilization (one set per spin)
    double* Wq;          // nspin x 3 x 16 x (N + 16) scratch of the blocked QR (GEMM panels and T factors)
    const double* expK;  // N*N (dense only)
    const double* invExpK;
    const double* cosq;  // N: Re e^{-i Q r_d} for displacement index d
    int8_t* field;       // L x N, (t,i) at t*N+i
    ChainState* st;
    double* acc;         // 16 doubles: DQMC_NUM_OBS sums, then scratch (DQMC_ACC_SCRATCH)
    double* Gt0;         // nspin x N*N   G(t,0) of the time-displaced sweep
    double* G0t;         // nspin x N*N   G(0,t)
    double* dacc;        // DQMC_DYN_ACC_SIZE(L, N) dynamic sums
    double* dscr;        // L x N + N: per-slice spin-density differences and the summed current vector
    const double* rhoq;  // N: cos(kx rx) - cos(kx ry) of displacement d = rx + side ry (stiffness factor)
    double* recG;        // L x nspin x N*N (RECORD)
    double* recSign;     // L (RECORD)
    uint8_t* recAcc;     // L x N (RECORD)
    double* recGt0;      // L x nspin x N*N (RECORD): G(t,0) per slot of the last time-displaced sweep
    double* recG0t;      // L x nspin x N*N (RECORD): G(0,t)
    double* xbuf;        // DQMC_PAIR_XBUF_DOUBLES (repulsive chains): spin-pair exchange blocks
    unsigned* xflag;     // DQMC_PAIR_FLAG_WORDS (repulsive chains): spin-pair flags, zeroed before a pair launch
    // dense chains whose exp(K) factors over the two lattice directions (side <= 16):
    // exp(-dtau K)[(x,y),(x',y')] = X[x][x'] Y[y][y'], site x + side y; four 16 x 16 column-major
    // matrices, zero padded: X, Y of exp(K), then X, Y of exp(K)^-1.  Null: dense N x N products.
    const double* kron;
};
#define DQMC_KRON_MAX_SIDE 16
