// refcpu — the oracle: a CPU FP64 restatement of the dqmc-framework equal-time
// DQMC sweep and its time-displaced sweep, used ONLY as test infrastructure (parity checker in tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py).  Nothing in
// the shipped library links, loads or calls this file.
//
// It restates, function by function, the reference files
//   src/lattice/square.cpp, src/model/{repulsive,attractive}_hubbard.cpp,
//   src/model/model_base.cpp, src/checkerboard/square.cpp, src/svd_stack.cpp,
//   include/utils/linear_algebra.hpp, include/utils/numerical_stable.hpp,
//   src/dqmc_walker.cpp, src/measure/measure_methods.cpp,
//   src/measure/measure_handler.cpp, include/measure/observable.h,
//   src/dqmc.cpp, src/random.cpp
// with the same formulas, the same operation order where it matters for the
// Markov chain (random-draw order, slice indexing, stabilization schedule,
// measurement slots) and the same quirks.  Third-party numerics that the
// reference takes from Eigen 3.4 / Intel MKL are re-derived here from their
// published algorithms:
//   * matrix exponential: Higham 2005 scaling-and-squaring Pade (the method of
//     Eigen's MatrixFunctions module),
//   * dgesvd: one-sided Jacobi SVD (Hestenes; Demmel-Veselic accuracy for
//     graded matrices) — the SVD is unique, so U S V^T agrees to rounding,
//   * FullPivHouseholderQR::solve and PartialPivLU::inverse: textbook
//     algorithms (Golub & Van Loan 5.4.2, 3.4.1),
//   * std::default_random_engine (libstdc++ = minstd_rand0) together with
//     std::generate_canonical / std::bernoulli_distribution: from the C++
//     standard's definitions [rand.eng.lcong], [rand.util.canonical].
//
// Parity pins for every function are listed in the tests (tests/test_oracle_*.py); the
// time-displaced functions (compute_dynamic_greens, compute_greens_b0, the orthogonal-U form
// compute_dynamic_greens_udt, Engine::sweep_dynamic, Engine::dynamic_measure) are pinned by
// U = 0 closed forms and by time-ordered traces in the many-body Fock space of a 2x2 lattice
// (tests/test_oracle_dynamic.py).
// One deliberate departure, on request (Engine::use_stable_sign; the CLI and the tests' binding
// turn it on): the initial configuration sign from the stabilized factorization of the stack
// (stable_green_sign) instead of sgn det G, which is rounding noise once G has eigenvalues below
// FP64 resolution; pinned against exact 50-digit determinants, against the reference's
// definition where that one is resolved, and against particle-hole symmetry
// (tests/test_oracle_sign.py).
// Built with -ffp-contract=off so the canonical-double computation of the
// random stream is not fused into FMAs.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace refcpu {

using std::vector;

// ----------------------------------------------------------------------------
// Column-major square matrix (same element layout as Eigen::MatrixXd).
// ----------------------------------------------------------------------------
struct Mat {
    int n = 0;
    vector<double> a;
    Mat() = default;
    explicit Mat(int n_) : n(n_), a((size_t)n_ * n_, 0.0) {}
    static Mat Identity(int n) {
        Mat m(n);
        for (int i = 0; i < n; ++i) m(i, i) = 1.0;
        return m;
    }
    double& operator()(int i, int j) { return a[(size_t)j * n + i]; }
    double operator()(int i, int j) const { return a[(size_t)j * n + i]; }
    double* col(int j) { return a.data() + (size_t)j * n; }
    const double* col(int j) const { return a.data() + (size_t)j * n; }
};

// C = A * B  (C must not alias A or B); axpy-ordered column loop.
static void gemm(const Mat& A, const Mat& B, Mat& C) {
    const int n = A.n;
    C.n = n;
    C.a.assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        double* c = C.col(j);
        for (int k = 0; k < n; ++k) {
            const double b = B(k, j);
            const double* a = A.col(k);
            for (int i = 0; i < n; ++i) c[i] += a[i] * b;
        }
    }
}
static Mat mul(const Mat& A, const Mat& B) { Mat C; gemm(A, B, C); return C; }
static Mat transpose(const Mat& A) {
    Mat T(A.n);
    for (int j = 0; j < A.n; ++j)
        for (int i = 0; i < A.n; ++i) T(j, i) = A(i, j);
    return T;
}

// ----------------------------------------------------------------------------
// Random numbers: std::default_random_engine of libstdc++ (minstd_rand0) and
// the draw of std::bernoulli_distribution (src/random.cpp:6-10,
// src/dqmc_walker.cpp:184, src/model/repulsive_hubbard.cpp:92-97).
// ----------------------------------------------------------------------------
struct Minstd0 {
    static constexpr uint64_t kA = 16807, kM = 2147483647ULL;
    uint64_t x = 1;
    // linear_congruential_engine::seed: state = s mod m, and 1 if that is 0.
    void seed(long long s) {
        // the reference passes an int that is converted to the engine's
        // unsigned result_type (unsigned long on LP64) before the reduction
        uint64_t u = (uint64_t)(int64_t)s;
        x = u % kM;
        if (x == 0) x = 1;
    }
    uint64_t next() { x = (kA * x) % kM; return x; }
};

// generate_canonical<double, 53>: R = max-min+1 = 2147483646, k = ceil(53/log2 R) = 2,
// S = sum_i (g_i - min) R^i, result S / R^k (clamped below 1).
static double canonical(Minstd0& g) {
    const double R = 2147483646.0;
    double sum = 0.0, tmp = 1.0;
    for (int k = 0; k < 2; ++k) {
        sum += (double)(g.next() - 1) * tmp;
        tmp *= R;
    }
    double ret = sum / tmp;
    if (ret >= 1.0) ret = std::nextafter(1.0, 0.0);
    return ret;
}
static bool bernoulli(Minstd0& g, double p) { return canonical(g) < p; }

// ----------------------------------------------------------------------------
// Lattice::Square (src/lattice/square.cpp:16-196)
// ----------------------------------------------------------------------------
struct Square {
    int L = 0, N = 0, nk = 0;
    vector<int> site_x, site_y;  // index2site
    vector<int> nn;              // N x 4, (i,dir) at i*4+dir; 0:+x 1:+y 2:-x 3:-y
    vector<int> disp;            // N x N, (i,j) at i*N+j  (pointing from i to j)
    vector<double> kx, ky;       // k stars
    vector<double> fourier;      // N x nk, (x,k) at x*nk+k
    Mat hopping;
    int gamma = 0, xpoint = 0, mpoint = 0;

    void init(int side) {
        L = side; N = side * side;
        site_x.resize(N); site_y.resize(N);
        for (int i = 0; i < N; ++i) { site_x[i] = i % L; site_y[i] = i / L; }  // :31-39
        nk = (int)((std::floor(L / 2.0) + 1) * (std::floor(L / 2.0) + 2) / 2);  // :47
        kx.clear(); ky.clear();
        for (int i = (int)std::ceil(L / 2.0); i <= L; ++i)                      // :58-65
            for (int j = (int)std::ceil(L / 2.0); j <= i; ++j) {
                kx.push_back((double)i / L * 2 * M_PI - M_PI);
                ky.push_back((double)j / L * 2 * M_PI - M_PI);
            }
        nn.resize((size_t)N * 4);
        for (int idx = 0; idx < N; ++idx) {                                    // :76-84
            const int x = idx % L, y = idx / L;
            nn[idx * 4 + 0] = ((x + 1) % L) + L * y;
            nn[idx * 4 + 2] = ((x - 1 + L) % L) + L * y;
            nn[idx * 4 + 1] = x + L * ((y + 1) % L);
            nn[idx * 4 + 3] = x + L * ((y - 1 + L) % L);
        }
        disp.resize((size_t)N * N);
        for (int i = 0; i < N; ++i)                                             // :91-104
            for (int j = 0; j < N; ++j) {
                const int dx = (site_x[j] - site_x[i] + L) % L;
                const int dy = (site_y[j] - site_y[i] + L) % L;
                disp[(size_t)i * N + j] = dx + dy * L;
            }
        gamma = 0;                                                             // :114-116
        xpoint = nk - (int)std::floor(L / 2.0) - 1;
        mpoint = nk - 1;
        fourier.resize((size_t)N * nk);
        for (int x = 0; x < N; ++x)                                             // :152-160
            for (int k = 0; k < nk; ++k)
                fourier[(size_t)x * nk + k] = std::cos((-site_x[x]) * kx[k] - site_y[x] * ky[k]);
        hopping = Mat(N);  // zero-initialised (the reference leaves it uninitialised, :166)
        for (int idx = 0; idx < N; ++idx) {                                    // :167-176
            const int xp = nn[idx * 4 + 0], yp = nn[idx * 4 + 1];
            hopping(idx, xp) += 1.0; hopping(xp, idx) += 1.0;
            hopping(idx, yp) += 1.0; hopping(yp, idx) += 1.0;
        }
    }
    double FourierFactor(int d, int k) const { return fourier[(size_t)d * nk + k]; }
};

// ----------------------------------------------------------------------------
// Dense linear algebra
// ----------------------------------------------------------------------------
// LU with partial pivoting, in place; returns permutation (row swaps) and the
// number of swaps.  Golub & Van Loan Alg. 3.4.1.
static void lu_factor(Mat& A, vector<int>& piv, int& nswaps) {
    const int n = A.n;
    piv.resize(n);
    nswaps = 0;
    for (int k = 0; k < n; ++k) {
        int p = k; double best = std::fabs(A(k, k));
        for (int i = k + 1; i < n; ++i) if (std::fabs(A(i, k)) > best) { best = std::fabs(A(i, k)); p = i; }
        piv[k] = p;
        if (p != k) { ++nswaps; for (int j = 0; j < n; ++j) std::swap(A(k, j), A(p, j)); }
        const double d = A(k, k);
        if (d != 0.0) for (int i = k + 1; i < n; ++i) A(i, k) /= d;
        for (int j = k + 1; j < n; ++j) {
            const double akj = A(k, j);
            if (akj == 0.0) continue;
            for (int i = k + 1; i < n; ++i) A(i, j) -= A(i, k) * akj;
        }
    }
}
// X = A^-1 B via the LU factors (B overwritten by X).
static void lu_solve(const Mat& LU, const vector<int>& piv, Mat& B) {
    const int n = LU.n;
    for (int c = 0; c < n; ++c) {
        double* b = B.col(c);
        for (int k = 0; k < n; ++k) if (piv[k] != k) std::swap(b[k], b[piv[k]]);
        for (int k = 0; k < n; ++k) { const double v = b[k]; for (int i = k + 1; i < n; ++i) b[i] -= LU(i, k) * v; }
        for (int k = n - 1; k >= 0; --k) { b[k] /= LU(k, k); const double v = b[k]; for (int i = 0; i < k; ++i) b[i] -= LU(i, k) * v; }
    }
}
// Eigen MatrixBase::inverse() for dynamic sizes = PartialPivLU solve with I.
static Mat inverse(const Mat& A) {
    Mat LU = A; vector<int> piv; int ns;
    lu_factor(LU, piv, ns);
    Mat X = Mat::Identity(A.n);
    lu_solve(LU, piv, X);
    return X;
}
// the sign of det(A) from its LU: interchange parity times the pivot signs (no product of
// pivots, so it neither underflows nor overflows)
static double det_sign(const Mat& A) {
    Mat LU = A; vector<int> piv; int ns;
    lu_factor(LU, piv, ns);
    double d = (ns % 2) ? -1.0 : 1.0;
    for (int i = 0; i < A.n; ++i) if (LU(i, i) < 0) d = -d;
    return d;
}
static double det_sign_times(const Mat& A) {  // returns the determinant (for the sign)
    Mat LU = A; vector<int> piv; int ns;
    lu_factor(LU, piv, ns);
    double d = (ns % 2) ? -1.0 : 1.0;
    for (int i = 0; i < A.n; ++i) d *= LU(i, i);
    return d;
}

// Full-pivoting Householder QR solve of H X = B (Eigen FullPivHouseholderQR::solve,
// include/utils/numerical_stable.hpp:151).  Golub & Van Loan 5.4.2 with complete pivoting.
static Mat fullpiv_qr_solve(const Mat& Hin, const Mat& Bin) {
    const int n = Hin.n;
    Mat A = Hin, B = Bin;
    vector<int> colperm(n);
    for (int i = 0; i < n; ++i) colperm[i] = i;
    vector<double> v(n);
    for (int k = 0; k < n; ++k) {
        int bi = k, bj = k; double best = -1.0;
        for (int j = k; j < n; ++j)
            for (int i = k; i < n; ++i) if (std::fabs(A(i, j)) > best) { best = std::fabs(A(i, j)); bi = i; bj = j; }
        if (bi != k) {  // row transposition applied to A and to the right-hand sides
            for (int j = 0; j < n; ++j) std::swap(A(k, j), A(bi, j));
            for (int j = 0; j < n; ++j) std::swap(B(k, j), B(bi, j));
        }
        if (bj != k) { for (int i = 0; i < n; ++i) std::swap(A(i, k), A(i, bj)); std::swap(colperm[k], colperm[bj]); }
        // Householder reflector for A(k:n, k)
        double norm2 = 0.0;
        for (int i = k; i < n; ++i) norm2 += A(i, k) * A(i, k);
        const double nrm = std::sqrt(norm2);
        if (nrm == 0.0) continue;
        const double alpha = (A(k, k) > 0) ? -nrm : nrm;
        for (int i = k; i < n; ++i) v[i] = A(i, k);
        v[k] -= alpha;
        double vnorm2 = 0.0;
        for (int i = k; i < n; ++i) vnorm2 += v[i] * v[i];
        if (vnorm2 == 0.0) continue;
        const double tau = 2.0 / vnorm2;
        for (int j = k; j < n; ++j) {
            double s = 0.0;
            for (int i = k; i < n; ++i) s += v[i] * A(i, j);
            s *= tau;
            for (int i = k; i < n; ++i) A(i, j) -= s * v[i];
        }
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int i = k; i < n; ++i) s += v[i] * B(i, j);
            s *= tau;
            for (int i = k; i < n; ++i) B(i, j) -= s * v[i];
        }
    }
    // back substitution R Y = Q^T P B, then X = colperm Y
    Mat X(n);
    for (int c = 0; c < n; ++c) {
        vector<double> y(n);
        for (int k = n - 1; k >= 0; --k) {
            double s = B(k, c);
            for (int j = k + 1; j < n; ++j) s -= A(k, j) * y[j];
            y[k] = s / A(k, k);
        }
        for (int k = 0; k < n; ++k) X(colperm[k], c) = y[k];
    }
    return X;
}

// One-sided Jacobi SVD A = U diag(s) V^T, s descending (stands in for
// LAPACKE_dgesvd('A','A'), include/utils/linear_algebra.hpp:41-80; V is returned, not V^T).
static void svd(const Mat& A, Mat& U, vector<double>& s, Mat& V) {
    const int n = A.n;
    Mat W = A;
    Mat Vm = Mat::Identity(n);
    const double eps = 2.220446049250313e-16;
    for (int sweep = 0; sweep < 80; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double alpha = 0.0, beta = 0.0, gamma = 0.0;
                const double* wp = W.col(p); const double* wq = W.col(q);
                for (int i = 0; i < n; ++i) { alpha += wp[i] * wp[i]; beta += wq[i] * wq[i]; gamma += wp[i] * wq[i]; }
                if (gamma == 0.0 || std::fabs(gamma) <= eps * std::sqrt(alpha * beta)) continue;
                rotated = true;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                double* Wp = W.col(p); double* Wq = W.col(q);
                for (int i = 0; i < n; ++i) { const double x = Wp[i], y = Wq[i]; Wp[i] = c * x - sn * y; Wq[i] = sn * x + c * y; }
                double* Vp = Vm.col(p); double* Vq = Vm.col(q);
                for (int i = 0; i < n; ++i) { const double x = Vp[i], y = Vq[i]; Vp[i] = c * x - sn * y; Vq[i] = sn * x + c * y; }
            }
        if (!rotated) break;
    }
    vector<double> norms(n);
    for (int j = 0; j < n; ++j) { double ss = 0.0; for (int i = 0; i < n; ++i) ss += W(i, j) * W(i, j); norms[j] = std::sqrt(ss); }
    vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return norms[x] > norms[y]; });
    U = Mat(n); V = Mat(n); s.assign(n, 0.0);
    for (int jj = 0; jj < n; ++jj) {
        const int j = order[jj];
        s[jj] = norms[j];
        for (int i = 0; i < n; ++i) { U(i, jj) = norms[j] > 0 ? W(i, j) / norms[j] : (i == jj ? 1.0 : 0.0); V(i, jj) = Vm(i, j); }
    }
}

// Householder QR with column pivoting, A P = Q R; returns U = Q (explicit),
// d = |diag R| and V with V^T = diag(d)^-1 R P^T.  Used only for the
// factorization-invariance pin (the GPU stacks use this factorization).
static void qrp_udv(const Mat& Ain, Mat& U, vector<double>& d, Mat& V) {
    const int n = Ain.n;
    Mat A = Ain;
    vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    vector<double> tau(n, 0.0), cn(n);
    for (int j = 0; j < n; ++j) { double s = 0.0; for (int i = 0; i < n; ++i) s += A(i, j) * A(i, j); cn[j] = s; }
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int j = k + 1; j < n; ++j) if (cn[j] > cn[p]) p = j;
        if (p != k) { for (int i = 0; i < n; ++i) std::swap(A(i, k), A(i, p)); std::swap(cn[k], cn[p]); std::swap(perm[k], perm[p]); }
        double norm2 = 0.0;
        for (int i = k; i < n; ++i) norm2 += A(i, k) * A(i, k);
        const double nrm = std::sqrt(norm2);
        if (nrm == 0.0) { tau[k] = 0.0; continue; }
        const double alpha = (A(k, k) > 0) ? -nrm : nrm;
        const double v0 = A(k, k) - alpha;
        // v = [1, A(k+1:,k)/v0], tau = -v0/alpha  (LAPACK dlarfg convention)
        for (int i = k + 1; i < n; ++i) A(i, k) /= v0;
        tau[k] = -v0 / alpha;
        A(k, k) = alpha;
        for (int j = k + 1; j < n; ++j) {
            double s = A(k, j);
            for (int i = k + 1; i < n; ++i) s += A(i, k) * A(i, j);
            s *= tau[k];
            A(k, j) -= s;
            for (int i = k + 1; i < n; ++i) A(i, j) -= s * A(i, k);
        }
        for (int j = k + 1; j < n; ++j) {  // exact recomputation of the trailing column norms
            double s = 0.0;
            for (int i = k + 1; i < n; ++i) s += A(i, j) * A(i, j);
            cn[j] = s;
        }
    }
    // explicit Q by backward accumulation
    U = Mat::Identity(n);
    for (int k = n - 1; k >= 0; --k) {
        if (tau[k] == 0.0) continue;
        for (int j = k; j < n; ++j) {
            double s = U(k, j);
            for (int i = k + 1; i < n; ++i) s += A(i, k) * U(i, j);
            s *= tau[k];
            U(k, j) -= s;
            for (int i = k + 1; i < n; ++i) U(i, j) -= s * A(i, k);
        }
    }
    d.assign(n, 0.0);
    for (int k = 0; k < n; ++k) d[k] = std::fabs(A(k, k));
    // T = D^-1 R P^T ; V = T^T
    V = Mat(n);
    for (int r = 0; r < n; ++r)
        for (int c = r; c < n; ++c) V(perm[c], r) = A(r, c) / d[r];
}

// Pre-pivoted Householder QR: columns sorted once by decreasing norm, then
// QR without pivoting (Bauer, SciPost Phys. Core 2, 011 (2020)): A P = Q R,
// U = Q, d = |diag R|, V^T = diag(d)^-1 R P^T.  The GPU stacks use this
// factorization with blocked Householder panels.
static void qr_prepivot_udv(const Mat& Ain, Mat& U, vector<double>& d, Mat& V) {
    const int n = Ain.n;
    vector<double> cn(n);
    for (int j = 0; j < n; ++j) { double s = 0.0; for (int i = 0; i < n; ++i) s += Ain(i, j) * Ain(i, j); cn[j] = s; }
    vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return cn[a] > cn[b]; });
    Mat A(n);
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < n; ++i) A(i, c) = Ain(i, perm[c]);
    vector<double> tau(n, 0.0);
    for (int k = 0; k < n; ++k) {
        double norm2 = 0.0;
        for (int i = k; i < n; ++i) norm2 += A(i, k) * A(i, k);
        const double nrm = std::sqrt(norm2);
        if (nrm == 0.0) continue;
        const double alpha = (A(k, k) > 0) ? -nrm : nrm;
        const double v0 = A(k, k) - alpha;
        for (int i = k + 1; i < n; ++i) A(i, k) /= v0;
        tau[k] = -v0 / alpha;
        A(k, k) = alpha;
        for (int j = k + 1; j < n; ++j) {
            double s = A(k, j);
            for (int i = k + 1; i < n; ++i) s += A(i, k) * A(i, j);
            s *= tau[k];
            A(k, j) -= s;
            for (int i = k + 1; i < n; ++i) A(i, j) -= s * A(i, k);
        }
    }
    U = Mat::Identity(n);
    for (int k = n - 1; k >= 0; --k) {
        if (tau[k] == 0.0) continue;
        for (int j = k; j < n; ++j) {
            double s = U(k, j);
            for (int i = k + 1; i < n; ++i) s += A(i, k) * U(i, j);
            s *= tau[k];
            U(k, j) -= s;
            for (int i = k + 1; i < n; ++i) U(i, j) -= s * A(i, k);
        }
    }
    d.assign(n, 0.0);
    for (int k = 0; k < n; ++k) d[k] = std::fabs(A(k, k));
    V = Mat(n);
    for (int r = 0; r < n; ++r)
        for (int c = r; c < n; ++c) V(perm[c], r) = A(r, c) / d[r];
}

// Matrix exponential, Higham 2005 scaling & squaring with Pade degree
// 3/5/7/9/13 selected by the 1-norm (the method of Eigen's MatrixExponential,
// used at src/model/repulsive_hubbard.cpp:64-65).
static Mat expm(const Mat& Ain) {
    const int n = Ain.n;
    double l1 = 0.0;
    for (int j = 0; j < n; ++j) { double s = 0.0; for (int i = 0; i < n; ++i) s += std::fabs(Ain(i, j)); l1 = std::max(l1, s); }
    const int degs[5] = {3, 5, 7, 9, 13};
    const double theta[5] = {1.495585217958292e-002, 2.539398330063230e-001, 9.504178996162932e-001,
                             2.097847961257068e+000, 5.371920351148152e+000};
    int m = 13, squarings = 0;
    for (int q = 0; q < 4; ++q) if (l1 < theta[q]) { m = degs[q]; break; }
    Mat A = Ain;
    if (m == 13 && l1 > theta[4]) {
        squarings = std::max(0, (int)std::ceil(std::log2(l1 / theta[4])));
        const double sc = std::ldexp(1.0, -squarings);
        for (auto& x : A.a) x *= sc;
    }
    // Pade coefficients c_j = (2m-j)! m! / ((2m)! j! (m-j)!)
    vector<double> c(m + 1);
    c[0] = 1.0;
    for (int j = 1; j <= m; ++j) c[j] = c[j - 1] * (double)(m - j + 1) / ((double)j * (double)(2 * m - j + 1));
    vector<Mat> pw(m + 1);  // even powers as needed
    pw[0] = Mat::Identity(n);
    Mat A2 = mul(A, A);
    Mat Upoly(n), Vpoly(n);
    // U = A * sum_{odd j} c_j A^{j-1},  V = sum_{even j} c_j A^j
    Mat P = Mat::Identity(n);  // A^{2i}
    for (int j = 0; j <= m; j += 2) {
        for (size_t e = 0; e < P.a.size(); ++e) Vpoly.a[e] += c[j] * P.a[e];
        if (j + 1 <= m) for (size_t e = 0; e < P.a.size(); ++e) Upoly.a[e] += c[j + 1] * P.a[e];
        if (j + 2 <= m) P = mul(P, A2);
    }
    Upoly = mul(A, Upoly);
    Mat num(n), den(n);
    for (size_t e = 0; e < num.a.size(); ++e) { num.a[e] = Vpoly.a[e] + Upoly.a[e]; den.a[e] = Vpoly.a[e] - Upoly.a[e]; }
    Mat LU = den; vector<int> piv; int ns;
    lu_factor(LU, piv, ns);
    lu_solve(LU, piv, num);
    for (int s = 0; s < squarings; ++s) num = mul(num, num);
    return num;
}

// ----------------------------------------------------------------------------
// Utils::SvdStack (src/svd_stack.cpp:15-87)
// ----------------------------------------------------------------------------
enum Factorization { kSVD = 0, kQRP = 1, kQRPre = 2 };

struct SvdStack {
    int dim = 0, len = 0, fact = kSVD;
    vector<Mat> U, V;
    vector<vector<double>> S;
    SvdStack() = default;
    SvdStack(int d, int cap, int f) : dim(d), len(0), fact(f), U(cap), V(cap), S(cap) {}
    bool empty() const { return len == 0; }
    void clear() { len = 0; }
    void decompose(const Mat& M, int slot) {
        if (fact == kSVD) svd(M, U[slot], S[slot], V[slot]);
        else if (fact == kQRP) qrp_udv(M, U[slot], S[slot], V[slot]);
        else qr_prepivot_udv(M, U[slot], S[slot], V[slot]);
    }
    void push(const Mat& m) {                                                  // :34-61
        if (len == 0) decompose(m, 0);
        else {
            Mat t = mul(m, U[len - 1]);
            const vector<double>& s = S[len - 1];
            for (int j = 0; j < dim; ++j) for (int i = 0; i < dim; ++i) t(i, j) *= s[j];
            decompose(t, len);
        }
        ++len;
    }
    void pop() { --len; }                                                      // :63-67
    vector<double> SingularValues() const { return S[len - 1]; }               // :70-73
    Mat MatrixU() const { return U[len - 1]; }                                 // :75-78
    Mat MatrixV() const {                                                      // :80-87
        Mat r = V[0];
        for (int i = 1; i < len; ++i) r = mul(r, V[i]);
        return r;
    }
};

// Sign of det G(0) = 1 / det(1 + M) from the stabilized factorization of one stack holding the
// whole product M = U diag(S) V^T (the right stack at initialization, the left one empty):
//   det(1 + M) = det(U) det(Db)^-1 det(Db U^T + Ds V^T),  Db = 1 / max(S, 1), Ds = min(S, 1),
// so sign det G = sign det U x sign det(Db U^T + Ds V^T), signs of well-conditioned matrices.
// The reference takes the sign of det G itself (initial_config_sign, src/dqmc_walker.cpp:162),
// which is rounding noise once G has eigenvalues below FP64 resolution (low temperature, large
// U: at beta 10, U 6, 16x16 half filling it comes out negative for some configurations, where
// particle-hole symmetry fixes it to +1).  This is the engine's definition; the oracle computes
// it on request (stable_sign) and is pinned to the reference's where that one is reliable.
static double stable_green_sign(const SvdStack& st) {
    const Mat U = st.MatrixU();
    const vector<double> S = st.SingularValues();
    const Mat V = st.MatrixV();
    const int n = U.n;
    Mat H(n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const double db = S[i] > 1.0 ? 1.0 / S[i] : 1.0, ds = S[i] > 1.0 ? 1.0 : S[i];
            H(i, j) = db * U(j, i) + ds * V(j, i);
        }
    return det_sign(U) * det_sign(H);
}

// ----------------------------------------------------------------------------
// Utils::NumericalStable (include/utils/numerical_stable.hpp)
// ----------------------------------------------------------------------------
static double matrix_compare_error(const Mat& u, const Mat& v) {               // :34-47
    double e = 0.0;
    for (int i = 0; i < u.n; ++i) for (int j = 0; j < u.n; ++j) e = std::max(e, std::fabs(u(i, j) - v(i, j)));
    return e;
}
static void div_dvec_max_min(const vector<double>& d, vector<double>& dmax, vector<double>& dmin) {  // :57-71
    const int n = (int)d.size();
    dmax.resize(n); dmin.resize(n);
    for (int i = 0; i < n; ++i) {
        if (d[i] >= 1.0) { dmin[i] = 1.0; dmax[i] = d[i]; }
        if (d[i] < 1.0) { dmax[i] = 1.0; dmin[i] = d[i]; }
    }
}
static void mult_v_invd_u(const Mat& v, const vector<double>& d, const Mat& u, Mat& z) {  // :79-98
    const int n = v.n;
    z = Mat(n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double zt = 0.0;
            for (int k = 0; k < n; ++k) zt += v(j, k) * u(k, i) / d[k];
            z(j, i) = zt;
        }
}
// (1 + U S V^T)^-1 with H = Sbi U^T + Ss V^T, G = H^-1 Sbi U^T  (:132-152)
static Mat compute_greens_00_bb(const Mat& U, const vector<double>& S, const Mat& V) {
    const int n = U.n;
    vector<double> Sbi(n), Ss(n);
    for (int i = 0; i < n; ++i) {
        if (S[i] > 1) { Sbi[i] = 1.0 / S[i]; Ss[i] = 1.0; }
        else { Sbi[i] = 1.0; Ss[i] = S[i]; }
    }
    Mat H(n), R(n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            H(i, j) = Sbi[i] * U(j, i) + Ss[i] * V(j, i);
            R(i, j) = Sbi[i] * U(j, i);
        }
    return fullpiv_qr_solve(H, R);
}
static Mat compute_equaltime_greens(const SvdStack& left, const SvdStack& right) {  // :186-239
    const int n = left.dim;
    if (left.empty()) {
        if (right.fact == kSVD) return compute_greens_00_bb(right.MatrixV(), right.SingularValues(), right.MatrixU());
        // non-orthogonal V: (1 + V S U^T)^-1 = ((1 + U S V^T)^-1)^T
        return transpose(compute_greens_00_bb(right.MatrixU(), right.SingularValues(), right.MatrixV()));
    }
    if (right.empty()) return compute_greens_00_bb(left.MatrixU(), left.SingularValues(), left.MatrixV());
    const Mat ul = left.MatrixU(); const vector<double> dl = left.SingularValues(); const Mat vl = left.MatrixV();
    const Mat ur = right.MatrixU(); const vector<double> dr = right.SingularValues(); const Mat vr = right.MatrixV();
    vector<double> dlmax, dlmin, drmax, drmin;
    div_dvec_max_min(dl, dlmax, dlmin);
    div_dvec_max_min(dr, drmax, drmin);
    Mat A = mul(transpose(ul), ur);
    Mat B = mul(transpose(vl), vr);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            A(i, j) = A(i, j) / (dlmax[i] * drmax[j]);
            B(i, j) = B(i, j) * dlmin[i] * drmin[j];
        }
    Mat tmp(n);
    for (size_t e = 0; e < tmp.a.size(); ++e) tmp.a[e] = A.a[e] + B.a[e];
    Mat Z1, G;
    mult_v_invd_u(ur, drmax, inverse(tmp), Z1);
    mult_v_invd_u(Z1, dlmax, transpose(ul), G);
    return G;
}
// v d u (:100-113)
static void mult_v_d_u(const Mat& v, const vector<double>& d, const Mat& u, Mat& z) {
    const int n = v.n;
    z = Mat(n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double zt = 0.0;
            for (int k = 0; k < n; ++k) zt += v(j, k) * u(k, i) * d[k];
            z(j, i) = zt;
        }
}
// (1 + U S V^T)^-1 U S V^T with H = Sbi U^T + Ss V^T, G = H^-1 Ss V^T  (:159-179)
static Mat compute_greens_b0(const Mat& U, const vector<double>& S, const Mat& V) {
    const int n = U.n;
    vector<double> Sbi(n), Ss(n);
    for (int i = 0; i < n; ++i) {
        if (S[i] > 1) { Sbi[i] = 1.0 / S[i]; Ss[i] = 1.0; }
        else { Sbi[i] = 1.0; Ss[i] = S[i]; }
    }
    Mat H(n), R(n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            H(i, j) = Sbi[i] * U(j, i) + Ss[i] * V(j, i);
            R(i, j) = Ss[i] * V(j, i);
        }
    return fullpiv_qr_solve(H, R);
}
// The reference's time-displaced G(t,0), G(0,t) from the two stacks (:246-325).  Its
// middle-case g0t inverts V_L and V_R by transposition, so it holds for SVD stacks.
static void compute_dynamic_greens(const SvdStack& left, const SvdStack& right, Mat& gt0, Mat& g0t) {
    const int n = left.dim;
    if (left.empty()) {
        gt0 = compute_greens_00_bb(right.MatrixV(), right.SingularValues(), right.MatrixU());
        g0t = gt0;
        for (int i = 0; i < n; ++i) g0t(i, i) -= 1.0;
        return;
    }
    if (right.empty()) {
        gt0 = compute_greens_b0(left.MatrixU(), left.SingularValues(), left.MatrixV());
        g0t = compute_greens_00_bb(left.MatrixU(), left.SingularValues(), left.MatrixV());
        for (double& x : g0t.a) x = -x;
        return;
    }
    const Mat ul = left.MatrixU(); const vector<double> dl = left.SingularValues(); const Mat vl = left.MatrixV();
    const Mat ur = right.MatrixU(); const vector<double> dr = right.SingularValues(); const Mat vr = right.MatrixV();
    vector<double> dlmax, dlmin, drmax, drmin;
    div_dvec_max_min(dl, dlmax, dlmin);
    div_dvec_max_min(dr, drmax, drmin);
    Mat A = mul(transpose(ul), ur), B = mul(transpose(vl), vr);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            A(i, j) = A(i, j) / (dlmax[i] * drmax[j]);
            B(i, j) = B(i, j) * dlmin[i] * drmin[j];
        }
    Mat tmp(n);
    for (size_t e = 0; e < tmp.a.size(); ++e) tmp.a[e] = A.a[e] + B.a[e];
    Mat Z;
    mult_v_invd_u(ur, drmax, inverse(tmp), Z);
    mult_v_d_u(Z, dlmin, transpose(vl), gt0);
    Mat X = mul(transpose(vr), vl), Y = mul(transpose(ur), ul);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            X(i, j) = X(i, j) / (drmax[i] * dlmax[j]);
            Y(i, j) = Y(i, j) * drmin[i] * dlmin[j];
        }
    for (size_t e = 0; e < tmp.a.size(); ++e) tmp.a[e] = X.a[e] + Y.a[e];
    Mat mvl = vl;
    for (double& x : mvl.a) x = -x;
    mult_v_invd_u(mvl, dlmax, inverse(tmp), Z);
    mult_v_d_u(Z, drmin, transpose(ur), g0t);
}
// The same three Green's functions for stacks whose V is not orthogonal (the QR-based
// UDV of the engine), from orthogonality of U alone.  With T = V^T:
//   1 + U_L D_L T_L T_R^T D_R U_R^T = U_L D_Lmax (A + B) D_Rmax U_R^T,
//   A = D_Lmax^-1 U_L^T U_R D_Rmax^-1,  B = D_Lmin T_L T_R^T D_Rmin,
// and by G(0,t) = -B(beta,t) G(t,t) = -[(1 + B(t,0)^T B(beta,t)^T)^-1 B(beta,t)^T]^T the
// matrix of G(0,t) is (A + B)^T, so one solve (A + B) [X1 X2] = [D_Lmax^-1 U_L^T, D_Lmin T_L]
// gives G(t,t) = U_R D_Rmax^-1 X1, G(t,0) = U_R D_Rmax^-1 X2, G(0,t) = -T_R^T D_Rmin X1.
// At t = beta (right stack empty): H = Sbi U^T + Ss V^T, H [X1 X2] = [Sbi U^T, Ss V^T],
// G(t,t) = X1, G(t,0) = X2, G(0,t) = -X1.
static void compute_dynamic_greens_udt(const SvdStack& left, const SvdStack& right, Mat& gtt, Mat& gt0, Mat& g0t) {
    const int n = left.dim;
    Mat H(n), R1(n), R2(n);
    if (right.empty()) {
        const Mat U = left.MatrixU(), V = left.MatrixV();
        const vector<double> S = left.SingularValues();
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double sbi = S[i] > 1 ? 1.0 / S[i] : 1.0, ss = S[i] > 1 ? 1.0 : S[i];
                H(i, j) = sbi * U(j, i) + ss * V(j, i);
                R1(i, j) = sbi * U(j, i);
                R2(i, j) = ss * V(j, i);
            }
        gtt = fullpiv_qr_solve(H, R1);
        gt0 = fullpiv_qr_solve(H, R2);
        g0t = gtt;
        for (double& x : g0t.a) x = -x;
        return;
    }
    const Mat ul = left.MatrixU(); const vector<double> dl = left.SingularValues(); const Mat vl = left.MatrixV();
    const Mat ur = right.MatrixU(); const vector<double> dr = right.SingularValues(); const Mat vr = right.MatrixV();
    vector<double> dlmax, dlmin, drmax, drmin;
    div_dvec_max_min(dl, dlmax, dlmin);
    div_dvec_max_min(dr, drmax, drmin);
    const Mat A = mul(transpose(ul), ur), B = mul(transpose(vl), vr);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            H(i, j) = A(i, j) / (dlmax[i] * drmax[j]) + B(i, j) * dlmin[i] * drmin[j];
            R1(i, j) = ul(j, i) / dlmax[i];
            R2(i, j) = vl(j, i) * dlmin[i];
        }
    Mat LU = H; vector<int> piv; int ns;
    lu_factor(LU, piv, ns);
    lu_solve(LU, piv, R1);
    lu_solve(LU, piv, R2);
    gtt = Mat(n); gt0 = Mat(n); g0t = Mat(n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double a = 0.0, b = 0.0, c = 0.0;
            for (int k = 0; k < n; ++k) {
                a += ur(i, k) / drmax[k] * R1(k, j);
                b += ur(i, k) / drmax[k] * R2(k, j);
                c += vr(i, k) * drmin[k] * R1(k, j);
            }
            gtt(i, j) = a; gt0(i, j) = b; g0t(i, j) = -c;
        }
}

// ----------------------------------------------------------------------------
// Model (src/model/{model_base,repulsive_hubbard,attractive_hubbard}.cpp) and
// CheckerBoard::Square (src/checkerboard/square.cpp).
// ----------------------------------------------------------------------------
enum ModelKind { kRepulsive = 0, kAttractive = 1 };

struct Engine;  // fwd

struct Model {
    int kind = kRepulsive;
    double t = 1.0, U = 4.0, mu = 0.0;
    int N = 0, L = 0;
    double dtau = 0.0, alpha = 0.0;
    Mat expK, invExpK, transExpK;
    vector<double> field;  // (t,i) at t*N+i
    bool cb = false;
    int side = 0;
    double Pk[16], Pinv[16];  // 4x4 plaquettes, row-major

    void init(const Square& lat, double beta, int time_size, bool checkerboard) {
        N = lat.N; L = time_size; side = lat.L;
        dtau = beta / time_size;                                   // src/dqmc_walker.cpp:30
        alpha = std::acosh(std::exp(0.5 * dtau * U));              // :50
        field.assign((size_t)L * N, 1.0);
        Mat K(N);
        for (int j = 0; j < N; ++j) for (int i = 0; i < N; ++i) K(i, j) = -t * lat.hopping(i, j) + (i == j ? mu : 0.0);
        Mat A1(N), A2(N);
        for (size_t e = 0; e < K.a.size(); ++e) { A1.a[e] = -dtau * K.a[e]; A2.a[e] = +dtau * K.a[e]; }
        expK = expm(A1); invExpK = expm(A2); transExpK = transpose(expK);  // :64-68
        cb = checkerboard;
        // src/checkerboard/square.cpp:28-50
        double ch = std::cosh(dtau * t), sh = std::sinh(dtau * t);
        double red[16] = {ch * ch, ch * sh, ch * sh, sh * sh, ch * sh, ch * ch, sh * sh, ch * sh,
                          ch * sh, sh * sh, ch * ch, ch * sh, sh * sh, ch * sh, ch * sh, ch * ch};
        double f = std::exp(0.5 * dtau * mu);
        for (int e = 0; e < 16; ++e) Pk[e] = f * red[e];
        ch = std::cosh(-dtau * t); sh = std::sinh(-dtau * t);
        double red2[16] = {ch * ch, ch * sh, ch * sh, sh * sh, ch * sh, ch * ch, sh * sh, ch * sh,
                           ch * sh, sh * sh, ch * ch, ch * sh, sh * sh, ch * sh, ch * sh, ch * ch};
        f = std::exp(-0.5 * dtau * mu);
        for (int e = 0; e < 16; ++e) Pinv[e] = f * red2[e];
    }
    double s(int tt, int i) const { return field[(size_t)tt * N + i]; }

    void set_fields_random(Minstd0& g) {                           // :86-99
        for (int tt = 0; tt < L; ++tt)
            for (int i = 0; i < N; ++i) field[(size_t)tt * N + i] = bernoulli(g, 0.5) ? +1.0 : -1.0;
    }

    // --- checkerboard plaquette products (src/checkerboard/square.cpp:66-127)
    void plaq_idx(int x, int y, int idx[4]) const {
        idx[0] = (x % side) + side * (y % side);
        idx[1] = ((x + 1) % side) + side * (y % side);
        idx[2] = (x % side) + side * ((y + 1) % side);
        idx[3] = ((x + 1) % side) + side * ((y + 1) % side);
    }
    void plaq_left(Mat& m, const double* P, int x, int y) const {
        int idx[4]; plaq_idx(x, y, idx);
        for (int j = 0; j < N; ++j) {
            double v[4] = {m(idx[0], j), m(idx[1], j), m(idx[2], j), m(idx[3], j)};
            for (int r = 0; r < 4; ++r) m(idx[r], j) = P[r * 4 + 0] * v[0] + P[r * 4 + 1] * v[1] + P[r * 4 + 2] * v[2] + P[r * 4 + 3] * v[3];
        }
    }
    void plaq_right(Mat& m, const double* P, int x, int y) const {
        int idx[4]; plaq_idx(x, y, idx);
        for (int i = 0; i < N; ++i) {
            double v[4] = {m(i, idx[0]), m(i, idx[1]), m(i, idx[2]), m(i, idx[3])};
            for (int c = 0; c < 4; ++c) m(i, idx[c]) = v[0] * P[0 * 4 + c] + v[1] * P[1 * 4 + c] + v[2] * P[2 * 4 + c] + v[3] * P[3 * 4 + c];
        }
    }
    void sub_left(Mat& m, const double* P, int start) const {
        for (int x = start; x < side; x += 2) for (int y = start; y < side; y += 2) plaq_left(m, P, x, y);
    }
    void sub_right(Mat& m, const double* P, int start) const {
        for (int x = start; x < side; x += 2) for (int y = start; y < side; y += 2) plaq_right(m, P, x, y);
    }
    // the five hooks of ModelBase::link / link(checkerboard) (src/model/model_base.cpp:7-55)
    void mult_expK_left(Mat& g) const {
        if (!cb) { g = mul(expK, g); return; }
        sub_left(g, Pk, 1); sub_left(g, Pk, 0);                     // :130-147  B then A
    }
    void mult_expK_right(Mat& g) const {
        if (!cb) { g = mul(g, expK); return; }
        sub_right(g, Pk, 0); sub_right(g, Pk, 1);                   // :170-187  A then B
    }
    void mult_inv_expK_left(Mat& g) const {
        if (!cb) { g = mul(invExpK, g); return; }
        sub_left(g, Pinv, 0); sub_left(g, Pinv, 1);                 // :150-167  A then B
    }
    void mult_inv_expK_right(Mat& g) const {
        if (!cb) { g = mul(g, invExpK); return; }
        sub_right(g, Pinv, 1); sub_right(g, Pinv, 0);               // :190-207  B then A
    }
    void mult_trans_expK_left(Mat& g) const {
        if (!cb) { g = mul(transExpK, g); return; }
        sub_left(g, Pk, 0); sub_left(g, Pk, 1);                     // :209-226  A then B
    }

    double vsign(int spin) const { return kind == kAttractive ? 1.0 : (double)spin; }
    int eff(int tt) const { return tt == 0 ? L - 1 : tt - 1; }     // :170

    void mult_B_from_left(Mat& g, int tt, int spin) const {        // :158-175
        const int e = eff(tt);
        mult_expK_left(g);
        for (int i = 0; i < N; ++i) { const double f = std::exp(+vsign(spin) * alpha * s(e, i)); for (int j = 0; j < N; ++j) g(i, j) *= f; }
    }
    void mult_B_from_right(Mat& g, int tt, int spin) const {       // :178-192
        const int e = eff(tt);
        for (int i = 0; i < N; ++i) { const double f = std::exp(+vsign(spin) * alpha * s(e, i)); double* c = g.col(i); for (int j = 0; j < N; ++j) c[j] *= f; }
        mult_expK_right(g);
    }
    void mult_invB_from_left(Mat& g, int tt, int spin) const {     // :195-209
        const int e = eff(tt);
        for (int i = 0; i < N; ++i) { const double f = std::exp(-vsign(spin) * alpha * s(e, i)); for (int j = 0; j < N; ++j) g(i, j) *= f; }
        mult_inv_expK_left(g);
    }
    void mult_invB_from_right(Mat& g, int tt, int spin) const {    // :212-226
        const int e = eff(tt);
        mult_inv_expK_right(g);
        for (int i = 0; i < N; ++i) { const double f = std::exp(-vsign(spin) * alpha * s(e, i)); double* c = g.col(i); for (int j = 0; j < N; ++j) c[j] *= f; }
    }
    void mult_transB_from_left(Mat& g, int tt, int spin) const {   // :229-243
        const int e = eff(tt);
        for (int i = 0; i < N; ++i) { const double f = std::exp(+vsign(spin) * alpha * s(e, i)); for (int j = 0; j < N; ++j) g(i, j) *= f; }
        mult_trans_expK_left(g);
    }

    // get_update_ratio (repulsive :112-124, attractive :112-125)
    double ratio(const Mat& gu, const Mat& gd, int tt, int i) const {
        const double si = s(tt, i);
        if (kind == kRepulsive)
            return (1 + (1 - gu(i, i)) * (std::exp(-2 * alpha * si) - 1)) * (1 + (1 - gd(i, i)) * (std::exp(+2 * alpha * si) - 1));
        return std::exp(2 * alpha * si) * (1 + (1 - gu(i, i)) * (std::exp(-2 * alpha * si) - 1)) *
               (1 + (1 - gd(i, i)) * (std::exp(-2 * alpha * si) - 1));
    }
    // update_greens_function (repulsive :127-155, attractive :128-157)
    void update_greens(Mat& gu, Mat& gd, int tt, int i) const {
        const double si = s(tt, i);
        const double fu = (std::exp(-2 * alpha * si) - 1) / (1 + (1 - gu(i, i)) * (std::exp(-2 * alpha * si) - 1));
        double fd;
        if (kind == kRepulsive) fd = (std::exp(+2 * alpha * si) - 1) / (1 + (1 - gd(i, i)) * (std::exp(+2 * alpha * si) - 1));
        else fd = fu;
        rank1(gu, fu, i);
        rank1(gd, fd, i);
    }
    // G -= f * G(:,i) * (e_i^T - G(i,:)), outer product evaluated into a temporary
    void rank1(Mat& g, double f, int i) const {
        vector<double> c(N), r(N);
        for (int a = 0; a < N; ++a) { c[a] = g(a, i); r[a] = (a == i ? 1.0 : 0.0) - g(i, a); }
        for (int b = 0; b < N; ++b) {
            const double rb = r[b];
            for (int a = 0; a < N; ++a) g(a, b) -= f * c[a] * rb;
        }
    }
    void flip(int tt, int i) { field[(size_t)tt * N + i] = -field[(size_t)tt * N + i]; }  // :102-109
};

// ----------------------------------------------------------------------------
// Equal-time observables (src/measure/measure_methods.cpp:22-255) — index order
// of the accumulator arrays used everywhere in this project.
// ----------------------------------------------------------------------------
enum Obs {
    kSign = 0, kFilling, kDoubleOcc, kKinetic, kMomentumDist, kLocalSpin, kSDW, kCDW, kSwave, kNumObs
};

// ----------------------------------------------------------------------------
// QuantumMonteCarlo::DqmcWalker (src/dqmc_walker.cpp)
// ----------------------------------------------------------------------------
struct Engine {
    Square lat;
    Model model;
    int L = 0, N = 0, pace = 10, fact = kSVD, momentum = 0;
    double beta = 0.0;
    Minstd0 rng;
    Mat gu, gd;
    vector<Mat> vgu, vgd;      // vec_green_tt
    vector<double> vsign;      // vec_config_sign
    double sign = 1.0, wrap_error = 0.0;
    bool use_stable_sign = false;   // initial sign from the stabilized factorization (stable_green_sign)
    double sign_reference = 1.0, sign_stable = 1.0;   // both definitions at the last init
    SvdStack lu_, ld_, ru_, rd_;
    vector<unsigned char> accepts;  // last sweep, (eff_t, i) at eff_t*N+i
    double acc[kNumObs] = {0}, cnt[kNumObs] = {0};
    // time-displaced path: G(t,0), G(0,t) and their per-slice records (vec_green_t0/0t)
    Mat gt0u, gt0d, g0tu, g0td;
    vector<Mat> vgt0u, vgt0d, vg0tu, vg0td;
    bool dyn_general = false;   // use the orthogonal-U form of the recompute even with SVD stacks
    // dynamic accumulators: G(k,t) over every k-star (nk x L, (k,t) at k + nk t), D(t), DSS(t)
    // per slice; rho_s; sign
    vector<double> dgk, ddos, ddss;
    double drho = 0.0, dsign = 0.0, dcount = 0.0;

    void setup(int side, int time_size, double beta_, int pace_, int kind, bool cb, double t, double U, double mu,
               int momentum_index, int factorization) {
        lat.init(side);
        N = lat.N; L = time_size; beta = beta_; pace = pace_; fact = factorization;
        model.kind = kind; model.t = t; model.U = U; model.mu = mu;
        model.init(lat, beta, time_size, cb);
        momentum = momentum_index < 0 ? lat.mpoint : momentum_index;
        vgu.assign(L, Mat(N)); vgd.assign(L, Mat(N)); vsign.assign(L, 0.0);
        accepts.assign((size_t)L * N, 0);
        vgt0u.assign(L, Mat(N)); vgt0d.assign(L, Mat(N)); vg0tu.assign(L, Mat(N)); vg0td.assign(L, Mat(N));
        dgk.assign((size_t)lat.nk * L, 0.0); ddos.assign(L, 0.0); ddss.assign(L, 0.0);
    }

    void initial_svd_stacks() {                                    // :109-134
        lu_ = SvdStack(N, L, fact); ld_ = SvdStack(N, L, fact);
        ru_ = SvdStack(N, L, fact); rd_ = SvdStack(N, L, fact);
        Mat tu = Mat::Identity(N), td = Mat::Identity(N);
        for (int t = L; t >= 1; --t) {
            model.mult_transB_from_left(tu, t, +1);
            model.mult_transB_from_left(td, t, -1);
            if ((t - 1) % pace == 0) {
                ru_.push(tu); rd_.push(td);
                tu = Mat::Identity(N); td = Mat::Identity(N);
            }
        }
    }
    void initial_greens() {                                        // :137-147
        gu = compute_equaltime_greens(lu_, ru_);
        gd = compute_equaltime_greens(ld_, rd_);
    }
    void initial_sign() {                                          // :150-163
        sign_reference = (det_sign_times(gu) * det_sign_times(gd) >= 0) ? +1.0 : -1.0;
        sign_stable = stable_green_sign(ru_) * stable_green_sign(rd_);
        sign = use_stable_sign ? sign_stable : sign_reference;
    }
    void init() { initial_svd_stacks(); initial_greens(); initial_sign(); wrap_error = 0.0; }

    void metropolis(int t) {                                       // :173-197
        const int e = (t == 0) ? L - 1 : t - 1;
        for (int i = 0; i < N; ++i) {
            const double r = model.ratio(gu, gd, e, i);
            const bool ok = bernoulli(rng, std::min(1.0, std::fabs(r)));
            accepts[(size_t)e * N + i] = ok ? 1 : 0;
            if (ok) {
                model.update_greens(gu, gd, e, i);
                model.flip(e, i);
                sign = (r >= 0) ? +sign : -sign;
            }
        }
    }
    void wrap_0_to_beta(int t) {                                   // :208-217
        const int e = (t == L) ? 1 : t + 1;
        model.mult_B_from_left(gu, e, +1); model.mult_invB_from_right(gu, e, +1);
        model.mult_B_from_left(gd, e, -1); model.mult_invB_from_right(gd, e, -1);
    }
    void wrap_beta_to_0(int t) {                                   // :228-237
        const int e = (t == 0) ? L : t;
        model.mult_B_from_right(gu, e, +1); model.mult_invB_from_left(gu, e, +1);
        model.mult_B_from_right(gd, e, -1); model.mult_invB_from_left(gd, e, -1);
    }
    void stabilize_swap(SvdStack& popu, SvdStack& popd, SvdStack& pushu, SvdStack& pushd, const Mat& tu, const Mat& td) {
        popu.pop(); popd.pop();
        pushu.push(tu); pushd.push(td);
    }
    void fresh_with_error() {
        Mat fu = compute_equaltime_greens(lu_, ru_);
        Mat fd = compute_equaltime_greens(ld_, rd_);
        const double eu = matrix_compare_error(fu, gu), ed = matrix_compare_error(fd, gd);
        wrap_error = std::max(wrap_error, std::max(eu, ed));
        gu = fu; gd = fd;
    }
    void sweep_0_to_beta() {                                       // :246-325
        Mat tu = Mat::Identity(N), td = Mat::Identity(N);
        for (int t = 1; t <= L; ++t) {
            wrap_0_to_beta(t - 1);
            metropolis(t);
            vgu[t - 1] = gu; vgd[t - 1] = gd; vsign[t - 1] = sign;
            model.mult_B_from_left(tu, t, +1);
            model.mult_B_from_left(td, t, -1);
            if (t % pace == 0 || t == L) {
                stabilize_swap(ru_, rd_, lu_, ld_, tu, td);
                fresh_with_error();
                vgu[t - 1] = gu; vgd[t - 1] = gd;
                tu = Mat::Identity(N); td = Mat::Identity(N);
            }
        }
        vgu[L - 1] = gu; vgd[L - 1] = gd;
    }
    void sweep_beta_to_0() {                                       // :334-412
        Mat tu = Mat::Identity(N), td = Mat::Identity(N);
        for (int t = L; t >= 1; --t) {
            if (t % pace == 0 && t != L) {
                stabilize_swap(lu_, ld_, ru_, rd_, tu, td);
                fresh_with_error();
                tu = Mat::Identity(N); td = Mat::Identity(N);
            }
            metropolis(t);
            vgu[t - 1] = gu; vgd[t - 1] = gd; vsign[t - 1] = sign;
            model.mult_transB_from_left(tu, t, +1);
            model.mult_transB_from_left(td, t, -1);
            wrap_beta_to_0(t);
        }
        lu_.pop(); ld_.pop();
        ru_.push(tu); rd_.push(td);
        gu = compute_equaltime_greens(lu_, ru_);
        gd = compute_equaltime_greens(ld_, rd_);
        vgu[L - 1] = gu; vgd[L - 1] = gd;
    }

    // DqmcWalker::sweep_for_dynamic_greens (src/dqmc_walker.cpp:424-524): from t = 0 (left stack
    // empty, right stack full, as a backward sweep leaves it) to beta without field updates,
    // wrapping G(t,t), G(t,0) = B(t) G(t-1,0) and G(0,t) = G(0,t-1) B(t)^-1, recomputed from the
    // stacks at every stabilization (the wrap errors of G(t,0) and G(0,t) are collected).
    void fresh_dynamic(const SvdStack& l, const SvdStack& r, Mat& gtt, Mat& gt0, Mat& g0t) {
        if (fact == kSVD && !dyn_general) {
            gtt = compute_equaltime_greens(l, r);
            compute_dynamic_greens(l, r, gt0, g0t);
        } else {
            compute_dynamic_greens_udt(l, r, gtt, gt0, g0t);
        }
    }
    void sweep_dynamic() {
        gt0u = gu; gt0d = gd;
        g0tu = gu; g0td = gd;
        for (int i = 0; i < N; ++i) { g0tu(i, i) -= 1.0; g0td(i, i) -= 1.0; }
        Mat tu = Mat::Identity(N), td = Mat::Identity(N);
        for (int t = 1; t <= L; ++t) {
            wrap_0_to_beta(t - 1);
            vgu[t - 1] = gu; vgd[t - 1] = gd; vsign[t - 1] = sign;
            model.mult_B_from_left(gt0u, t, +1); model.mult_B_from_left(gt0d, t, -1);
            vgt0u[t - 1] = gt0u; vgt0d[t - 1] = gt0d;
            model.mult_invB_from_right(g0tu, t, +1); model.mult_invB_from_right(g0td, t, -1);
            vg0tu[t - 1] = g0tu; vg0td[t - 1] = g0td;
            model.mult_B_from_left(tu, t, +1);
            model.mult_B_from_left(td, t, -1);
            if (t % pace == 0 || t == L) {
                stabilize_swap(ru_, rd_, lu_, ld_, tu, td);
                Mat fu, fd, au, ad, bu, bd;
                fresh_dynamic(lu_, ru_, fu, au, bu);
                fresh_dynamic(ld_, rd_, fd, ad, bd);
                wrap_error = std::max(wrap_error, std::max(matrix_compare_error(au, gt0u), matrix_compare_error(ad, gt0d)));
                wrap_error = std::max(wrap_error, std::max(matrix_compare_error(bu, g0tu), matrix_compare_error(bd, g0td)));
                gu = fu; gd = fd; gt0u = au; gt0d = ad; g0tu = bu; g0td = bd;
                vgu[t - 1] = gu; vgd[t - 1] = gd;
                vgt0u[t - 1] = gt0u; vgt0d[t - 1] = gt0d; vg0tu[t - 1] = g0tu; vg0td[t - 1] = g0td;
                tu = Mat::Identity(N); td = Mat::Identity(N);
            }
        }
    }
    // MeasureHandler::dynamic_measure over the stored slices (src/measure/measure_methods.cpp:264-461):
    // G(k,t) for the momentum list of every k-star (the reference's "KstarsAll"), density of
    // states D(t), superfluid stiffness rho_s, dynamic spin susceptibility, sign.  Slot t = 0
    // takes the slice-(L-1) records, as there.
    void dynamic_measure() {
        const double Nd = (double)N, cs = sign;
        const Mat& g00u = vgu[L - 1]; const Mat& g00d = vgd[L - 1];
        for (int t = 0; t < L; ++t) {
            const int e = (t == 0) ? L - 1 : t - 1;
            // greens_functions and density_of_states (:278-325): the spin average of G(t,0), G(0,0) at t = 0
            const Mat& au = (t == 0) ? vgu[L - 1] : vgt0u[t - 1];
            const Mat& ad = (t == 0) ? vgd[L - 1] : vgt0d[t - 1];
            double tr = 0.0;
            for (int i = 0; i < N; ++i) {
                tr += 0.5 * (au(i, i) + ad(i, i));
                for (int j = 0; j < N; ++j)
                    for (int k = 0; k < lat.nk; ++k)
                        dgk[(size_t)t * lat.nk + k] += cs * 0.5 * (au(j, i) + ad(j, i)) / Nd *
                                                       lat.FourierFactor(lat.disp[(size_t)i * N + j], k);
            }
            ddos[t] += cs * tr / Nd;
            // superfluid_stiffness (:335-405) and dynamic_spin_susceptibility (:413-461)
            const Mat& gttu = vgu[e]; const Mat& gttd = vgd[e];
            const Mat& gt0u_ = vgt0u[e]; const Mat& gt0d_ = vgt0d[e];
            const Mat& g0tu_ = vg0tu[e]; const Mat& g0td_ = vg0td[e];
            double rho = 0.0;
            for (int i = 0; i < N; ++i) {
                const int ipx = lat.nn[i * 4 + 0];
                for (int j = 0; j < N; ++j) {
                    const int jpx = lat.nn[j * 4 + 0];
                    const int dsp = lat.disp[(size_t)i * N + j];
                    const int rx = dsp % lat.L, ry = dsp / lat.L;
                    const double ff = lat.FourierFactor(rx, 1) - lat.FourierFactor(ry, 1);
                    rho += model.t * model.t * cs * ff * (
                        -(gttu(j, jpx) - gttu(jpx, j) + gttd(j, jpx) - gttd(jpx, j)) *
                         (g00u(i, ipx) - g00u(ipx, i) + g00d(i, ipx) - g00d(ipx, i))
                        - g0tu_(ipx, jpx) * gt0u_(j, i) - g0td_(ipx, jpx) * gt0d_(j, i)
                        + g0tu_(i, jpx) * gt0u_(j, ipx) + g0td_(i, jpx) * gt0d_(j, ipx)
                        + g0tu_(ipx, j) * gt0u_(jpx, i) + g0td_(ipx, j) * gt0d_(jpx, i)
                        - g0tu_(i, j) * gt0u_(jpx, ipx) - g0td_(i, j) * gt0d_(jpx, ipx));
                }
            }
            drho += 0.25 * rho / (Nd * Nd);
            double dss = 0.0;
            for (int i = 0; i < N; ++i) {
                const double gcttu = 1.0 - gttu(i, i), gcttd = 1.0 - gttd(i, i);
                const double gc00u = 1.0 - g00u(i, i), gc00d = 1.0 - g00d(i, i);
                dss += 0.25 * cs / Nd * (gcttu * gc00u - g0tu_(i, i) * gt0u_(i, i) + gcttd * gc00d - g0td_(i, i) * gt0d_(i, i)
                                         - gcttu * gc00d - gc00u * gcttd);
            }
            ddss[t] += dss;
        }
        dsign += cs;
        dcount += 1.0;
    }
    void clear_dynamic() {
        std::fill(dgk.begin(), dgk.end(), 0.0); std::fill(ddos.begin(), ddos.end(), 0.0);
        std::fill(ddss.begin(), ddss.end(), 0.0);
        drho = dsign = dcount = 0.0;
    }

    // MeasureHandler::equaltime_measure over the stored slices
    void equaltime_measure() {
        const double Nd = (double)N;
        const double tt = model.t;
        for (int t = 0; t < L; ++t) {
            const Mat& u = vgu[t]; const Mat& d = vgd[t];
            const double cs = vsign[t];
            // filling_number :42-46
            double tru = 0.0, trd = 0.0;
            for (int i = 0; i < N; ++i) { tru += u(i, i); trd += d(i, i); }
            acc[kFilling] += cs * (2 - (tru + trd) / Nd);
            // double_occupancy :58-69
            double docc = 0.0;
            for (int i = 0; i < N; ++i) docc += (1 - u(i, i)) * (1 - d(i, i));
            acc[kDoubleOcc] += cs * docc / Nd;
            // kinetic_energy :81-98
            double kin = 0.0;
            for (int i = 0; i < N; ++i)
                for (int dir = 0; dir < 2; ++dir) kin += u(i, lat.nn[i * 4 + dir]) + d(i, lat.nn[i * 4 + dir]);
            acc[kKinetic] += cs * (2 * tt) * kin / Nd;
            // local_spin_corr :111-122
            double ls = 0.0;
            for (int i = 0; i < N; ++i) ls += cs * (u(i, i) + d(i, i) - 2 * u(i, i) * d(i, i));
            acc[kLocalSpin] += ls / Nd;
            // momentum_distribution :135-151
            double md = 0.0;
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) md += (u(j, i) + d(j, i)) * lat.FourierFactor(lat.disp[(size_t)i * N + j], momentum);
            acc[kMomentumDist] += cs * (1 - 0.5 * md / Nd);
            // sdw / cdw :164-219, s-wave :235-254  (gc = I - g^T)
            double sdw = 0.0, cdw = 0.0, sw = 0.0;
            for (int i = 0; i < N; ++i) {
                const double guc_ii = 1.0 - u(i, i), gdc_ii = 1.0 - d(i, i);
                for (int j = 0; j < N; ++j) {
                    const double dij = (i == j) ? 1.0 : 0.0;
                    const double guc_jj = 1.0 - u(j, j), gdc_jj = 1.0 - d(j, j);
                    const double guc_ij = dij - u(j, i), gdc_ij = dij - d(j, i);
                    const double ff = lat.FourierFactor(lat.disp[(size_t)i * N + j], momentum);
                    sdw += cs * ff * (+guc_ii * guc_jj + guc_ij * u(i, j) + gdc_ii * gdc_jj + gdc_ij * d(i, j)
                                      - gdc_ii * guc_jj - guc_ii * gdc_jj);
                    cdw += cs * ff * (+guc_ii * guc_jj + guc_ij * u(i, j) + gdc_ii * gdc_jj + gdc_ij * d(i, j)
                                      + gdc_ii * guc_jj + guc_ii * gdc_jj);
                    sw += cs * (guc_ij * gdc_ij);
                }
            }
            acc[kSDW] += sdw / (Nd * Nd);
            acc[kCDW] += cdw / (Nd * Nd);
            acc[kSwave] += sw / Nd;
            for (int o = 1; o < kNumObs; ++o) cnt[o] += 1.0;
        }
        // equaltime_sign :22-30
        for (int t = 0; t < L; ++t) acc[kSign] += vsign[t];
        cnt[kSign] += L;
    }
    void clear_measure() { for (int o = 0; o < kNumObs; ++o) { acc[o] = 0.0; cnt[o] = 0.0; } }
};

}  // namespace refcpu

// ============================================================================
// C interface (ctypes) — test infrastructure only.
// ============================================================================
using namespace refcpu;

extern "C" {

void* rc_create(int side, int time_size, double beta, int pace, int kind, int cb, double t, double U, double mu,
                int momentum_index, int factorization) {
    Engine* e = new Engine();
    e->setup(side, time_size, beta, pace, kind, cb != 0, t, U, mu, momentum_index, factorization);
    return e;
}
void rc_destroy(void* h) { delete (Engine*)h; }
int rc_space_size(void* h) { return ((Engine*)h)->N; }
void rc_seed(void* h, long long s) { ((Engine*)h)->rng.seed(s); }
unsigned long long rc_rng_state(void* h) { return ((Engine*)h)->rng.x; }
void rc_set_rng_state(void* h, unsigned long long x) { ((Engine*)h)->rng.x = x; }
void rc_fields_random(void* h) { Engine* e = (Engine*)h; e->model.set_fields_random(e->rng); }
void rc_set_fields(void* h, const double* f) { Engine* e = (Engine*)h; std::memcpy(e->model.field.data(), f, sizeof(double) * e->model.field.size()); }
void rc_get_fields(void* h, double* f) { Engine* e = (Engine*)h; std::memcpy(f, e->model.field.data(), sizeof(double) * e->model.field.size()); }
void rc_init(void* h) { ((Engine*)h)->init(); }
void rc_sweep_forward(void* h) { ((Engine*)h)->sweep_0_to_beta(); }
void rc_sweep_backward(void* h) { ((Engine*)h)->sweep_beta_to_0(); }
void rc_get_green(void* h, int spin, double* out) {
    Engine* e = (Engine*)h; const Mat& g = spin > 0 ? e->gu : e->gd;
    std::memcpy(out, g.a.data(), sizeof(double) * g.a.size());
}
void rc_set_green(void* h, int spin, const double* in) {
    Engine* e = (Engine*)h; Mat& g = spin > 0 ? e->gu : e->gd;
    std::memcpy(g.a.data(), in, sizeof(double) * g.a.size());
}
void rc_get_slice_green(void* h, int slot, int spin, double* out) {
    Engine* e = (Engine*)h; const Mat& g = spin > 0 ? e->vgu[slot] : e->vgd[slot];
    std::memcpy(out, g.a.data(), sizeof(double) * g.a.size());
}
double rc_get_slice_sign(void* h, int slot) { return ((Engine*)h)->vsign[slot]; }
// a stored slot's G of both spins and its sign (the input of equaltime_measure)
void rc_set_slice(void* h, int slot, const double* gu, const double* gd, double sign) {
    Engine* e = (Engine*)h;
    std::memcpy(e->vgu[slot].a.data(), gu, sizeof(double) * e->vgu[slot].a.size());
    std::memcpy(e->vgd[slot].a.data(), gd, sizeof(double) * e->vgd[slot].a.size());
    e->vsign[slot] = sign;
}
double rc_sign(void* h) { return ((Engine*)h)->sign; }
void rc_set_stable_sign(void* h, int on) { ((Engine*)h)->use_stable_sign = on != 0; }
double rc_init_sign(void* h, int stable) { Engine* e = (Engine*)h; return stable ? e->sign_stable : e->sign_reference; }
double rc_wrap_error(void* h) { return ((Engine*)h)->wrap_error; }
void rc_get_accepts(void* h, unsigned char* out) { Engine* e = (Engine*)h; std::memcpy(out, e->accepts.data(), e->accepts.size()); }
void rc_measure(void* h) { ((Engine*)h)->equaltime_measure(); }
void rc_get_measure(void* h, double* sums, double* counts) {
    Engine* e = (Engine*)h;
    for (int o = 0; o < kNumObs; ++o) { sums[o] = e->acc[o]; counts[o] = e->cnt[o]; }
}
void rc_clear_measure(void* h) { ((Engine*)h)->clear_measure(); }
// time-displaced path
void rc_sweep_dynamic(void* h) { ((Engine*)h)->sweep_dynamic(); }
void rc_set_dynamic_formula(void* h, int general) { ((Engine*)h)->dyn_general = general != 0; }
// which: 0 G(t,0), 1 G(0,t)
void rc_get_dynamic_green(void* h, int slot, int spin, int which, double* out) {
    Engine* e = (Engine*)h;
    const Mat& g = which == 0 ? (spin > 0 ? e->vgt0u[slot] : e->vgt0d[slot]) : (spin > 0 ? e->vg0tu[slot] : e->vg0td[slot]);
    std::memcpy(out, g.a.data(), sizeof(double) * g.a.size());
}
void rc_dynamic_measure(void* h) { ((Engine*)h)->dynamic_measure(); }
// out: G(k,t)[nk x L, (k,t) at k + nk t], D(t)[L], DSS(t)[L], rho_s, sign sum, count
void rc_get_dynamic_measure(void* h, double* out) {
    Engine* e = (Engine*)h;
    const int L = e->L, nk = e->lat.nk;
    for (size_t q = 0; q < e->dgk.size(); ++q) out[q] = e->dgk[q];
    out += (size_t)nk * L;
    for (int t = 0; t < L; ++t) { out[t] = e->ddos[t]; out[L + t] = e->ddss[t]; }
    out[2 * L] = e->drho; out[2 * L + 1] = e->dsign; out[2 * L + 2] = e->dcount;
}
void rc_clear_dynamic(void* h) { ((Engine*)h)->clear_dynamic(); }
double rc_alpha(void* h) { return ((Engine*)h)->model.alpha; }
// which: 0 expK, 1 invExpK, 2 transExpK
void rc_get_expK(void* h, int which, double* out) {
    Engine* e = (Engine*)h; const Mat& m = which == 0 ? e->model.expK : (which == 1 ? e->model.invExpK : e->model.transExpK);
    std::memcpy(out, m.a.data(), sizeof(double) * m.a.size());
}
void rc_set_expK(void* h, const double* expK, const double* invExpK) {
    Engine* e = (Engine*)h; const size_t n = e->model.expK.a.size();
    std::memcpy(e->model.expK.a.data(), expK, sizeof(double) * n);
    std::memcpy(e->model.invExpK.a.data(), invExpK, sizeof(double) * n);
    e->model.transExpK = transpose(e->model.expK);
}
// op: 0 B_left, 1 B_right, 2 invB_left, 3 invB_right, 4 transB_left, 5..9 the bare expK hooks in the same order
void rc_mult(void* h, int op, int t, int spin, double* m) {
    Engine* e = (Engine*)h; Mat g(e->N);
    std::memcpy(g.a.data(), m, sizeof(double) * g.a.size());
    switch (op) {
        case 0: e->model.mult_B_from_left(g, t, spin); break;
        case 1: e->model.mult_B_from_right(g, t, spin); break;
        case 2: e->model.mult_invB_from_left(g, t, spin); break;
        case 3: e->model.mult_invB_from_right(g, t, spin); break;
        case 4: e->model.mult_transB_from_left(g, t, spin); break;
        case 5: e->model.mult_expK_left(g); break;
        case 6: e->model.mult_expK_right(g); break;
        case 7: e->model.mult_inv_expK_left(g); break;
        case 8: e->model.mult_inv_expK_right(g); break;
        case 9: e->model.mult_trans_expK_left(g); break;
    }
    std::memcpy(m, g.a.data(), sizeof(double) * g.a.size());
}
double rc_update_ratio(void* h, int t, int i) { Engine* e = (Engine*)h; return e->model.ratio(e->gu, e->gd, t, i); }
void rc_update_greens(void* h, int t, int i) { Engine* e = (Engine*)h; e->model.update_greens(e->gu, e->gd, t, i); }
void rc_flip(void* h, int t, int i) { ((Engine*)h)->model.flip(t, i); }
// lattice tables: nn N*4, disp N*N, kstars nk*2, fourier N*nk, hopping N*N; returns nk
int rc_lattice(int side, int* nn, int* disp, double* kstars, double* fourier, double* hopping, int* points) {
    Square s; s.init(side);
    if (nn) std::memcpy(nn, s.nn.data(), sizeof(int) * s.nn.size());
    if (disp) std::memcpy(disp, s.disp.data(), sizeof(int) * s.disp.size());
    if (kstars) for (int k = 0; k < s.nk; ++k) { kstars[2 * k] = s.kx[k]; kstars[2 * k + 1] = s.ky[k]; }
    if (fourier) std::memcpy(fourier, s.fourier.data(), sizeof(double) * s.fourier.size());
    if (hopping) std::memcpy(hopping, s.hopping.a.data(), sizeof(double) * s.hopping.a.size());
    if (points) { points[0] = s.gamma; points[1] = s.xpoint; points[2] = s.mpoint; }
    return s.nk;
}
// n bernoulli(p) outcomes of a freshly seeded engine
void rc_rng_bernoulli(long long seed, int n, double p, unsigned char* out) {
    Minstd0 g; g.seed(seed);
    for (int k = 0; k < n; ++k) out[k] = bernoulli(g, p) ? 1 : 0;
}
void rc_rng_canonical(long long seed, int n, double* out) {
    Minstd0 g; g.seed(seed);
    for (int k = 0; k < n; ++k) out[k] = canonical(g);
}
void rc_expm(int n, const double* a, double* out) {
    Mat A(n); std::memcpy(A.a.data(), a, sizeof(double) * n * n);
    Mat E = expm(A); std::memcpy(out, E.a.data(), sizeof(double) * n * n);
}
// factorization: 0 svd, 1 column-pivoted QR, 2 pre-pivoted QR.  out: U, s, V
void rc_udv(int n, int factorization, const double* a, double* U, double* s, double* V) {
    Mat A(n); std::memcpy(A.a.data(), a, sizeof(double) * n * n);
    Mat Um, Vm; vector<double> sv;
    if (factorization == 0) svd(A, Um, sv, Vm);
    else if (factorization == 1) qrp_udv(A, Um, sv, Vm);
    else qr_prepivot_udv(A, Um, sv, Vm);
    std::memcpy(U, Um.a.data(), sizeof(double) * n * n);
    std::memcpy(V, Vm.a.data(), sizeof(double) * n * n);
    std::memcpy(s, sv.data(), sizeof(double) * n);
}
void rc_greens_00_bb(int n, const double* U, const double* S, const double* V, double* out) {
    Mat Um(n), Vm(n); std::memcpy(Um.a.data(), U, sizeof(double) * n * n); std::memcpy(Vm.a.data(), V, sizeof(double) * n * n);
    vector<double> s(S, S + n);
    Mat G = compute_greens_00_bb(Um, s, Vm);
    std::memcpy(out, G.a.data(), sizeof(double) * n * n);
}
void rc_inverse(int n, const double* a, double* out) {
    Mat A(n); std::memcpy(A.a.data(), a, sizeof(double) * n * n);
    Mat X = inverse(A); std::memcpy(out, X.a.data(), sizeof(double) * n * n);
}
void rc_fullpiv_solve(int n, const double* h, const double* b, double* out) {
    Mat H(n), B(n); std::memcpy(H.a.data(), h, sizeof(double) * n * n); std::memcpy(B.a.data(), b, sizeof(double) * n * n);
    Mat X = fullpiv_qr_solve(H, B); std::memcpy(out, X.a.data(), sizeof(double) * n * n);
}

}  // extern "C"

// ============================================================================
// Command-line driver: the CPU baseline (Dqmc::thermalize / measure / analyse,
// src/dqmc.cpp:48-158), one chain per process.
// ============================================================================
#ifdef REFCPU_MAIN
int main(int argc, char** argv) {
    int side = 4, slices = 80, pace = 10, kind = 0, cb = 0, warm = 100, bins = 10, bin_size = 20, between = 0;
    double beta = 4.0, t = 1.0, U = 4.0, mu = 0.0, max_seconds = -1.0;
    long long seed = 42;
    int fact = 0, dynamic = 0, stable_sign = 1;   // the engine's sign definition (--stable-sign 0: the reference's)
    for (int a = 1; a + 1 < argc; a += 2) {
        std::string k = argv[a]; const char* v = argv[a + 1];
        if (k == "--side") side = atoi(v); else if (k == "--slices") slices = atoi(v);
        else if (k == "--beta") beta = atof(v); else if (k == "--pace") pace = atoi(v);
        else if (k == "--model") kind = atoi(v); else if (k == "--cb") cb = atoi(v);
        else if (k == "--t") t = atof(v); else if (k == "--U") U = atof(v); else if (k == "--mu") mu = atof(v);
        else if (k == "--warm") warm = atoi(v); else if (k == "--bins") bins = atoi(v);
        else if (k == "--bin-size") bin_size = atoi(v); else if (k == "--between") between = atoi(v);
        else if (k == "--seed") seed = atoll(v); else if (k == "--fact") fact = atoi(v);
        else if (k == "--max-seconds") max_seconds = atof(v);
        else if (k == "--dynamic") dynamic = atoi(v);
        else if (k == "--stable-sign") stable_sign = atoi(v);
    }
    Engine e;
    e.setup(side, slices, beta, pace, kind, cb != 0, t, U, mu, -1, fact);
    e.use_stable_sign = stable_sign != 0;
    e.rng.seed(seed);
    e.model.set_fields_random(e.rng);
    e.init();
    auto t0 = std::chrono::steady_clock::now();
    long sweeps = 0;
    auto elapsed = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    bool stop = false;
    for (int s = 0; s < warm / 2 && !stop; ++s) {
        e.sweep_0_to_beta(); e.sweep_beta_to_0(); sweeps += 2;
        if (max_seconds > 0 && elapsed() > max_seconds) stop = true;
    }
    vector<vector<double>> bin_vals, dyn_bins;   // dyn: |sign|, rho_s, gk (nk x L), dos (L), dss (L)
    for (int b = 0; b < bins && !stop; ++b) {
        for (int s = 0; s < bin_size / 2; ++s) {   // Dqmc::sweep_forth_and_back (src/dqmc.cpp:48-70)
            if (dynamic) { e.sweep_dynamic(); e.dynamic_measure(); }
            else { e.sweep_0_to_beta(); e.equaltime_measure(); }
            e.sweep_beta_to_0(); e.equaltime_measure();
            sweeps += 2;
        }
        vector<double> v(kNumObs);
        const double sg = e.acc[kSign] / e.cnt[kSign];
        for (int o = 1; o < kNumObs; ++o) v[o] = e.acc[o] / (e.cnt[o] * sg);
        v[0] = std::fabs(sg);
        bin_vals.push_back(v);
        e.clear_measure();
        if (dynamic) {   // MeasureHandler::normalize_stats, dynamic branch (src/measure/measure_handler.cpp:191-206)
            const double dsg = e.dsign / e.dcount, den = e.dcount * dsg;
            vector<double> w = {std::fabs(dsg), e.drho / den};
            for (double x : e.dgk) w.push_back(x / den);
            for (double x : e.ddos) w.push_back(x / den);
            for (double x : e.ddss) w.push_back(x / den);
            dyn_bins.push_back(w);
            e.clear_dynamic();
        }
        for (int s = 0; s < between / 2; ++s) { e.sweep_0_to_beta(); e.sweep_beta_to_0(); sweeps += 2; }
        if (max_seconds > 0 && elapsed() > max_seconds) stop = true;
    }
    const double secs = elapsed();
    printf("{\"sweeps\": %ld, \"seconds\": %.6f, \"sweeps_per_s\": %.6f, \"wrap_error\": %.6e, \"bins\": %zu",
           sweeps, secs, sweeps / secs, e.wrap_error, bin_vals.size());
    const char* names[kNumObs] = {"sign", "filling", "double_occupancy", "kinetic_energy", "momentum_distribution",
                                  "local_spin_corr", "sdw", "cdw", "swave"};
    const size_t nb = bin_vals.size();
    for (int o = 0; o < kNumObs && nb > 1; ++o) {
        double m = 0.0, m2 = 0.0;
        for (auto& v : bin_vals) { m += v[o]; m2 += v[o] * v[o]; }
        m /= nb; m2 /= nb;
        const double err = std::sqrt(std::max(0.0, m2 - m * m)) / std::sqrt((double)nb - 1);
        printf(", \"%s\": [%.12f, %.12f]", names[o], m, err);
    }
    if (dynamic && dyn_bins.size() > 1) {
        const size_t nd = dyn_bins.size(), nv = dyn_bins[0].size();
        vector<double> m(nv, 0.0), err(nv, 0.0);
        for (size_t q = 0; q < nv; ++q) {
            double a = 0.0, a2 = 0.0;
            for (auto& w : dyn_bins) { a += w[q]; a2 += w[q] * w[q]; }
            a /= nd; a2 /= nd;
            m[q] = a;
            err[q] = std::sqrt(std::max(0.0, a2 - a * a)) / std::sqrt((double)nd - 1);
        }
        auto arr = [&](const char* name, size_t o, size_t n) {
            printf(", \"%s\": [[", name);
            for (size_t q = 0; q < n; ++q) printf("%s%.12e", q ? ", " : "", m[o + q]);
            printf("], [");
            for (size_t q = 0; q < n; ++q) printf("%s%.12e", q ? ", " : "", err[o + q]);
            printf("]]");
        };
        const size_t nk = e.dgk.size() / slices;
        arr("dynamic_sign", 0, 1);
        arr("superfluid_stiffness", 1, 1);
        arr("greens_functions", 2, nk * slices);
        arr("density_of_states", 2 + nk * slices, slices);
        arr("dynamic_spin_susceptibility", 2 + nk * slices + slices, slices);
    }
    printf("}\n");
    return 0;
}
#endif
