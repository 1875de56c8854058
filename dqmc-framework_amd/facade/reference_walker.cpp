// The reference's walker written against the facade: DqmcWalker::initial_svd_stacks,
// initial_greens_functions, initial_config_sign, metropolis_update, wrap_from_0_to_beta,
// wrap_from_beta_to_0, sweep_from_0_to_beta and sweep_from_beta_to_0
// (src/dqmc_walker.cpp:109-412) with the same calls in the same order, on the MI355X engine:
// the Green's functions stay on the device (the walker's GreensFunc handles), the model's
// per-site plugin calls (get_update_ratio, update_greens_function, update_bosonic_field) and
// the wraps act on them there, the Metropolis draws come from the global engine
// (Utils::Random::Engine), and the stacks are the facade's Utils::SvdStack with
// NumericalStable::compute_equaltime_greens, as the reference's walker keeps them.
//
//   reference_walker side time_size beta pace model(0|1) checkerboard(0|1) U mu seed sweeps
//
// Runs `sweeps` single-direction sweeps (0 -> beta first) and prints one JSON line per sweep:
// accept bits (t-major, site-minor), fields, configuration sign, wrap error and the final G of
// both spins (column-major).
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "dqmc_framework.hpp"

namespace {

using Matrix = std::vector<double>;

Matrix identity(int n) {
    Matrix m((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) m[(size_t)i * n + i] = 1.0;
    return m;
}

// sign of det(A) by partial-pivot LU (Eigen's determinant(), src/dqmc_walker.cpp:162)
double det_sign(Matrix a, int n) {
    double sg = 1.0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int r = k + 1; r < n; ++r)
            if (std::fabs(a[(size_t)k * n + r]) > std::fabs(a[(size_t)k * n + p])) p = r;
        if (p != k) {
            sg = -sg;
            for (int c = 0; c < n; ++c) std::swap(a[(size_t)c * n + k], a[(size_t)c * n + p]);
        }
        const double piv = a[(size_t)k * n + k];
        if (piv < 0) sg = -sg;
        for (int r = k + 1; r < n; ++r) {
            const double l = a[(size_t)k * n + r] / piv;
            for (int c = k + 1; c < n; ++c) a[(size_t)c * n + r] -= l * a[(size_t)c * n + k];
        }
    }
    return sg;
}

double max_abs_diff(const Matrix& a, const Matrix& b) {
    double e = 0.0;
    for (size_t k = 0; k < a.size(); ++k) e = std::max(e, std::fabs(a[k] - b[k]));
    return e;
}

struct Walker {
    QuantumMonteCarlo::DqmcWalker& w;
    Model::ModelBase& model;
    int N, L, pace;
    double sign = 1.0, wrap_error = 0.0;
    std::unique_ptr<Utils::SvdStack> left_up, left_dn, right_up, right_dn;
    std::vector<uint8_t> accepts;   // of the running sweep, (t, i) at t * N + i

    Walker(QuantumMonteCarlo::DqmcWalker& w_, Model::ModelBase& m, int N_, int L_, int p_)
        : w(w_), model(m), N(N_), L(L_), pace(p_), accepts((size_t)L_ * N_, 0) {
        const int len = (L + pace - 1) / pace;
        left_up.reset(new Utils::SvdStack(N, len));
        left_dn.reset(new Utils::SvdStack(N, len));
        right_up.reset(new Utils::SvdStack(N, len));
        right_dn.reset(new Utils::SvdStack(N, len));
    }

    // src/dqmc_walker.cpp:109-133
    void initial_svd_stacks() {
        Matrix up = identity(N), dn = identity(N);
        for (int t = L; t >= 1; --t) {
            model.mult_transB_from_left(up, t, +1);
            model.mult_transB_from_left(dn, t, -1);
            if ((t - 1) % pace == 0) {
                right_up->push(up);
                right_dn->push(dn);
                up = identity(N);
                dn = identity(N);
            }
        }
    }
    // src/dqmc_walker.cpp:136-163
    void initial_greens_functions_and_sign() {
        Matrix gu, gd;
        Utils::NumericalStable::compute_equaltime_greens(*left_up, *right_up, gu);
        Utils::NumericalStable::compute_equaltime_greens(*left_dn, *right_dn, gd);
        w.GreenttUp() = gu;
        w.GreenttDn() = gd;
        sign = det_sign(gu, N) * det_sign(gd, N) >= 0 ? +1.0 : -1.0;
    }

    // src/dqmc_walker.cpp:173-197
    void metropolis_update(int t) {
        const int eff_t = (t == 0) ? L - 1 : t - 1;
        for (int i = 0; i < N; ++i) {
            const double r = model.get_update_ratio(w, eff_t, i);
            const bool ok = std::bernoulli_distribution(std::min(1.0, std::abs(r)))(Utils::Random::Engine());
            accepts[(size_t)eff_t * N + i] = ok ? 1 : 0;
            if (ok) {
                model.update_greens_function(w, eff_t, i);
                model.update_bosonic_field(eff_t, i);
                sign = (r >= 0) ? +sign : -sign;
            }
        }
    }
    // src/dqmc_walker.cpp:208-237
    void wrap_from_0_to_beta(int t) {
        const int eff_t = (t == L) ? 1 : t + 1;
        model.mult_B_from_left(w.GreenttUp(), eff_t, +1);
        model.mult_invB_from_right(w.GreenttUp(), eff_t, +1);
        model.mult_B_from_left(w.GreenttDn(), eff_t, -1);
        model.mult_invB_from_right(w.GreenttDn(), eff_t, -1);
    }
    void wrap_from_beta_to_0(int t) {
        const int eff_t = (t == 0) ? L : t;
        model.mult_B_from_right(w.GreenttUp(), eff_t, +1);
        model.mult_invB_from_left(w.GreenttUp(), eff_t, +1);
        model.mult_B_from_right(w.GreenttDn(), eff_t, -1);
        model.mult_invB_from_left(w.GreenttDn(), eff_t, -1);
    }
    // the fresh G from the stacks replaces the wrapped one; the wrap error is collected
    void stabilize() {
        Matrix gu, gd;
        Utils::NumericalStable::compute_equaltime_greens(*left_up, *right_up, gu);
        Utils::NumericalStable::compute_equaltime_greens(*left_dn, *right_dn, gd);
        const Matrix wu = w.GreenttUp(), wd = w.GreenttDn();
        wrap_error = std::max(wrap_error, std::max(max_abs_diff(gu, wu), max_abs_diff(gd, wd)));
        w.GreenttUp() = gu;
        w.GreenttDn() = gd;
    }
    // src/dqmc_walker.cpp:246-325
    void sweep_from_0_to_beta() {
        Matrix up = identity(N), dn = identity(N);
        for (int t = 1; t <= L; ++t) {
            wrap_from_0_to_beta(t - 1);
            metropolis_update(t);
            model.mult_B_from_left(up, t, +1);
            model.mult_B_from_left(dn, t, -1);
            if (t % pace == 0 || t == L) {
                right_up->pop();
                right_dn->pop();
                left_up->push(up);
                left_dn->push(dn);
                stabilize();
                up = identity(N);
                dn = identity(N);
            }
        }
    }
    // src/dqmc_walker.cpp:334-412
    void sweep_from_beta_to_0() {
        Matrix up = identity(N), dn = identity(N);
        for (int t = L; t >= 1; --t) {
            if (t % pace == 0 && t != L) {
                left_up->pop();
                left_dn->pop();
                right_up->push(up);
                right_dn->push(dn);
                stabilize();
                up = identity(N);
                dn = identity(N);
            }
            metropolis_update(t);
            model.mult_transB_from_left(up, t, +1);
            model.mult_transB_from_left(dn, t, -1);
            wrap_from_beta_to_0(t);
        }
        left_up->pop();
        left_dn->pop();
        right_up->push(up);
        right_dn->push(dn);
        Matrix gu, gd;
        Utils::NumericalStable::compute_equaltime_greens(*left_up, *right_up, gu);
        Utils::NumericalStable::compute_equaltime_greens(*left_dn, *right_dn, gd);
        w.GreenttUp() = gu;
        w.GreenttDn() = gd;
    }
};

void print_list(const char* name, const std::vector<double>& v, bool last = false) {
    std::printf("\"%s\": [", name);
    for (size_t k = 0; k < v.size(); ++k) std::printf("%s%.17g", k ? ", " : "", v[k]);
    std::printf("]%s", last ? "" : ", ");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 11) {
        std::fprintf(stderr, "usage: %s side time_size beta pace model cb U mu seed sweeps\n", argv[0]);
        return 2;
    }
    const int side = std::atoi(argv[1]), time_size = std::atoi(argv[2]);
    const double beta = std::atof(argv[3]);
    const int pace = std::atoi(argv[4]), model_type = std::atoi(argv[5]), cb = std::atoi(argv[6]);
    const double U = std::atof(argv[7]), mu = std::atof(argv[8]);
    const int seed = std::atoi(argv[9]), sweeps = std::atoi(argv[10]);

    Utils::Random::set_seed(seed);
    Lattice::Square lattice;
    lattice.set_lattice_params({side, side});
    lattice.initial();
    std::unique_ptr<Model::ModelBase> model;
    if (model_type == 1) model = std::make_unique<Model::AttractiveHubbard>();
    else model = std::make_unique<Model::RepulsiveHubbard>();
    model->set_model_params(1.0, U, mu);
    QuantumMonteCarlo::DqmcWalker walker;
    walker.set_physical_params(beta, time_size);
    walker.set_stabilization_pace(pace);
    Measure::MeasureHandler meas;
    meas.set_measure_params(0, 0, 0, 0);
    CheckerBoard::Square checkerboard;
    if (cb) checkerboard.set_checkerboard_params(lattice, *model, walker);
    QuantumMonteCarlo::DqmcInitializer::initial_modules(lattice, *model, walker, meas, cb ? &checkerboard : nullptr);
    model->set_bosonic_fields_to_random();
    QuantumMonteCarlo::DqmcInitializer::initial_dqmc(lattice, *model, walker, meas);

    const int N = side * side;
    Walker wk(walker, *model, N, time_size, pace);
    wk.initial_svd_stacks();
    wk.initial_greens_functions_and_sign();
    for (int s = 0; s < sweeps; ++s) {
        std::fill(wk.accepts.begin(), wk.accepts.end(), 0);
        if (s % 2 == 0) wk.sweep_from_0_to_beta();
        else wk.sweep_from_beta_to_0();
        std::vector<double> acc(wk.accepts.begin(), wk.accepts.end());
        std::vector<double> fields(model->BosonicFields().begin(), model->BosonicFields().end());
        std::printf("{\"sweep\": %d, \"sign\": %.17g, \"wrap_error\": %.17g, ", s, wk.sign, wk.wrap_error);
        print_list("accepts", acc);
        print_list("fields", fields);
        print_list("green_up", walker.GreenttUp());
        print_list("green_dn", walker.GreenttDn(), true);
        std::printf("}\n");
        std::fflush(stdout);
    }
    return 0;
}
