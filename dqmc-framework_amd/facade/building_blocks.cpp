// The facade's building blocks driven the way the reference's walker drives them
// (DqmcWalker::initial_svd_stacks and initial_greens_functions, src/dqmc_walker.cpp:109-147):
// the modules are set up as in the reference's main, the fields drawn at random, and then
//   B1, B0      the identity multiplied from the left by B(t) for t = 1 and t = 0 (the slice at
//               beta), spin up (ModelBase::mult_B_from_left)
//   g0_stack    G(0) for spin up from a right stack built with mult_transB_from_left and
//               SvdStack::push every `pace` slices, then NumericalStable::compute_equaltime_greens
//               with an empty left stack
//   g0_engine   G(0) for spin up from the engine's own initialization (initial_dqmc)
// printed as one JSON line (matrices column-major).
//
//   building_blocks side time_size beta pace model(0|1) checkerboard(0|1) U mu seed
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "dqmc_framework.hpp"

static std::vector<double> identity(int n) {
    std::vector<double> m((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) m[(size_t)i * n + i] = 1.0;
    return m;
}

static void print_array(const char* key, const std::vector<double>& v, bool last = false) {
    std::printf("\"%s\": [", key);
    for (size_t e = 0; e < v.size(); ++e) std::printf(e ? ", %.17g" : "%.17g", v[e]);
    std::printf(last ? "]" : "], ");
}

int main(int argc, char** argv) {
    if (argc != 10) {
        std::fprintf(stderr, "usage: %s side time_size beta pace model cb U mu seed\n", argv[0]);
        return 2;
    }
    const int side = std::atoi(argv[1]), time_size = std::atoi(argv[2]);
    const double beta = std::atof(argv[3]);
    const int pace = std::atoi(argv[4]), model_type = std::atoi(argv[5]), cb = std::atoi(argv[6]);
    const double U = std::atof(argv[7]), mu = std::atof(argv[8]);
    const int seed = std::atoi(argv[9]);

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
    meas.set_measure_params(0, 1, 2, 0);
    meas.set_observables({"double_occupancy"});
    CheckerBoard::Square checkerboard;
    if (cb) checkerboard.set_checkerboard_params(lattice, *model, walker);
    QuantumMonteCarlo::DqmcInitializer::initial_modules(lattice, *model, walker, meas, cb ? &checkerboard : nullptr);
    model->set_bosonic_fields_to_random();

    const int n = lattice.SpaceSize();
    std::vector<double> b1 = identity(n), b0 = identity(n);
    model->mult_B_from_left(b1, 1, +1);
    model->mult_B_from_left(b0, 0, +1);

    // src/dqmc_walker.cpp:118-133: the right stack from B(t)^T products, t = L .. 1
    Utils::SvdStack left(n, (time_size + pace - 1) / pace), right(n, (time_size + pace - 1) / pace);
    std::vector<double> tmp = identity(n);
    for (int t = time_size; t >= 1; --t) {
        model->mult_transB_from_left(tmp, t, +1);
        if ((t - 1) % pace == 0) {
            right.push(tmp);
            tmp = identity(n);
        }
    }
    std::vector<double> g0_stack;
    Utils::NumericalStable::compute_equaltime_greens(left, right, g0_stack);

    QuantumMonteCarlo::DqmcInitializer::initial_dqmc(lattice, *model, walker, meas);
    const std::vector<double> g0_engine = walker.GreenttUp();

    std::vector<double> fields(model->BosonicFields().begin(), model->BosonicFields().end());
    std::printf("{\"stack_length\": %d, ", right.StackLength());
    print_array("fields", fields);
    print_array("B1", b1);
    print_array("B0", b0);
    print_array("g0_stack", g0_stack);
    print_array("g0_engine", g0_engine, true);
    std::printf("}\n");
    return 0;
}
