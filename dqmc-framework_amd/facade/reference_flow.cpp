// A driver written the way the reference's main drives its modules
// (src/dqmc_main.cpp:110-200: set_seed, initial_modules, random fields,
// initial_dqmc, thermalize, measure, analyse), on the MI355X engine through
// the C++ facade.  Prints one JSON line with the equal-time observables.
//
//   reference_flow side time_size beta pace model(0|1) checkerboard(0|1) U mu seed warmup bins bin_size between
//                  [dynamic [out_dir]]
//
// With out_dir the driver also writes the reference's output files there (src/dqmc_main.cpp:270-360):
// fields.out, kstars.out, tgrids.out, swave.out / swave.bins.out and, with dynamic = 1,
// dos, greens and dss (.out / .bins.out).
// With dynamic = 1 the handler also measures the four dynamic observables over every k-star
// (the reference's momentum_list = "KstarsAll"), so each measuring forth-and-back starts with
// the time-displaced sweep (src/dqmc.cpp:48-70).
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <utility>

#include "dqmc_framework.hpp"

int main(int argc, char** argv) {
    if (argc < 14 || argc > 16) {
        std::fprintf(stderr,
                     "usage: %s side time_size beta pace model cb U mu seed warmup bins bin_size between [dynamic [out_dir]]\n",
                     argv[0]);
        return 2;
    }
    const bool dynamic = argc >= 15 && std::atoi(argv[14]) != 0;
    const std::string out_dir = argc == 16 ? argv[15] : "";
    const int side = std::atoi(argv[1]), time_size = std::atoi(argv[2]);
    const double beta = std::atof(argv[3]);
    const int pace = std::atoi(argv[4]), model_type = std::atoi(argv[5]), cb = std::atoi(argv[6]);
    const double U = std::atof(argv[7]), mu = std::atof(argv[8]);
    const int seed = std::atoi(argv[9]), warmup = std::atoi(argv[10]), bins = std::atoi(argv[11]);
    const int bin_size = std::atoi(argv[12]), between = std::atoi(argv[13]);

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
    meas.set_measure_params(warmup, bins, bin_size, between);
    Measure::ObsList obs = {"filling_number", "double_occupancy", "kinetic_energy", "momentum_distribution",
                            "local_spin_corr", "spin_density_structure_factor", "charge_density_structure_factor",
                            "s_wave_pairing_corr"};
    if (dynamic) {
        for (const auto& n : Measure::DynamicNames()) obs.push_back(n);
        std::vector<int> ks(lattice.kStarsNum());
        for (int k = 0; k < lattice.kStarsNum(); ++k) ks[k] = k;
        meas.set_measured_momentum_list(ks);
    }
    meas.set_observables(obs);
    meas.set_measured_momentum(lattice.MPointIndex());

    CheckerBoard::Square checkerboard;
    if (cb) checkerboard.set_checkerboard_params(lattice, *model, walker);
    QuantumMonteCarlo::DqmcInitializer::initial_modules(lattice, *model, walker, meas, cb ? &checkerboard : nullptr);
    model->set_bosonic_fields_to_random();
    QuantumMonteCarlo::DqmcInitializer::initial_dqmc(lattice, *model, walker, meas);

    QuantumMonteCarlo::Dqmc::thermalize(walker, *model, lattice, meas);
    QuantumMonteCarlo::Dqmc::measure(walker, *model, lattice, meas);
    QuantumMonteCarlo::Dqmc::analyse(meas);

    if (!out_dir.empty()) {
        using QuantumMonteCarlo::DqmcIO;
        DqmcIO::output_bosonic_fields(out_dir + "/fields.out", *model);
        DqmcIO::output_k_stars(out_dir + "/kstars.out", lattice);
        DqmcIO::output_imaginary_time_grids(out_dir + "/tgrids.out", walker);
        DqmcIO::output_observable(out_dir + "/swave.out", meas, "s_wave_pairing_corr");
        DqmcIO::output_observable_in_bins(out_dir + "/swave.bins.out", meas, "s_wave_pairing_corr");
        if (dynamic) {
            for (const auto& [file, name] : {std::pair<std::string, std::string>{"dos", "density_of_states"},
                                             {"greens", "greens_functions"}, {"dss", "dynamic_spin_susceptibility"}}) {
                DqmcIO::output_observable(out_dir + "/" + file + ".out", meas, name);
                DqmcIO::output_observable_in_bins(out_dir + "/" + file + ".bins.out", meas, name);
            }
        }
    }

    const std::vector<double> g = walker.GreenttUp();
    double tr = 0.0;
    for (int i = 0; i < side * side; ++i) tr += g[(size_t)i * side * side + i];
    std::printf("{\"wrap_error\": %.6e, \"sign\": %.1f, \"trace_g_up\": %.15f", walker.WrapError(), walker.ConfigSign(), tr);
    for (const auto& name : Measure::EqualTimeNames())
        std::printf(", \"%s\": [%.15f, %.15f]", name.c_str(), meas.mean(name), meas.error(name));
    if (dynamic) {
        for (const char* name : {"dynamic_sign", "superfluid_stiffness", "greens_functions", "density_of_states",
                                 "dynamic_spin_susceptibility"}) {
            const auto m = meas.dynamic_mean(name), e = meas.dynamic_error(name);
            std::printf(", \"%s\": [[", name);
            for (size_t q = 0; q < m.size(); ++q) std::printf("%s%.15e", q ? ", " : "", m[q]);
            std::printf("], [");
            for (size_t q = 0; q < e.size(); ++q) std::printf("%s%.15e", q ? ", " : "", e[q]);
            std::printf("]]");
        }
    }
    std::printf("}\n");
    return 0;
}
