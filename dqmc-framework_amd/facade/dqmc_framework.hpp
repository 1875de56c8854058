// dqmc_framework.hpp — C++ host mirror of dqmc-framework's plugin interface
// for the equal-time sweep, running on the MI355X engine through the C-ABI
// of include/dqmc_amd.h.
//
// Class names, method names, argument meaning and call order follow the
// reference so that a driver written against it reads the same:
//   Lattice::Square                 include/lattice/square.h, src/lattice/square.cpp
//   Model::RepulsiveHubbard /
//   Model::AttractiveHubbard        include/model/*_hubbard.h (set_model_params, link)
//   CheckerBoard::Square            include/checkerboard/square.h
//   QuantumMonteCarlo::DqmcWalker   include/dqmc_walker.h:107-184
//   Measure::MeasureHandler         include/measure/measure_handler.h:60-118
//   QuantumMonteCarlo::DqmcInitializer::initial_modules / initial_dqmc
//                                   include/dqmc_initializer.h:47-78
//   QuantumMonteCarlo::Dqmc::thermalize / measure / analyse   include/dqmc.h:61-73
//   QuantumMonteCarlo::DqmcIO       include/dqmc_io.h (file outputs and the field file)
//   Utils::Random::set_seed         include/random.h:12-17
//   Utils::SvdStack,
//   Utils::NumericalStable::compute_equaltime_greens   include/svd_stack.h:70-91,
//                                   include/utils/numerical_stable.hpp:186-238
//   ModelBase::mult_{B,invB}_from_{left,right}, mult_transB_from_left
//                                   include/model/model_base.h:111-163
// Matrices are column-major std::vector<double> (Eigen::MatrixXd layout).
// Errors: the reference asserts or prints and exits; here a failing engine
// call throws std::runtime_error with the engine's message.
#pragma once

#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dqmc_amd.h"

namespace Utils {
// The reference's global engine (std::default_random_engine, src/random.cpp:6).
class Random {
   public:
    static std::default_random_engine& Engine() {
        static std::default_random_engine e(0);
        return e;
    }
    static void set_seed(const int seed) { Engine().seed(seed); }
    // state of the engine (minstd_rand0 keeps one integer)
    static unsigned long long state() {
        std::ostringstream os;
        os << Engine();
        return std::stoull(os.str());
    }
    static void set_state(unsigned long long x) {
        std::istringstream is(std::to_string(x));
        is >> Engine();
    }
};
}  // namespace Utils

namespace dqmc_detail {
inline void check(int rc, dqmc_ctx_t ctx, const char* what) {
    if (rc != DQMC_OK) throw std::runtime_error(std::string(what) + ": " + (ctx ? dqmc_context_error(ctx) : "") +
                                                " (status " + std::to_string(rc) + ")");
}
inline dqmc_ctx_t context() {
    static dqmc_ctx_t ctx = [] {
        dqmc_ctx_t c = nullptr;
        if (dqmc_context_create(0, &c) != DQMC_OK) throw std::runtime_error("dqmc_context_create(0) failed");
        return c;
    }();
    return ctx;
}
}  // namespace dqmc_detail

namespace Utils {
// Utils::SvdStack (include/svd_stack.h:70-91, src/svd_stack.cpp:15-87) on the engine's UDV push:
// each level keeps U, D and the cumulative V of the product pushed so far, so MatrixV() is the
// reference's product V_0 V_1 ... of the levels.  The engine factors by pre-pivoted QR: U is
// orthogonal and D graded like singular values, V carries the triangular factors (not orthogonal).
class SvdStack {
    int m_dim{0}, m_len{0};
    std::vector<std::vector<double>> m_u, m_d, m_v;

   public:
    SvdStack() = default;
    SvdStack(int mat_dim, int stack_length)
        : m_dim(mat_dim), m_u(stack_length), m_d(stack_length), m_v(stack_length) {}
    bool empty() const { return m_len == 0; }
    int MatDim() const { return m_dim; }
    int StackLength() const { return m_len; }
    std::vector<double> SingularValues() const { return top(m_d); }
    std::vector<double> MatrixU() const { return top(m_u); }
    std::vector<double> MatrixV() const { return top(m_v); }
    void clear() { m_len = 0; }
    // adds `matrix` from the left: the stack's product becomes matrix * (previous product)
    void push(const std::vector<double>& matrix) {
        const size_t n2 = (size_t)m_dim * m_dim;
        if (matrix.size() != n2) throw std::invalid_argument("SvdStack::push: matrix size");
        if (m_len >= (int)m_u.size()) throw std::length_error("SvdStack::push: stack is full");
        std::vector<double> U(n2), D(m_dim), V(n2);
        const bool e = empty();
        dqmc_ctx_t ctx = dqmc_detail::context();
        dqmc_detail::check(dqmc_k_udv_push(ctx, m_dim, matrix.data(), e ? nullptr : m_u[m_len - 1].data(),
                                           e ? nullptr : m_d[m_len - 1].data(), e ? nullptr : m_v[m_len - 1].data(),
                                           U.data(), D.data(), V.data()),
                           ctx, "SvdStack::push");
        m_u[m_len] = std::move(U);
        m_d[m_len] = std::move(D);
        m_v[m_len] = std::move(V);
        ++m_len;
    }
    void pop() {
        if (m_len == 0) throw std::logic_error("SvdStack::pop: stack is empty");
        --m_len;
    }

   private:
    std::vector<double> top(const std::vector<std::vector<double>>& level) const {
        if (m_len == 0) throw std::logic_error("SvdStack: stack is empty");
        return level[m_len - 1];
    }
};

// Utils::NumericalStable::compute_equaltime_greens (include/utils/numerical_stable.hpp:186-238):
// gtt = (1 + B_left B_right^T)^-1 from the two stacks' products, on the device
class NumericalStable {
   public:
    static void compute_equaltime_greens(SvdStack& left, SvdStack& right, std::vector<double>& gtt) {
        if (left.empty() && right.empty()) throw std::logic_error("compute_equaltime_greens: both stacks are empty");
        if (!left.empty() && !right.empty() && left.MatDim() != right.MatDim())
            throw std::invalid_argument("compute_equaltime_greens: stack dimensions differ");
        const int n = left.empty() ? right.MatDim() : left.MatDim();
        gtt.assign((size_t)n * n, 0.0);
        std::vector<double> ul, dl, vl, ur, dr, vr;
        if (!left.empty()) { ul = left.MatrixU(); dl = left.SingularValues(); vl = left.MatrixV(); }
        if (!right.empty()) { ur = right.MatrixU(); dr = right.SingularValues(); vr = right.MatrixV(); }
        auto ptr = [](const std::vector<double>& v) { return v.empty() ? nullptr : v.data(); };
        dqmc_ctx_t ctx = dqmc_detail::context();
        dqmc_detail::check(dqmc_k_greens_recompute(ctx, n, ptr(ul), ptr(dl), ptr(vl), ptr(ur), ptr(dr), ptr(vr),
                                                   gtt.data()),
                           ctx, "NumericalStable::compute_equaltime_greens");
    }
};
}  // namespace Utils

namespace Lattice {
using LatticeIntVec = std::vector<int>;

class LatticeBase {
   protected:
    bool m_initial_status{false};
    int m_space_dim{}, m_side_length{}, m_space_size{}, m_coordination_number{}, m_num_k_stars{};

   public:
    virtual ~LatticeBase() = default;
    virtual void set_lattice_params(const LatticeIntVec& side_length_vec) = 0;
    virtual void initial() = 0;
    bool InitialStatus() const { return m_initial_status; }
    int SpaceDim() const { return m_space_dim; }
    int SpaceSize() const { return m_space_size; }
    int SideLength() const { return m_side_length; }
    int CoordinationNumber() const { return m_coordination_number; }
    int kStarsNum() const { return m_num_k_stars; }
};

// 2-D periodic square lattice (src/lattice/square.cpp:16-196)
class Square : public LatticeBase {
    int m_gamma{}, m_x{}, m_m{};

   public:
    void set_lattice_params(const LatticeIntVec& v) override {
        if (v.size() != 2 || v[0] != v[1] || v[0] < 2) throw std::invalid_argument("square lattice needs {L, L}, L >= 2");
        m_space_dim = 2;
        m_coordination_number = 4;
        m_side_length = v[0];
        m_space_size = v[0] * v[1];
    }
    void initial() override {
        const int h = m_side_length / 2;
        m_num_k_stars = (h + 1) * (h + 2) / 2;
        m_gamma = 0;
        m_x = m_num_k_stars - h - 1;
        m_m = m_num_k_stars - 1;
        m_initial_status = true;
    }
    int GammaPointIndex() const { return m_gamma; }
    int XPointIndex() const { return m_x; }
    int MPointIndex() const { return m_m; }
    int NearestNeighbour(int i, int dir) const {
        const int L = m_side_length, x = i % L, y = i / L;
        switch (dir) {
            case 0: return (x + 1) % L + L * y;
            case 1: return x + L * ((y + 1) % L);
            case 2: return (x - 1 + L) % L + L * y;
            default: return x + L * ((y - 1 + L) % L);
        }
    }
    int Displacement(int i, int j) const {
        const int L = m_side_length;
        return (j % L - i % L + L) % L + L * ((j / L - i / L + L) % L);
    }
};
}  // namespace Lattice

namespace QuantumMonteCarlo {
class DqmcWalker;
class DqmcInitializer;
class DqmcIO;

// The walker's equal-time Green's function of one spin, resident on the device: the
// reference's GreensFunc& that the model changes in place (include/dqmc_walker.h:64-72,
// src/model/repulsive_hubbard.cpp:127-155).  Converting it to a host matrix copies it from the
// device; assigning a host matrix copies it to the device.  The attractive model keeps one
// matrix on the device (G_dn equals G_up bit for bit in the reference,
// src/model/attractive_hubbard.cpp:149,175): its spin-down handle mirrors the spin-up one, and
// products and assignments on the mirror are left out, since the reference applies the same
// ones to both matrices.
class GreensFunc {
    dqmc_batch_t m_batch{nullptr};
    int m_spin{+1}, m_dim{0};
    bool m_mirror{false};
    friend class DqmcWalker;
    friend class DqmcInitializer;

   public:
    int rows() const { return m_dim; }
    int cols() const { return m_dim; }
    int spin() const { return m_spin; }
    bool mirror() const { return m_mirror; }
    dqmc_batch_t batch() const { return m_batch; }
    std::vector<double> host() const {
        if (!m_batch) throw std::logic_error("GreensFunc: walker not initialized");
        std::vector<double> g((size_t)m_dim * m_dim);
        dqmc_detail::check(dqmc_get_green(m_batch, 0, m_spin, g.data()), dqmc_detail::context(), "GreensFunc::host");
        return g;
    }
    operator std::vector<double>() const { return host(); }
    GreensFunc& operator=(const std::vector<double>& g) {
        if (!m_batch) throw std::logic_error("GreensFunc: walker not initialized");
        if (g.size() != (size_t)m_dim * m_dim) throw std::invalid_argument("GreensFunc: matrix size");
        if (!m_mirror)
            dqmc_detail::check(dqmc_set_green(m_batch, 0, m_spin, g.data()), dqmc_detail::context(), "GreensFunc::assign");
        return *this;
    }
    // element (i, j) (one copy of the matrix per call)
    double operator()(int i, int j) const { return host()[(size_t)j * m_dim + i]; }
};
}  // namespace QuantumMonteCarlo

namespace CheckerBoard {
class CheckerBoardBase {
   public:
    virtual ~CheckerBoardBase() = default;
};
// The break-up itself runs on the device; this object records that the model
// is linked to it (src/model/model_base.cpp:48-55).
class Square : public CheckerBoardBase {
   public:
    template <class L, class M, class W>
    void set_checkerboard_params(const L& lattice, const M&, const W&) {
        if (lattice.SideLength() % 2 != 0) throw std::invalid_argument("checkerboard break-up needs an even side length");
    }
    void initial() {}
};
}  // namespace CheckerBoard

namespace Model {
class ModelBase {
   protected:
    int m_model_type{DQMC_MODEL_REPULSIVE};
    double m_hopping_t{1.0}, m_onsite_u{4.0}, m_chemical_potential{0.0};
    bool m_checkerboard{false};
    int m_space_size{}, m_time_size{};
    std::vector<int8_t> m_bosonic_field;  // (t, i) at t * N + i
    dqmc_batch_t m_batch{nullptr};        // the chain on the device once initialized (fields live there)
    friend class QuantumMonteCarlo::DqmcWalker;
    friend class QuantumMonteCarlo::DqmcInitializer;
    friend class QuantumMonteCarlo::DqmcIO;

   public:
    virtual ~ModelBase() = default;
    virtual void set_model_params(double hopping_t, double onsite_u, double chemical_potential) {
        if (hopping_t < 0.0 || onsite_u < 0.0) throw std::invalid_argument("hopping_t and onsite_u must be >= 0");
        m_hopping_t = hopping_t;
        m_onsite_u = onsite_u;
        m_chemical_potential = chemical_potential;
    }
    double HoppingT() const { return m_hopping_t; }
    double ChemicalPotential() const { return m_chemical_potential; }
    double OnSiteU() const { return m_onsite_u; }
    void link() { m_checkerboard = false; }
    void link(const CheckerBoard::CheckerBoardBase&) { m_checkerboard = true; }
    // src/model/repulsive_hubbard.cpp:86-99: t-outer, i-inner, bernoulli(0.5) on the global engine
    void set_bosonic_fields_to_random() {
        std::bernoulli_distribution b(0.5);
        for (auto& f : m_bosonic_field) f = b(Utils::Random::Engine()) ? +1 : -1;
    }
    const std::vector<int8_t>& BosonicFields() const { return m_bosonic_field; }

    // Model::ModelBase::initial / initial_params / initial_KV_matrices (include/model/model_base.h:120-122,
    // src/model/repulsive_hubbard.cpp:44-72): the HS coupling alpha = acosh(exp(dtau |U| / 2)) and
    // exp(-dtau K), exp(+dtau K) of the dense kinetic term as the engine computes them (the
    // plane-wave closed form of K = -t H + mu I)
    void initial(const Lattice::LatticeBase& lattice, const QuantumMonteCarlo::DqmcWalker& walker);
    void initial_params(const Lattice::LatticeBase& lattice, const QuantumMonteCarlo::DqmcWalker& walker);
    void initial_KV_matrices(const Lattice::LatticeBase& lattice, const QuantumMonteCarlo::DqmcWalker& walker);
    double Alpha() const { return m_alpha; }
    const std::vector<double>& ExpK() const { return m_expK; }
    const std::vector<double>& InvExpK() const { return m_inv_expK; }

    // The per-site plugin calls of DqmcWalker::metropolis_update (src/dqmc_walker.cpp:173-197;
    // include/model/model_base.h:142-150) on the walker's resident G and the device fields; t is
    // the field row (the walker's eff_t), 0 <= t < L.
    double get_update_ratio(QuantumMonteCarlo::DqmcWalker& walker, int time_index, int space_index) const;
    void update_greens_function(QuantumMonteCarlo::DqmcWalker& walker, int time_index, int space_index);
    void update_bosonic_field(int time_index, int space_index);

    // the same products on the walker's resident G, in place (DqmcWalker::wrap_from_0_to_beta,
    // wrap_from_beta_to_0: src/dqmc_walker.cpp:208-237); the handle's spin must be the call's
    void mult_B_from_left(QuantumMonteCarlo::GreensFunc& g, int time_index, int spin) const {
        mult_dev(g, time_index, spin, 0, "mult_B_from_left");
    }
    void mult_B_from_right(QuantumMonteCarlo::GreensFunc& g, int time_index, int spin) const {
        mult_dev(g, time_index, spin, 1, "mult_B_from_right");
    }
    void mult_invB_from_left(QuantumMonteCarlo::GreensFunc& g, int time_index, int spin) const {
        mult_dev(g, time_index, spin, 2, "mult_invB_from_left");
    }
    void mult_invB_from_right(QuantumMonteCarlo::GreensFunc& g, int time_index, int spin) const {
        mult_dev(g, time_index, spin, 3, "mult_invB_from_right");
    }
    void mult_transB_from_left(QuantumMonteCarlo::GreensFunc& g, int time_index, int spin) const {
        mult_dev(g, time_index, spin, 4, "mult_transB_from_left");
    }

    // B(t) = exp(-dtau V_sigma(t)) exp(-dtau K) products on a caller's N x N matrix, in place
    // (src/model/repulsive_hubbard.cpp:158-245, src/model/attractive_hubbard.cpp:159-245): t is
    // the reference's time index in [0, L] (field row t - 1, t = 0 the slice at beta), spin +-1.
    // They run on the device with the chain's current fields (after initial_modules).
    void mult_B_from_left(std::vector<double>& green, int time_index, int spin) const {
        mult(green, time_index, spin, 0, "mult_B_from_left");
    }
    void mult_B_from_right(std::vector<double>& green, int time_index, int spin) const {
        mult(green, time_index, spin, 1, "mult_B_from_right");
    }
    void mult_invB_from_left(std::vector<double>& green, int time_index, int spin) const {
        mult(green, time_index, spin, 2, "mult_invB_from_left");
    }
    void mult_invB_from_right(std::vector<double>& green, int time_index, int spin) const {
        mult(green, time_index, spin, 3, "mult_invB_from_right");
    }
    void mult_transB_from_left(std::vector<double>& green, int time_index, int spin) const {
        mult(green, time_index, spin, 4, "mult_transB_from_left");
    }

   private:
    int m_side{};
    double m_beta{};
    double m_alpha{};
    std::vector<double> m_expK, m_inv_expK;
    void mult_dev(QuantumMonteCarlo::GreensFunc& g, int t, int spin, int op, const char* who) const {
        if (g.spin() != spin) throw std::invalid_argument(std::string(who) + ": the matrix belongs to the other spin");
        if (g.mirror()) return;
        dqmc_detail::check(dqmc_chain_mult_green(g.batch(), 0, t, spin, op), dqmc_detail::context(), who);
    }
    dqmc_chain_params chain_params(double beta, int time_size) const {
        dqmc_chain_params p{};
        p.model = m_model_type;
        p.side = m_side;
        p.time_size = time_size;
        p.stabilization_pace = time_size;
        p.checkerboard = m_checkerboard ? 1 : 0;
        p.delay = 1;
        p.momentum_index = -1;
        p.precision = DQMC_PREC_FP64;
        p.beta = beta;
        p.hopping_t = m_hopping_t;
        p.onsite_u = m_onsite_u;
        p.chemical_potential = m_chemical_potential;
        return p;
    }
    void mult(std::vector<double>& g, int t, int spin, int op, const char* who) const {
        if (m_side < 2 || g.size() != (size_t)m_space_size * m_space_size)
            throw std::invalid_argument(std::string(who) + ": model not initialized or matrix size");
        dqmc_chain_params p{};
        p.model = m_model_type;
        p.side = m_side;
        p.time_size = m_time_size;
        p.stabilization_pace = m_time_size;   // one stack level: the product needs no stack
        p.checkerboard = m_checkerboard ? 1 : 0;
        p.delay = 1;
        p.momentum_index = -1;
        p.precision = DQMC_PREC_FP64;
        p.beta = m_beta;
        p.hopping_t = m_hopping_t;
        p.onsite_u = m_onsite_u;
        p.chemical_potential = m_chemical_potential;
        dqmc_ctx_t ctx = dqmc_detail::context();
        std::vector<int8_t> f = m_bosonic_field;   // the device holds the fields once the chain runs there
        if (m_batch) dqmc_detail::check(dqmc_get_fields(m_batch, 0, f.data()), ctx, who);
        dqmc_detail::check(dqmc_k_wrap(ctx, &p, f.data(), nullptr, nullptr, t, spin, op, g.data()), ctx, who);
    }
};
class RepulsiveHubbard : public ModelBase {
   public:
    RepulsiveHubbard() { m_model_type = DQMC_MODEL_REPULSIVE; }
};
class AttractiveHubbard : public ModelBase {
   public:
    AttractiveHubbard() { m_model_type = DQMC_MODEL_ATTRACTIVE; }
};
}  // namespace Model

namespace Measure {
class MeasureHandler;
}

namespace QuantumMonteCarlo {

// The walker owns one chain of the engine (a batch of one) once initialized.
class DqmcWalker {
    int m_space_size{}, m_time_size{}, m_stabilization_pace{10}, m_delay{16};
    double m_beta{}, m_time_interval{};
    bool m_is_equaltime{false}, m_is_dynamic{false};
    dqmc_batch_t m_batch{nullptr};
    GreensFunc m_green_tt_up, m_green_tt_dn;
    friend class DqmcInitializer;
    friend class Measure::MeasureHandler;
    friend class Model::ModelBase;

    std::vector<double> fetch_dynamic(int slot, int spin, int which) const {
        std::vector<double> g((size_t)m_space_size * m_space_size);
        dqmc_detail::check(dqmc_get_dynamic_green(m_batch, 0, slot, spin, which, g.data()), dqmc_detail::context(),
                           which == 0 ? "DqmcWalker::Greent0" : "DqmcWalker::Green0t");
        return g;
    }

    std::vector<double> fetch(int slot, int spin) const {
        std::vector<double> g((size_t)m_space_size * m_space_size);
        const int rc = slot < 0 ? dqmc_get_green(m_batch, 0, spin, g.data())
                                : dqmc_get_slice_green(m_batch, 0, slot, spin, g.data());
        dqmc_detail::check(rc, dqmc_detail::context(), "DqmcWalker::Greentt");
        return g;
    }
    std::vector<std::vector<double>> fetch_all(int spin) const {
        std::vector<std::vector<double>> v;
        for (int t = 0; t < m_time_size; ++t) v.push_back(fetch(t, spin));
        return v;
    }
    void sweep_once() {
        dqmc_detail::check(dqmc_sweep(m_batch, 1, m_is_equaltime ? DQMC_SWEEP_RECORD : 0), dqmc_detail::context(),
                           "DqmcWalker::sweep");
    }

   public:
    ~DqmcWalker() {
        if (m_batch) dqmc_batch_destroy(m_batch);
    }
    void set_physical_params(double beta, int time_size) {
        if (beta <= 0.0 || time_size < 1) throw std::invalid_argument("beta > 0 and time_size >= 1");
        m_beta = beta;
        m_time_size = time_size;
        m_time_interval = beta / time_size;
    }
    void set_stabilization_pace(int p) {
        if (p < 1) throw std::invalid_argument("stabilization pace must be >= 1");
        m_stabilization_pace = p;
    }
    // delayed-update depth of the engine (no counterpart in the reference)
    void set_delay(int k) { m_delay = k; }
    int TimeSize() const { return m_time_size; }
    double Beta() const { return m_beta; }
    double TimeInterval() const { return m_time_interval; }
    int StabilizationPace() const { return m_stabilization_pace; }

    void sweep_from_0_to_beta(Model::ModelBase&) {
        dqmc_chain_info i{};
        dqmc_get_info(m_batch, 0, &i);
        if (i.pos != 0) throw std::logic_error("sweep_from_0_to_beta: walker is not at slice 0");
        sweep_once();
    }
    void sweep_from_beta_to_0(Model::ModelBase&) {
        dqmc_chain_info i{};
        dqmc_get_info(m_batch, 0, &i);
        if (i.pos != m_time_size) throw std::logic_error("sweep_from_beta_to_0: walker is not at slice beta");
        sweep_once();
    }
    // src/dqmc_walker.cpp:424-524.  The engine accumulates the dynamic observables while this
    // sweep runs (Measure::MeasureHandler::dynamic_measure is then bookkeeping only), so a walker
    // whose handler measures dynamic observables is expected to pair it with dynamic_measure, as
    // Dqmc::sweep_forth_and_back does.
    void sweep_for_dynamic_greens(Model::ModelBase&) {
        if (!m_is_dynamic) return;   // as the reference: nothing to do without dynamic observables
        dqmc_chain_info i{};
        dqmc_get_info(m_batch, 0, &i);
        if (i.pos != 0) throw std::logic_error("sweep_for_dynamic_greens: walker is not at slice 0");
        dqmc_detail::check(dqmc_sweep(m_batch, 1, DQMC_SWEEP_DYNAMIC | DQMC_SWEEP_MEASURE | DQMC_SWEEP_RECORD),
                           dqmc_detail::context(), "DqmcWalker::sweep_for_dynamic_greens");
    }
    // accessors (include/dqmc_walker.h:136-164).  GreenttUp()/GreenttDn() give the resident
    // matrices by reference (a device handle; the const forms copy them to the host); the
    // per-slot ones copy the slots kept by the last sweep.
    GreensFunc& GreenttUp() { return m_green_tt_up; }
    GreensFunc& GreenttDn() { return m_green_tt_dn; }
    std::vector<double> GreenttUp() const { return fetch(-1, +1); }
    std::vector<double> GreenttDn() const { return fetch(-1, -1); }
    std::vector<double> GreenttUp(int t) const { return fetch(t, +1); }
    std::vector<double> GreenttDn(int t) const { return fetch(t, -1); }
    // DqmcWalker::vecGreenttUp/Dn (include/dqmc_walker.h:141-142): G of every slot of the last sweep
    std::vector<std::vector<double>> vecGreenttUp() const { return fetch_all(+1); }
    std::vector<std::vector<double>> vecGreenttDn() const { return fetch_all(-1); }
    // time-displaced G(t,0), G(0,t) of slot t (t+1 slices from 0) of the last dynamic sweep
    std::vector<double> Greent0Up(int t) const { return fetch_dynamic(t, +1, 0); }
    std::vector<double> Greent0Dn(int t) const { return fetch_dynamic(t, -1, 0); }
    std::vector<double> Green0tUp(int t) const { return fetch_dynamic(t, +1, 1); }
    std::vector<double> Green0tDn(int t) const { return fetch_dynamic(t, -1, 1); }
    double ConfigSign() const {
        dqmc_chain_info i{};
        dqmc_get_info(m_batch, 0, &i);
        return i.sign;
    }
    std::vector<double> vecConfigSign() const {
        std::vector<double> s(m_time_size);
        dqmc_detail::check(dqmc_get_slice_signs(m_batch, 0, s.data()), dqmc_detail::context(), "vecConfigSign");
        return s;
    }
    double WrapError() const {
        dqmc_chain_info i{};
        dqmc_get_info(m_batch, 0, &i);
        return i.wrap_error;
    }
};
}  // namespace QuantumMonteCarlo

namespace Model {
inline void ModelBase::initial(const Lattice::LatticeBase& lattice, const QuantumMonteCarlo::DqmcWalker& walker) {
    initial_params(lattice, walker);
    initial_KV_matrices(lattice, walker);
}
inline void ModelBase::initial_params(const Lattice::LatticeBase& lattice, const QuantumMonteCarlo::DqmcWalker& walker) {
    m_side = lattice.SideLength();
    m_space_size = lattice.SpaceSize();
    m_time_size = walker.TimeSize();
    m_beta = walker.Beta();
    m_alpha = std::acosh(std::exp(0.5 * walker.TimeInterval() * std::fabs(m_onsite_u)));
}
inline void ModelBase::initial_KV_matrices(const Lattice::LatticeBase&, const QuantumMonteCarlo::DqmcWalker& walker) {
    m_expK.clear();
    m_inv_expK.clear();
    if (m_checkerboard) return;   // the break-up has no matrices (CheckerBoard::Square::initial)
    dqmc_ctx_t ctx = dqmc_detail::context();
    const dqmc_chain_params p = chain_params(walker.Beta(), walker.TimeSize());
    dqmc_workspace_t w = nullptr;
    dqmc_detail::check(dqmc_workspace_create(ctx, &p, &w), ctx, "initial_KV_matrices");
    m_expK.resize((size_t)m_space_size * m_space_size);
    m_inv_expK.resize(m_expK.size());
    const int rc = dqmc_workspace_expk(w, m_expK.data(), m_inv_expK.data());
    dqmc_workspace_destroy(w);
    dqmc_detail::check(rc, ctx, "initial_KV_matrices");
}
inline double ModelBase::get_update_ratio(QuantumMonteCarlo::DqmcWalker& walker, int t, int i) const {
    double r = 0.0;
    dqmc_detail::check(dqmc_chain_update_ratio(walker.m_batch, 0, t, i, &r), dqmc_detail::context(), "get_update_ratio");
    return r;
}
inline void ModelBase::update_greens_function(QuantumMonteCarlo::DqmcWalker& walker, int t, int i) {
    dqmc_detail::check(dqmc_chain_update_green(walker.m_batch, 0, t, i), dqmc_detail::context(), "update_greens_function");
}
inline void ModelBase::update_bosonic_field(int t, int i) {
    if (!m_batch) throw std::logic_error("update_bosonic_field: chain not initialized");
    dqmc_detail::check(dqmc_chain_flip_field(m_batch, 0, t, i), dqmc_detail::context(), "update_bosonic_field");
    int8_t& f = m_bosonic_field[(size_t)t * m_space_size + i];
    f = (int8_t)(-f);
}
}  // namespace Model

namespace Measure {
using ObsList = std::vector<std::string>;

// names of the equal-time observables (src/measure/observable_handler.cpp:11-20) in engine order
inline const std::vector<std::string>& EqualTimeNames() {
    static const std::vector<std::string> n = {"equaltime_sign", "filling_number", "double_occupancy",
                                               "kinetic_energy", "momentum_distribution", "local_spin_corr",
                                               "spin_density_structure_factor", "charge_density_structure_factor",
                                               "s_wave_pairing_corr"};
    return n;
}
// names of the dynamic observables (src/measure/observable_handler.cpp:23-29)
inline const std::vector<std::string>& DynamicNames() {
    static const std::vector<std::string> n = {"greens_functions", "density_of_states", "superfluid_stiffness",
                                               "dynamic_spin_susceptibility"};
    return n;
}

class MeasureHandler {
    int m_sweeps_warmup{}, m_bin_num{}, m_bin_size{}, m_sweeps_between_bins{};
    ObsList m_obs_list{};
    int m_momentum{-1};
    std::vector<int> m_momentum_list{};
    bool m_is_equaltime{false}, m_is_dynamic{false};
    dqmc_batch_t m_batch{nullptr};
    int m_time_size{};
    std::vector<std::vector<double>> m_bins;
    std::vector<double> m_mean, m_err;
    // dynamic bins: |sign|, rho_s, G(k,t) (k fastest), D(t), DSS(t)
    std::vector<std::vector<double>> m_dyn_bins;
    std::vector<double> m_dyn_mean, m_dyn_err;
    friend class QuantumMonteCarlo::DqmcInitializer;

   public:
    void set_measure_params(int sweeps_warmup, int bin_num, int bin_size, int sweeps_between_bins) {
        if (sweeps_warmup < 0 || bin_num < 0 || bin_size < 0) throw std::invalid_argument("negative measuring params");
        m_sweeps_warmup = sweeps_warmup;
        m_bin_num = bin_num;
        m_bin_size = bin_size;
        m_sweeps_between_bins = sweeps_between_bins;
    }
    void set_observables(ObsList obs) {
        m_is_equaltime = m_is_dynamic = false;
        for (const auto& o : obs) {
            bool eq = false, dyn = false;
            for (const auto& n : EqualTimeNames()) eq = eq || (o == n);
            for (const auto& n : DynamicNames()) dyn = dyn || (o == n);
            if (!eq && !dyn) throw std::invalid_argument("unsupported observable '" + o + "'");
            m_is_equaltime = m_is_equaltime || eq;
            m_is_dynamic = m_is_dynamic || dyn;
        }
        m_obs_list = std::move(obs);
    }
    void set_measured_momentum(int momentum_index) { m_momentum = momentum_index; }
    // momentum indices (k-stars) of G(k,t) (MeasureHandler::set_measured_momentum_list)
    void set_measured_momentum_list(std::vector<int> k) { m_momentum_list = std::move(k); }
    const std::vector<int>& MomentumList() const { return m_momentum_list; }
    bool isWarmUp() const { return m_sweeps_warmup != 0; }
    bool isEqualTime() const { return m_is_equaltime; }
    bool isDynamic() const { return m_is_dynamic; }
    int WarmUpSweeps() const { return m_sweeps_warmup; }
    int SweepsBetweenBins() const { return m_sweeps_between_bins; }
    int BinsNum() const { return m_bin_num; }
    int BinsSize() const { return m_bin_size; }
    int Momentum() const { return m_momentum; }

    // src/measure/measure_handler.cpp:136-150: every method over the L stored slices
    void equaltime_measure(const QuantumMonteCarlo::DqmcWalker&, const Model::ModelBase&, const Lattice::LatticeBase&) {
        dqmc_detail::check(dqmc_measure_recorded(m_batch), dqmc_detail::context(), "equaltime_measure");
    }
    // src/measure/measure_handler.cpp:153-167: the engine has accumulated these during the
    // time-displaced sweep (DqmcWalker::sweep_for_dynamic_greens)
    void dynamic_measure(const QuantumMonteCarlo::DqmcWalker&, const Model::ModelBase&, const Lattice::LatticeBase&) {}
    // :170-240 — normalize by counts and <sign>, keep |<sign>|, store the bin
    void normalize_stats() {
        if (m_is_equaltime) {
            double sums[DQMC_NUM_OBSERVABLES], counts[DQMC_NUM_OBSERVABLES];
            dqmc_detail::check(dqmc_get_measure(m_batch, 0, sums, counts), dqmc_detail::context(), "normalize_stats");
            std::vector<double> v(DQMC_NUM_OBSERVABLES);
            const double sg = sums[0] / counts[0];
            for (int o = 1; o < DQMC_NUM_OBSERVABLES; ++o) v[o] = sums[o] / (counts[o] * sg);
            v[0] = std::fabs(sg);
            m_bins.push_back(v);
        }
        if (m_is_dynamic) {
            const int L = m_time_size, nk = (int)m_momentum_list.size();
            std::vector<double> gk((size_t)nk * L), dos(L), dss(L), sc(3);
            dqmc_detail::check(dqmc_get_dynamic_measure(m_batch, 0, nk, m_momentum_list.data(), gk.data(), dos.data(),
                                                        dss.data(), sc.data()),
                               dqmc_detail::context(), "normalize_stats");
            const double sg = sc[1] / sc[2], den = sc[2] * sg;
            std::vector<double> w = {std::fabs(sg), sc[0] / den};
            for (double x : gk) w.push_back(x / den);
            for (double x : dos) w.push_back(x / den);
            for (double x : dss) w.push_back(x / den);
            m_dyn_bins.push_back(w);
        }
    }
    void write_stats_to_bins(int) {}
    void clear_temporary() { dqmc_detail::check(dqmc_clear_measure(m_batch), dqmc_detail::context(), "clear_temporary"); }
    // include/measure/observable.h:169-193
    void analyse_stats() {
        auto stats = [](const std::vector<std::vector<double>>& bins, std::vector<double>& mean, std::vector<double>& err) {
            const size_t nb = bins.size(), nv = nb ? bins[0].size() : 0;
            mean.assign(nv, 0.0);
            err.assign(nv, 0.0);
            for (size_t o = 0; o < nv; ++o) {
                double m = 0.0, m2 = 0.0;
                for (auto& b : bins) { m += b[o]; m2 += b[o] * b[o]; }
                m /= nb; m2 /= nb;
                mean[o] = m;
                err[o] = nb > 1 ? std::sqrt(std::max(0.0, m2 - m * m)) / std::sqrt((double)nb - 1) : 0.0;
            }
        };
        stats(m_bins, m_mean, m_err);
        stats(m_dyn_bins, m_dyn_mean, m_dyn_err);
    }
    double mean(const std::string& name) const { return m_mean.at(index(name)); }
    double error(const std::string& name) const { return m_err.at(index(name)); }
    // dynamic results: "dynamic_sign" and "superfluid_stiffness" one value, "greens_functions"
    // n_k x L (k fastest), "density_of_states" and "dynamic_spin_susceptibility" L values
    std::vector<double> dynamic_mean(const std::string& name) const { return slice(m_dyn_mean, name); }
    std::vector<double> dynamic_error(const std::string& name) const { return slice(m_dyn_err, name); }
    std::vector<double> slice(const std::vector<double>& v, const std::string& name) const {
        const size_t L = m_time_size, nk = m_momentum_list.size();
        size_t o = 0, n = 1;
        if (name == "dynamic_sign") o = 0;
        else if (name == "superfluid_stiffness") o = 1;
        else if (name == "greens_functions") { o = 2; n = nk * L; }
        else if (name == "density_of_states") { o = 2 + nk * L; n = L; }
        else if (name == "dynamic_spin_susceptibility") { o = 2 + nk * L + L; n = L; }
        else throw std::invalid_argument("unknown dynamic observable " + name);
        if (v.size() < o + n) throw std::logic_error("no dynamic statistics");
        return std::vector<double>(v.begin() + o, v.begin() + o + n);
    }
    // bins of one observable, bin-major (the element order of dynamic_mean for dynamic ones)
    std::vector<double> bins(const std::string& name) const {
        std::vector<double> out;
        bool dyn = false;
        for (const auto& n : DynamicNames()) dyn = dyn || n == name;
        if (dyn) {
            for (const auto& b : m_dyn_bins) {
                const auto v = slice(b, name);
                out.insert(out.end(), v.begin(), v.end());
            }
        } else {
            const int o = index(name);
            for (const auto& b : m_bins) out.push_back(b[o]);
        }
        return out;
    }
    int BinsCount(const std::string& name) const {
        for (const auto& n : DynamicNames())
            if (n == name) return (int)m_dyn_bins.size();
        return (int)m_bins.size();
    }
    int TimeSize() const { return m_time_size; }
    static int index(const std::string& name) {
        const auto& n = EqualTimeNames();
        for (size_t k = 0; k < n.size(); ++k)
            if (n[k] == name) return (int)k;
        throw std::invalid_argument("unknown observable " + name);
    }
};
}  // namespace Measure

namespace QuantumMonteCarlo {

class DqmcInitializer {
   public:
    // src/dqmc_initializer.cpp:341-392
    static void initial_modules(Lattice::LatticeBase& lattice, Model::ModelBase& model, DqmcWalker& walker,
                                Measure::MeasureHandler& meas, CheckerBoard::CheckerBoardBase* checkerboard = nullptr) {
        if (!lattice.InitialStatus()) lattice.initial();
        walker.m_space_size = lattice.SpaceSize();
        walker.m_is_equaltime = meas.isEqualTime();
        walker.m_is_dynamic = meas.isDynamic();
        meas.m_time_size = walker.TimeSize();
        model.m_space_size = lattice.SpaceSize();
        model.m_time_size = walker.TimeSize();
        model.m_side = lattice.SideLength();
        model.m_beta = walker.Beta();
        model.m_bosonic_field.assign((size_t)walker.TimeSize() * lattice.SpaceSize(), 1);
        if (checkerboard) model.link(*checkerboard); else model.link();
        model.initial(lattice, walker);   // src/dqmc_initializer.cpp:361
    }
    // src/dqmc_initializer.cpp:395-408: stacks, G(0), sign — the chain moves to the device here
    static void initial_dqmc(Lattice::LatticeBase& lattice, Model::ModelBase& model, DqmcWalker& walker,
                             Measure::MeasureHandler& meas) {
        dqmc_ctx_t ctx = dqmc_detail::context();
        dqmc_chain_params p{};
        p.model = model.m_model_type;
        p.side = lattice.SideLength();
        p.time_size = walker.TimeSize();
        p.stabilization_pace = walker.StabilizationPace();
        p.checkerboard = model.m_checkerboard ? 1 : 0;
        p.delay = walker.m_delay;
        p.momentum_index = meas.Momentum();
        p.precision = DQMC_PREC_FP64;
        p.beta = walker.Beta();
        p.hopping_t = model.HoppingT();
        p.onsite_u = model.OnSiteU();
        p.chemical_potential = model.ChemicalPotential();
        p.seed = 1;
        if (walker.m_batch) dqmc_batch_destroy(walker.m_batch);
        walker.m_batch = nullptr;
        dqmc_detail::check(dqmc_batch_create(ctx, 1, &p, (walker.m_is_equaltime || walker.m_is_dynamic) ? 1 : 0,
                                             &walker.m_batch),
                           ctx, "dqmc_batch_create");
        dqmc_detail::check(dqmc_set_fields(walker.m_batch, 0, model.m_bosonic_field.data()), ctx, "dqmc_set_fields");
        // the device continues the global engine's stream (Metropolis draws follow the field draws)
        dqmc_detail::check(dqmc_set_rng_state(walker.m_batch, 0, Utils::Random::state()), ctx, "dqmc_set_rng_state");
        dqmc_detail::check(dqmc_init(walker.m_batch), ctx, "dqmc_init");
        meas.m_batch = walker.m_batch;
        model.m_batch = walker.m_batch;
        for (int s = 0; s < 2; ++s) {
            GreensFunc& g = s == 0 ? walker.m_green_tt_up : walker.m_green_tt_dn;
            g.m_batch = walker.m_batch;
            g.m_spin = s == 0 ? +1 : -1;
            g.m_dim = lattice.SpaceSize();
            g.m_mirror = s == 1 && p.model == DQMC_MODEL_ATTRACTIVE;
        }
    }
};

class Dqmc {
   public:
    // src/dqmc.cpp:48-70
    static void sweep_forth_and_back(DqmcWalker& walker, Model::ModelBase& model, Lattice::LatticeBase& lattice,
                                     Measure::MeasureHandler& meas) {
        if (meas.isDynamic()) {
            walker.sweep_for_dynamic_greens(model);
            meas.dynamic_measure(walker, model, lattice);
        } else {
            walker.sweep_from_0_to_beta(model);
            if (meas.isEqualTime()) meas.equaltime_measure(walker, model, lattice);
        }
        walker.sweep_from_beta_to_0(model);
        if (meas.isEqualTime()) meas.equaltime_measure(walker, model, lattice);
    }
    // src/dqmc.cpp:73-105
    static void thermalize(DqmcWalker& walker, Model::ModelBase& model, Lattice::LatticeBase&,
                           Measure::MeasureHandler& meas) {
        for (int s = 1; s <= meas.WarmUpSweeps() / 2; ++s) {
            walker.sweep_from_0_to_beta(model);
            walker.sweep_from_beta_to_0(model);
        }
    }
    // src/dqmc.cpp:108-151
    static void measure(DqmcWalker& walker, Model::ModelBase& model, Lattice::LatticeBase& lattice,
                        Measure::MeasureHandler& meas) {
        if (!meas.isEqualTime() && !meas.isDynamic()) return;
        for (int bin = 0; bin < meas.BinsNum(); ++bin) {
            for (int s = 1; s <= meas.BinsSize() / 2; ++s) sweep_forth_and_back(walker, model, lattice, meas);
            meas.normalize_stats();
            meas.write_stats_to_bins(bin);
            meas.clear_temporary();
            for (int s = 0; s < meas.SweepsBetweenBins() / 2; ++s) {
                walker.sweep_from_0_to_beta(model);
                walker.sweep_from_beta_to_0(model);
            }
        }
    }
    static void analyse(Measure::MeasureHandler& meas) { meas.analyse_stats(); }
};

// QuantumMonteCarlo::DqmcIO (include/dqmc_io.h:267-629): the file outputs of the reference's
// main (src/dqmc_main.cpp:270-360) and its field file, written by the engine library's host code
// in the reference's formats.  Path-based: the reference takes an open std::ofstream.
class DqmcIO {
    static void io(int rc, const std::string& what, const std::string& path) {
        if (rc != DQMC_OK) throw std::runtime_error(what + "('" + path + "') failed with status " + std::to_string(rc));
    }

   public:
    // the current fields: from the device once the chain is initialized
    static void output_bosonic_fields(const std::string& path, const Model::ModelBase& model) {
        std::vector<int8_t> f = model.m_bosonic_field;
        if (model.m_batch)
            dqmc_detail::check(dqmc_get_fields(model.m_batch, 0, f.data()), dqmc_detail::context(), "dqmc_get_fields");
        io(dqmc_write_fields_file(path.c_str(), model.m_time_size, model.m_space_size, f.data()), "output_bosonic_fields",
           path);
    }
    // before DqmcInitializer::initial_dqmc, as src/dqmc_main.cpp:148-160 does
    static void read_bosonic_fields_from_file(const std::string& path, Model::ModelBase& model) {
        io(dqmc_read_fields_file(path.c_str(), model.m_time_size, model.m_space_size, model.m_bosonic_field.data()),
           "read_bosonic_fields_from_file", path);
    }
    static void output_k_stars(const std::string& path, const Lattice::LatticeBase& lattice) {
        io(dqmc_write_kstars_file(path.c_str(), lattice.SideLength()), "output_k_stars", path);
    }
    static void output_imaginary_time_grids(const std::string& path, const DqmcWalker& walker) {
        io(dqmc_write_tgrids_file(path.c_str(), walker.TimeSize(), walker.Beta()), "output_imaginary_time_grids", path);
    }
    // rank of each observable as the reference stores it (scalar, vector over t, matrix k x t)
    static void shape(const Measure::MeasureHandler& meas, const std::string& name, int& rank, int& rows, int& cols) {
        rank = 0; rows = 1; cols = 1;
        if (name == "density_of_states" || name == "dynamic_spin_susceptibility") { rank = 1; rows = meas.TimeSize(); }
        else if (name == "greens_functions") {
            rank = 2; rows = (int)meas.MomentumList().size(); cols = meas.TimeSize();
        }
    }
    static void output_observable(const std::string& path, const Measure::MeasureHandler& meas, const std::string& name) {
        int rank, rows, cols;
        shape(meas, name, rank, rows, cols);
        std::vector<double> m, e;
        if (rank == 0 && name != "superfluid_stiffness") { m = {meas.mean(name)}; e = {meas.error(name)}; }
        else { m = meas.dynamic_mean(name); e = meas.dynamic_error(name); }
        io(dqmc_write_observable_file(path.c_str(), rank, rows, cols, m.data(), e.data()), "output_observable", path);
    }
    static void output_observable_in_bins(const std::string& path, const Measure::MeasureHandler& meas,
                                          const std::string& name) {
        int rank, rows, cols;
        shape(meas, name, rank, rows, cols);
        const std::vector<double> b = meas.bins(name);
        io(dqmc_write_observable_bins_file(path.c_str(), rank, meas.BinsCount(name), rows, cols, b.data()),
           "output_observable_in_bins", path);
    }
};

}  // namespace QuantumMonteCarlo
