// Host-side WRAPD classes with the reference's names and call structure
// (src/Settings.hpp, src/Lame.hpp, src/ModelSettings.hpp, src/RuntimeData.hpp,
// src/Elements.hpp, src/TetElements.hpp, src/System.hpp, src/AlgorithmData.hpp,
// src/GlobalStepBehavior.hpp, src/ADMM.hpp, src/Solver.hpp).  The element data and
// all per-element / per-vertex arithmetic live on the GPU behind include/wrapd_gpu.h;
// these classes keep the ADMM control, the logs and the timers.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/wrapd_gpu.h"
#include "host_error.hpp"

namespace wrapd {
namespace math {
// positions as (n,3) row-major doubles (the reference uses Eigen::MatrixX3d)
struct MatX3 {
    int rows_ = 0;
    std::vector<double> v;
    MatX3() = default;
    explicit MatX3(int r) : rows_(r), v(3ull * r, 0.0) {}
    int rows() const { return rows_; }
    double& operator()(int r, int c) { return v[3ull * r + c]; }
    double operator()(int r, int c) const { return v[3ull * r + c]; }
    double* data() { return v.data(); }
    const double* data() const { return v.data(); }
};
struct Vec3 { double x[3]; };
struct Vec4i { int x[4]; };
using VectorXd = std::vector<double>;

// src/Math.hpp:47-53: Triplet / Triplets / SpMat (a row-major sparse matrix; here compressed
// rows with sorted columns, duplicates summed by setFromTriplets)
struct Triplet {
    int r = 0, c = 0;
    double v = 0.0;
    Triplet() = default;
    Triplet(int row, int col, double value) : r(row), c(col), v(value) {}
    int row() const { return r; }
    int col() const { return c; }
    double value() const { return v; }
};
using Triplets = std::vector<Triplet>;

class SpMat {
 public:
    SpMat() = default;
    SpMat(int rows, int cols) { resize(rows, cols); }
    int rows() const { return m_rows; }
    int cols() const { return m_cols; }
    long long nonZeros() const { return static_cast<long long>(m_inner.size()); }
    void resize(int rows, int cols) {
        m_rows = rows;
        m_cols = cols;
        m_outer.assign(static_cast<size_t>(rows) + 1, 0);
        m_inner.clear();
        m_values.clear();
    }
    template <typename It>
    void setFromTriplets(It begin, It end) {
        std::vector<Triplet> t(begin, end);
        std::stable_sort(t.begin(), t.end(), [](const Triplet& a, const Triplet& b) {
            return a.r < b.r || (a.r == b.r && a.c < b.c);
        });
        resize(m_rows, m_cols);
        for (size_t k = 0; k < t.size(); ++k) {
            if (t[k].r < 0 || t[k].r >= m_rows || t[k].c < 0 || t[k].c >= m_cols)
                throw std::runtime_error("SpMat::setFromTriplets: index out of range");
            if (k > 0 && t[k].r == t[k - 1].r && t[k].c == t[k - 1].c) {
                m_values.back() += t[k].v;
                continue;
            }
            m_inner.push_back(t[k].c);
            m_values.push_back(t[k].v);
            m_outer[static_cast<size_t>(t[k].r) + 1]++;
        }
        for (int i = 0; i < m_rows; ++i) m_outer[i + 1] += m_outer[i];
    }
    // the stored entry (i, j); throws if the pattern has no such entry
    double& coeffRef(int i, int j) {
        const int* b = m_inner.data() + m_outer[i];
        const int* e = m_inner.data() + m_outer[i + 1];
        const int* p = std::lower_bound(b, e, j);
        if (p == e || *p != j) throw std::runtime_error("SpMat::coeffRef: entry not in the pattern");
        return m_values[static_cast<size_t>(p - m_inner.data())];
    }
    double coeff(int i, int j) const {
        const int* b = m_inner.data() + m_outer[i];
        const int* e = m_inner.data() + m_outer[i + 1];
        const int* p = std::lower_bound(b, e, j);
        return (p == e || *p != j) ? 0.0 : m_values[static_cast<size_t>(p - m_inner.data())];
    }
    void makeCompressed() {}
    void setZero() { std::fill(m_values.begin(), m_values.end(), 0.0); }
    const std::vector<int>& outerIndex() const { return m_outer; }   // row pointers (rows + 1)
    const std::vector<int>& innerIndex() const { return m_inner; }   // column of each entry
    const std::vector<double>& valuesVec() const { return m_values; }
    std::vector<double>& valuesVec() { return m_values; }
 private:
    int m_rows = 0, m_cols = 0;
    std::vector<int> m_outer{0}, m_inner;
    std::vector<double> m_values;
};
}  // namespace math

// src/Lame.hpp:36-58 (Young's modulus is a parameter here)
class Lame {
 public:
    static Lame preset(double poisson, double youngs = 1e7) { return Lame(youngs, poisson); }
    Lame(double k, double v) : m_mu(k / (2.0 * (1.0 + v))), m_lambda(k * v / ((1.0 + v) * (1.0 - 2.0 * v))) {}
    double mu() const { return m_mu; }
    double lambda() const { return m_lambda; }
    void mu(double v) { m_mu = v; }
    void lambda(double v) { m_lambda = v; }
 private:
    double m_mu, m_lambda;
};

// src/ModelSettings.hpp:31-95
class ModelSettings {
 public:
    enum class ElasticModel { ARAP, NH };
    ElasticModel elastic_model() const { return m_model; }
    void elastic_model(const std::string& v) {
        if (v == "arap") m_model = ElasticModel::ARAP;
        else if (v == "nh") m_model = ElasticModel::NH;
        else throw std::runtime_error("Error: Invalid input for -elastic_model");
    }
    void poisson(double v) { m_poisson = v; m_lame = Lame::preset(v, m_youngs); }
    void youngs(double v) { m_youngs = v; m_lame = Lame::preset(m_poisson, v); }
    double poisson() const { return m_poisson; }
    double youngs() const { return m_youngs; }
    const Lame& elastic_lame() const { return m_lame; }
    double elastic_beta_static() const { return m_beta_static; }
    void elastic_beta_static(double v) { m_beta_static = v; }
    double elastic_beta_min() const { return m_beta_min; }
    void elastic_beta_min(double v) { m_beta_min = v; }
    double elastic_beta_max() const { return m_beta_max; }
    void elastic_beta_max(double v) { m_beta_max = v; }
    int elastic_prox_iters() const { return m_prox_iters; }
    void elastic_prox_iters(int v) { m_prox_iters = v; }
 private:
    ElasticModel m_model = ElasticModel::NH;
    double m_poisson = 0.45, m_youngs = 1e7;
    Lame m_lame = Lame::preset(0.45, 1e7);
    double m_beta_static = 1.0, m_beta_min = 0.1, m_beta_max = 10.0;
    int m_prox_iters = 20;
};

// src/Settings.hpp:34-163 plus the device-path options
struct Settings {
    bool parse_args(int argc, char** argv);   // true if -help was given
    static std::string help_text();

    class IO {
     public:
        std::string rest_mesh() const { return m_rest; }
        void rest_mesh(const std::string& v) { m_rest = strip_extension(v); }
        std::string init_mesh() const { return m_init; }
        void init_mesh(const std::string& v) { m_init = strip_extension(v); }
        std::string soln_mesh() const { return m_soln; }
        void soln_mesh(const std::string& v) { m_soln = strip_extension(v); m_soln_provided = true; }
        std::string handles() const { return m_handles; }
        void handles(const std::string& v) { m_handles = v; }
        std::string output_dir() const { return m_outdir; }
        void output_dir(const std::string& v) { m_outdir = v; }
        int verbose() const { return m_verbose; }
        void verbose(int v) { m_verbose = v; }
        bool should_save_data() const { return m_save; }
        void should_save_data(bool v) { m_save = v; }
        bool soln_provided() const { return m_soln_provided; }
        static std::string strip_extension(const std::string& raw);
     private:
        std::string m_rest, m_init, m_soln, m_handles, m_outdir = "output";
        int m_verbose = 1;
        bool m_save = true;
        bool m_soln_provided = false;
    };
    IO m_io;

    double m_tol = 1.e-11;
    int m_max_admm_iters = 5000;
    double m_reweighting_delay = 50;
    enum class RotAwareness { ENABLED, DISABLED };
    RotAwareness m_rot_awareness = RotAwareness::ENABLED;
    enum class Reweighting { DISABLED, ENABLED };
    Reweighting m_reweighting = Reweighting::ENABLED;
    double m_gamma = 1.5;

    class GStep {
     public:
        enum class LineSearch { None, Backtracking };
        int max_iters() const { return m_max_iters; }
        void max_iters(int v) { m_max_iters = v; }
        LineSearch linesearch() const { return m_ls; }
        void linesearch(LineSearch v) { m_ls = v; }
        double kappa() const { return m_kappa; }
        void kappa(double v) { m_kappa = v; }
        int memory() const { return m_memory; }
        void memory(int v) { m_memory = v; }
     private:
        int m_max_iters = 100;
        LineSearch m_ls = LineSearch::None;
        double m_kappa = 1.0;
        int m_memory = 6;     // SplitLBFGS<...,6> (deps/mcloptlib/include/MCL/SplitLBFGS.hpp:41)
    };
    GStep m_gstep;
    ModelSettings m_model_settings;

    // device path
    int m_device = 0;
    int m_precision = WRAPD_FP64;
    int m_prox_mode = WRAPD_PROX_LBFGS;
    double m_newton_rtol = 0.0;      // 0: 1e-12 (FP64) / 1e-6 (FP32)
    double m_fallback_sigma = 0.2;
    double m_fp64_switch = 1e-8;     // mixed precision: FP64 local steps below this residual
    std::vector<double> m_mu_e, m_lambda_e;   // optional per-element material
};

// src/RuntimeData.hpp:30-44
// Flat byte streams for checkpoints (native layout, read back by the same build).
struct ByteWriter {
    std::vector<char> buf;
    template <typename T>
    void pod(const T& v) {
        const char* p = reinterpret_cast<const char*>(&v);
        buf.insert(buf.end(), p, p + sizeof(T));
    }
    template <typename T>
    void vec(const std::vector<T>& v) {
        pod(static_cast<unsigned long long>(v.size()));
        const char* p = reinterpret_cast<const char*>(v.data());
        buf.insert(buf.end(), p, p + v.size() * sizeof(T));
    }
};
struct ByteReader {
    const char* p = nullptr;
    const char* end = nullptr;
    template <typename T>
    void pod(T& v) {
        need(sizeof(T));
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
    }
    template <typename T>
    void vec(std::vector<T>& v) {
        unsigned long long n = 0;
        pod(n);
        need(n * sizeof(T));
        v.resize(n);
        std::memcpy(v.data(), p, n * sizeof(T));
        p += n * sizeof(T);
    }
    void need(size_t n) const {
        if (static_cast<size_t>(end - p) < n) throw std::runtime_error("checkpoint: truncated");
    }
};

struct RuntimeData {
    double setup_solve_ms = 0, local_ms = 0, global_ms = 0, refactor_ms = 0;
    int inner_iters = 0;
    double finish_solve_ms = 0;
    // device-path extras
    double prox_iters = 0, fallback_elems = 0, learned_accepted = 0, learned_used = 0;
    double fp64_switch_iter = -1;   // first iteration of the FP64 tail of a mixed run
};

class MicroTimer {
 public:
    void reset() { t0 = std::chrono::high_resolution_clock::now(); }
    double elapsed_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
    }
 private:
    std::chrono::high_resolution_clock::time_point t0 = std::chrono::high_resolution_clock::now();
};

// src/Elements.hpp:36-157: the reference's element interface.  The element data of the one
// implementation, TetElements, lives on the GPU, so the base class keeps only the count;
// every operation an implementation does not provide throws, as the reference's do.
class Elements {
 public:
    explicit Elements(int num_elements) : m_num_elements(num_elements) {}
    virtual ~Elements() {}
    int num_elements() const { return m_num_elements; }
    virtual int is_inverted(const math::MatX3& x_free, const math::MatX3& x_fix) {
        (void)x_free;
        (void)x_fix;
        return 0;
    }
    // src/Elements.hpp:56-69 (max over elements of w^2/cand and its reciprocal)
    virtual double max_wsq_ratio() { throw std::runtime_error("Error: Need to implement max_wsq_ratio"); }
    virtual void update_actual_weights() { throw std::runtime_error("Error: Need to implement update_actual_weights"); }
    virtual void initialize_dual_vars(const math::MatX3& X) {
        (void)X;
        throw std::runtime_error("Error: Need to implement initialize_dual_vars");
    }
    virtual void advance_dual_vars() { throw std::runtime_error("Error: Need to implement advance_dual_vars"); }
    virtual void midupdate_weights(const math::MatX3& X) {
        (void)X;
        throw std::runtime_error("Error: Need to implement midupdate_weights(X)");
    }
    virtual void midupdate_weights() { throw std::runtime_error("Error: Need to implement midupdate_weights()"); }
    virtual void update(bool update_candidate_weights) {
        (void)update_candidate_weights;
        throw std::runtime_error("Error: Need to implement EnergyTerm::update(MatX3, bool)");
    }
    virtual double polar_update(bool update_candidate_weights) {
        (void)update_candidate_weights;
        throw std::runtime_error("Error: Implement polar_update(bool)");
    }
    virtual math::MatX3 add_to_b() { throw std::runtime_error("Error: Need to implement add_to_b()"); }
    virtual void get_A_triplets(math::Triplets& triplets) {
        (void)triplets;
        throw std::runtime_error("Error: Need to implement get_A_triplets(triplets)");
    }
    virtual void update_A_coeffs(math::SpMat& A) {
        (void)A;
        throw std::runtime_error("Error: Need to implement update_A_coeffs(A)");
    }
    virtual double penalty_energy(bool lagged_u) const {
        (void)lagged_u;
        throw std::runtime_error("Error: Need to implement penalty_energy(lagged_u)");
    }
    virtual double gradient(const math::MatX3& Ft, math::MatX3& grad) {
        (void)Ft;
        (void)grad;
        throw std::runtime_error("Error: Need to implement gradient");
    }
    virtual double primal_residual_sq(Settings::RotAwareness rot_awareness) const {
        (void)rot_awareness;
        throw std::runtime_error("Error: Need to implement primal_residual_sq");
    }
    virtual double potential_energy_x() const { throw std::runtime_error("Error: Need to implement potential_energy_x"); }
    virtual double polar_potential_energy() const {
        throw std::runtime_error("Error: Need to implement polar_potential_energy");
    }

 protected:
    virtual double potential_energy(const math::MatX3& Ft) {
        (void)Ft;
        throw std::runtime_error("Error: Need to implement EnergyTerm::potential_energy(const MatX3 &Ft)");
    }
    const int m_num_elements;
};

// src/TetElements.hpp:165-247: the element set, resident on the GPU.
class TetElements : public Elements {
 public:
    // tets: indices into rest (n_v x 3); vertex_offset is added to the indices inds() reports
    // (create_tets_from_mesh, src/TetElements.hpp:285)
    TetElements(const std::vector<double>& rest, const std::vector<int>& tets, const std::vector<int>& pins,
                const Settings& settings, int vertex_offset = 0);
    ~TetElements() override;
    TetElements(const TetElements&) = delete;
    TetElements& operator=(const TetElements&) = delete;

    int num_free() const { return m_n_f; }
    wrapd_ctx* ctx() { return m_ctx; }
    const std::vector<math::Vec4i>& inds() const { return m_tet; }   // src/TetElements.hpp:207

    void update_fix_cache(const math::MatX3& x_fix);
    void set_x_free(const math::MatX3& x_free);
    void get_x_free(math::MatX3& x_free);
    double potential_energy_x() const override { return m_energy; }
    int inversion_count() const { return m_flips; }
    void update_defo_cache();                                  // energy + flips + symF at x
    void update_defo_cache(const math::MatX3& x_free);         // the same at the given x (it stays the current x)
    void initialize_dual_vars();                               // at the current x
    void initialize_dual_vars(const math::MatX3& X) override;  // at X (all vertices); the current x is kept
    // The device keeps no copies of the previous iterate's Z and U: the reference reads them
    // only in penalty_energy(true) and scales them with the weights in update_actual_weights.
    void advance_dual_vars() override {}
    double polar_update(bool update_candidate_weights) override;   // returns dAL of the local step
    void polar_update_launch(bool update_candidate_weights);   // polar_update in two halves
    double polar_update_finish();
    void update(bool update_candidate_weights) override;       // -ra 0
    void midupdate_weights(bool direct);
    void midupdate_weights(const math::MatX3& X) override;     // from F(X), X all vertices
    void midupdate_weights() override { midupdate_weights(false); }
    void update_actual_weights() override;
    double max_wsq_ratio() override;
    math::MatX3 add_to_b() override;                          // (n_free, 3)
    void get_A_triplets(math::Triplets& triplets) override;   // upper triangle, one triplet per entry
    void update_A_coeffs(math::SpMat& A) override;            // A(i, j) += entry, on A's pattern
    double penalty_energy(bool lagged_u) const override;      // lagged_u: not kept (throws)
    double primal_residual_sq(Settings::RotAwareness rot_awareness) const override;
    double polar_potential_energy() const override;
    double global_obj_value(const math::MatX3& x_free);       // src/TetElements.cpp:678-688 (x_free becomes current)
    double global_obj_grad(const math::MatX3& x_free, math::MatX3& grad);   // :691-738
    math::MatX3 global_obj_grad();                            // :742-787, at the current x
    void refactor();
    void global_step(double early_exit_tol, wrapd_global_out& out);
    void global_step_begin();     // queues the global step's first value and gradient
    void global_step_unaware(wrapd_global_out& out);
    void reset_history();
    double x_distance();
    void x_distance_launch();     // x_distance in two halves: the launch returns at once
    double x_distance_result();
    void set_x_ref(const math::MatX3& x);
    const wrapd_local_out& last_local() const { return m_last_local; }
    void set_fp64_tail(bool on) { m_fp64_tail = on; }
    // checkpoint: the element set's host state and the device state (wrapd_state_save)
    void save(ByteWriter& w);
    void load(ByteReader& r);   // then refactor() before the next global step

    double m_curr_pe = 0.0;
    int m_cand_count = 0;       // src/TetElements.cpp:404 (function-static counter in the reference)

 private:
    void check(int st, const char* what) const;
    std::vector<double> elem(const char* name, int width) const;   // device element array, row-major
    math::MatX3 all_to_free(const math::MatX3& X) const;
    const Settings m_settings;
    wrapd_ctx* m_ctx = nullptr;
    int m_n_e = 0, m_n_f = 0;
    std::vector<math::Vec4i> m_tet;   // vertex indices (+ vertex_offset)
    std::vector<int> m_free_rows;     // vertex of each free index
    double m_energy = 0.0;
    int m_flips = 0;
    wrapd_local_out m_last_local{};
    bool m_fp64_tail = false;   // mixed precision: FP64 local steps from here on
};

template <typename IN_SCALAR, typename TYPE>
inline void create_tets_from_mesh(std::shared_ptr<TYPE>& tet_elements, const IN_SCALAR* verts, const int* inds,
                                  int n_tets, const Settings& settings, const std::vector<int>& free_index_list,
                                  const std::vector<int>& fix_index_list, const int vertex_offset) {
    // src/TetElements.hpp:250-288; the GPU keeps whole-mesh arrays instead of per-tet copies
    // the vertices of the mesh are verts[0 .. n_v); the tets index them, and the stored
    // indices (inds()) are shifted by vertex_offset as the reference's are
    const int n_v = static_cast<int>(free_index_list.size());
    std::vector<double> rest(3ull * n_v);
    for (long long i = 0; i < 3ll * n_v; ++i) rest[i] = static_cast<double>(verts[i]);
    std::vector<int> tets(inds, inds + 4ll * n_tets);
    std::vector<int> pins;
    std::vector<std::pair<int, int>> fx;
    for (int v = 0; v < n_v; ++v) if (fix_index_list[v] >= 0) fx.emplace_back(fix_index_list[v], v);
    std::sort(fx.begin(), fx.end());
    for (auto& p : fx) pins.push_back(p.second);
    tet_elements = std::make_shared<TYPE>(rest, tets, pins, settings, vertex_offset);
}

class ConstraintSet {
 public:
    std::vector<int> m_pin_inds;
};

// src/System.hpp:41-174
class System {
 public:
    explicit System(const Settings& settings) : m_settings(settings), m_constraints(std::make_shared<ConstraintSet>()) {}
    int num_all_verts() const { return m_X.rows(); }
    const math::MatX3& X() const { return m_X; }
    void X(const math::MatX3& v) { m_X = v; }
    const math::MatX3& X_init() const { return m_X_init; }
    const math::MatX3& X_soln() const { return m_X_soln; }
    int add_nodes(const math::MatX3& X, const math::MatX3& X_init, const math::MatX3& X_soln);
    void set_pins(const std::vector<int>& inds, const std::vector<math::Vec3>& points);
    std::shared_ptr<ConstraintSet> constraint_set() { return m_constraints; }
    std::shared_ptr<TetElements>& tet_elements() { return m_tet_elements; }
    int num_tet_elements() const { return m_tet_elements ? m_tet_elements->num_elements() : 0; }
    bool possibly_update_weights(double gamma);
    double polar_local_update(bool update_candidate_weights) { return m_tet_elements->polar_update(update_candidate_weights); }
    void local_update(bool update_candidate_weights) { m_tet_elements->update(update_candidate_weights); }
    double problem_objective() const { return m_tet_elements->potential_energy_x(); }
    int inversion_count() const { return m_tet_elements->inversion_count(); }
    // src/System.cpp:66-100 and System.hpp:120-156
    math::MatX3 get_b() { return m_tet_elements->add_to_b(); }
    void get_A(math::SpMat& A);
    void update_A(math::SpMat& A);
    double penalty_energy(bool lagged_u) const { return m_tet_elements->penalty_energy(lagged_u); }
    double primal_residual_sq() const { return m_tet_elements->primal_residual_sq(m_settings.m_rot_awareness); }
    void midupdate_WtW(const math::MatX3& X) { m_tet_elements->midupdate_weights(X); }
    void midupdate_WtW() { m_tet_elements->midupdate_weights(); }
    double global_obj_value(const math::MatX3& x_free) { return m_tet_elements->global_obj_value(x_free); }
    double global_obj_grad(const math::MatX3& x_free, math::MatX3& grad, bool positions_have_changed) {
        if (positions_have_changed) return m_tet_elements->global_obj_grad(x_free, grad);
        grad = m_tet_elements->global_obj_grad();
        return 0.0;
    }
    void set_initial_pose() { m_X = m_X_init; }
    void initialize() {}
 private:
    const Settings m_settings;
    std::shared_ptr<TetElements> m_tet_elements;
    std::shared_ptr<ConstraintSet> m_constraints;
    math::MatX3 m_X, m_X_init, m_X_soln;
};

// src/AlgorithmData.hpp:37-146
class AlgorithmData {
 public:
    AlgorithmData(const Settings& s, std::shared_ptr<System> sys);
    void log_initial_data();
    void update_per_iter_data();
    void update_per_iter_runtime();
    void update_per_iter_reweighted(bool r) { m_reweighted.push_back(r ? 1 : 0); }
    void update_per_iter_delta_auglag(double ls, double gs) {
        m_per_iter_ls_delta_auglag.push_back(std::abs(ls));
        m_per_iter_gs_delta_auglag.push_back(std::abs(gs));
    }
    const std::vector<double>& per_iter_objectives() const { return m_objectives; }
    const std::vector<int>& per_iter_reweighted() const { return m_reweighted; }
    const std::vector<double>& per_iter_state_x_errors() const {
        resolve_x_error();
        return m_state_x_errors;
    }
    const std::vector<double>& per_iter_accumulated_time_s() const { return m_accumulated_time_s; }
    const std::vector<double>& per_iter_wall_time_s() const { return m_wall_time_s; }
    const std::vector<int>& per_iter_inner_iters() const { return m_per_iter_inner_iters; }
    const std::vector<int>& per_iter_flips() const { return m_flips_count; }
    double current_normalized_x_error() const {
        resolve_x_error();
        return m_state_x_errors[m_iter];
    }
    double current_normalized_objective_residual() const { return m_norm_obj_residuals[m_iter - 1]; }
    bool no_recent_flips() const { return m_flips_count[m_iter] == 0 && m_flips_count[m_iter - 1] == 0; }
    void print_initial_data() const;
    void print_curr_iter_data() const;
    void print_final_iter_data() const;
    void save(ByteWriter& w) const;   // the per-iteration state and logs (not the set-up lists)
    void load(ByteReader& r);
    int num_free_verts() const { return m_n_free; }
    int num_fix_verts() const { return static_cast<int>(m_fix_list.size()); }

    RuntimeData m_runtime;
    int m_iter = 0;
    double m_init_state_x_error = 1.0;
    double m_init_objective = 0.0;
    std::vector<double> m_norm_obj_residuals;
    std::vector<int> m_per_iter_inner_iters;
    std::vector<double> m_per_iter_ls_delta_auglag, m_per_iter_gs_delta_auglag;
    std::vector<int> m_free_list, m_fix_list;   // S_free / S_fix as index lists
    math::MatX3 m_curr_x_free, m_init_x_fix, m_curr_x_fix, m_soln_x_free;
    double m_init_auglag_obj = 0.0;
    double m_last_localstep_delta_auglag_obj = 0.0;
    MicroTimer m_wall;

 private:
    const Settings m_settings;
    std::shared_ptr<System> m_system;
    int m_n_free = 0;
    std::vector<double> m_objectives, m_accumulated_time_s, m_wall_time_s;
    std::vector<int> m_reweighted, m_flips_count;
    // Without a reference solution the x-error log only records, so each iteration's entry
    // is read back at the next iteration (or by its first reader): the host does not wait
    // for it before queuing the next iteration's kernels.
    mutable std::vector<double> m_state_x_errors;
    mutable int m_xerr_pending = -1;   // the entry still to be read, or -1
    void resolve_x_error() const;
};

// src/GlobalStepBehavior.hpp:76-268
class GlobalStepBehavior {
 public:
    GlobalStepBehavior(const Settings& s, std::shared_ptr<System> sys, std::shared_ptr<AlgorithmData> ad);
    virtual ~GlobalStepBehavior() {}
    virtual void global_step() = 0;
    virtual void begin_global_step() {}   // queue what global_step starts with, where that does not wait for the host
    virtual double last_step_delta_auglag() { return 1.e8; }
    virtual void initialize_weights();
    virtual bool update_weights();
    virtual void save(ByteWriter&) const {}
    virtual void load(ByteReader&) {}
 protected:
    const Settings m_settings;
    std::shared_ptr<System> m_system;
    std::shared_ptr<AlgorithmData> m_algorithm_data;
};

class UnawareGS : public GlobalStepBehavior {
 public:
    using GlobalStepBehavior::GlobalStepBehavior;
    void global_step() override;
};

class RotAwareGS : public GlobalStepBehavior {
 public:
    RotAwareGS(const Settings& s, std::shared_ptr<System> sys, std::shared_ptr<AlgorithmData> ad);
    void global_step() override;
    void begin_global_step() override;
    void initialize_weights() override;
    bool update_weights() override;
    double last_step_delta_auglag() override { return m_obj_aft - m_obj_bef; }
    void save(ByteWriter& w) const override { w.pod(m_obj_bef); w.pod(m_obj_aft); }
    void load(ByteReader& r) override { r.pod(m_obj_bef); r.pod(m_obj_aft); }
 private:
    double m_obj_bef = 0.0, m_obj_aft = 0.0;
};

// src/ADMM.hpp:39-61
class ADMM {
 public:
    bool initialize(std::shared_ptr<System> system, const Settings& settings);
    void solve(math::MatX3& X);
    AlgorithmData* algorithm_data() { return m_algorithm_data.get(); }
    // stepwise access (setup_solve, iterate, converged, finish_solve of the reference)
    void setup_solve(math::MatX3& X);
    void iterate();
    bool converged();
    void finish_solve(math::MatX3& X);
    // Checkpoint/resume (no counterpart in the reference): the whole ADMM state after the
    // last iterate(), to a file; resume into an ADMM set up (setup_solve) with the same
    // mesh and settings continues bit for bit.
    void checkpoint(const std::string& path);
    void resume(const std::string& path);
 private:
    Settings m_settings;
    std::shared_ptr<System> m_system;
    std::shared_ptr<GlobalStepBehavior> m_global_step_behavior;
    std::shared_ptr<AlgorithmData> m_algorithm_data;
    MicroTimer m_micro_timer;
    bool m_initialized = false;
};

// src/Solver.hpp:37-72
class Solver {
 public:
    explicit Solver(const Settings& settings) : m_system(std::make_shared<System>(settings)), m_settings(settings) {}
    virtual ~Solver() {}
    virtual void set_pins(const std::vector<int>& inds, const std::vector<math::Vec3>& points = {});
    virtual bool initialize(const Settings& settings);
    virtual void solve();
    const Settings& settings() { return m_settings; }
    AlgorithmData* algorithm_data() { return m_admm.algorithm_data(); }
    ADMM& admm() { return m_admm; }
    std::shared_ptr<System> m_system;
    Settings m_settings;
 protected:
    ADMM m_admm;
    bool initialized = false;
};

}  // namespace wrapd
