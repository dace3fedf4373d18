// Host-side WRAPD control (see wrapd_host.hpp).  Each method cites the reference lines
// whose behaviour it keeps.
#include "wrapd_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>

#include "host_error.hpp"

namespace wrapd {

// ---- Settings (src/Settings.cpp:31-209) ---------------------------------------------
std::string Settings::IO::strip_extension(const std::string& raw) {
    // src/Settings.hpp:81-91: parent path + stem
    std::string parent, name = raw;
    const size_t slash = raw.find_last_of('/');
    if (slash != std::string::npos) { parent = raw.substr(0, slash); name = raw.substr(slash + 1); }
    const size_t dot = name.find_last_of('.');
    if (dot != std::string::npos && dot > 0) name = name.substr(0, dot);
    return parent.empty() ? name : parent + "/" + name;
}

bool Settings::parse_args(int argc, char** argv) {
    if (argc == 1) { std::printf("%s", help_text().c_str()); return true; }
    std::vector<std::string> cmds;
    std::vector<std::string> vals;
    for (int i = 1; i < argc; ++i) {
        const std::string a(argv[i]);
        // a token starting with '-' is a flag unless a digit or '.' follows (src/Settings.cpp:43-54)
        const bool is_flag = a.size() > 1 && a[0] == '-' && !(std::isdigit(static_cast<unsigned char>(a[1])) || a[1] == '.');
        if (is_flag) {
            cmds.push_back(a);
            vals.emplace_back();
        } else {
            if (vals.empty()) throw std::runtime_error("Error: value without a flag: " + a);
            vals.back() += a + " ";
        }
    }
    for (size_t i = 0; i < cmds.size(); ++i) {
        const std::string& arg = cmds[i];
        std::stringstream val(vals[i]);
        auto need_int01 = [&](const char* what) {
            int v = -1;
            val >> v;
            if (v != 0 && v != 1) throw std::runtime_error(std::string("Error: Invalid value provided for argument ") + what);
            return v;
        };
        if (arg == "-help" || arg == "--help" || arg == "-h") { std::printf("%s", help_text().c_str()); return true; }
        else if (arg == "-rest") { std::string s; val >> s; m_io.rest_mesh(s); }
        else if (arg == "-init") { std::string s; val >> s; m_io.init_mesh(s); }
        else if (arg == "-soln") { std::string s; val >> s; m_io.soln_mesh(s); }
        else if (arg == "-handles") { std::string s; val >> s; m_io.handles(s); }
        else if (arg == "-out") { std::string s; val >> s; m_io.output_dir(s); }
        else if (arg == "-p") { int v; val >> v; m_io.verbose(v); }
        else if (arg == "-s") { int v; val >> v; m_io.should_save_data(v != 0); }
        else if (arg == "-it") { val >> m_max_admm_iters; }
        else if (arg == "-tol") { val >> m_tol; }
        else if (arg == "-rotaware" || arg == "-ra") {
            m_rot_awareness = need_int01("-rotaware (or -ra)") ? RotAwareness::ENABLED : RotAwareness::DISABLED;
        } else if (arg == "-reweighting" || arg == "-rw") {
            m_reweighting = need_int01("-reweighting (or -rw)") ? Reweighting::ENABLED : Reweighting::DISABLED;
        } else if (arg == "-rw_delay") { val >> m_reweighting_delay; }
        else if (arg == "-gamma") { val >> m_gamma; }
        else if (arg == "-global_it") { int v; val >> v; m_gstep.max_iters(v); }
        else if (arg == "-global_ls") {
            m_gstep.linesearch(need_int01("-global_linesearch") ? GStep::LineSearch::Backtracking : GStep::LineSearch::None);
        } else if (arg == "-kappa") { double v; val >> v; m_gstep.kappa(v); }
        else if (arg == "-model") { std::string s; val >> s; m_model_settings.elastic_model(s); }
        else if (arg == "-beta_static") { double v; val >> v; m_model_settings.elastic_beta_static(v); }
        else if (arg == "-beta_min") { double v; val >> v; m_model_settings.elastic_beta_min(v); }
        else if (arg == "-beta_max") { double v; val >> v; m_model_settings.elastic_beta_max(v); }
        else if (arg == "-poisson") { double v; val >> v; m_model_settings.poisson(v); }
        // extensions
        else if (arg == "-youngs" || arg == "-E") { double v; val >> v; m_model_settings.youngs(v); }
        else if (arg == "-lbfgs_m") { int v; val >> v; m_gstep.memory(v); }
        else if (arg == "-prox_iters") { int v; val >> v; m_model_settings.elastic_prox_iters(v); }
        else if (arg == "-precision") {
            std::string s; val >> s;
            if (s == "fp64") m_precision = WRAPD_FP64;
            else if (s == "fp32") m_precision = WRAPD_FP32;
            else throw std::runtime_error("Error: -precision must be fp64 or fp32");
        } else if (arg == "-prox") {
            std::string s; val >> s;
            if (s == "lbfgs") m_prox_mode = WRAPD_PROX_LBFGS;
            else if (s == "newton") m_prox_mode = WRAPD_PROX_NEWTON;
            else if (s == "learned") m_prox_mode = WRAPD_PROX_LEARNED;
            else throw std::runtime_error("Error: -prox must be lbfgs, newton or learned");
        } else if (arg == "-fp64_switch") { val >> m_fp64_switch; }
        else if (arg == "-device") { val >> m_device; }
        else throw std::runtime_error("Error: Unrecognized command-line argument: " + arg);
    }
    return false;
}

std::string Settings::help_text() {
    return "\n==========================================\nArgs:\n"
           " General: \n"
           " -rest: zero-energy rest mesh (path to tetmesh .node file)  [REQUIRED]\n"
           " -init: initial deformed mesh (path to tetmesh .node file)  [REQUIRED]\n"
           " -handles: list of constrained vertices                     [REQUIRED]\n"
           " -soln: pre-computed solution (path to tetmesh .node file)  [OPTIONAL]\n"
           " -p: print stats to terminal while running                  [Default: 1]       (0=disabled, 1=enabled) \n"
           " -s: save stats to file after running                       [Default: 1]       (0=disabled, 1=enabled) \n"
           " -it: maximum # ADMM iterations                             [Default: 5000]    (any integer > 0) \n"
           " -tol: tolerance for ADMM early exit                        [Default: 1.e-11]  (any small float > 0) \n"
           " -model: constitutive model for elasticity                  [Default: nh]      (nh=Neo-Hookean, arap=ARAP) \n"
           " -poisson: Poisson's ratio (NH model only)                  [Default: 0.45]    (any float in the range [0, 0.5)) \n"
           " -beta_static: weight^2 mult (no reweighting)               [Default: 1.0]     (any float >= 1.0) \n"
           "\n Rotation awareness: \n"
           " -ra: rotation awareness toggle                             [Default: 1]       (0=disabled, 1=enabled) \n"
           " -global_it: max # global step L-BFGS iters                 [Default: 100]     (any integer > 0) \n"
           " -global_ls: linesearch in the global step                  [Default: 0]       (0=none, 1=backtracking) \n"
           " -kappa: global step early exit coeff.                      [Default: 1.0]     (any float > 0)] \n"
           "\n Dynamic reweighting: \n"
           " -rw: dynamic reweighting toggle                            [Default: 1]       (0=disabled, 1=enabled) \n"
           " -rw_delay: # ADMM iters to use static rest weights         [Default: 50]      (any integer > 0) \n"
           " -gamma: threshold for reweighting                          [Default:  1.5]    (any float >= 1.0) \n"
           " -beta_min: smallest allowable weight^2 mult                [Default:  0.1]    (any float >= 1.0) \n"
           " -beta_max: largest allowable weight^2 mult                 [Default: 10.0]    (any float >= 1.0) \n"
           "\n MI355X device path: \n"
           " -youngs: Young's modulus                                   [Default: 1e7] \n"
           " -lbfgs_m: global-step L-BFGS memory (1..8)                 [Default: 6] \n"
           " -prox_iters: prox iteration cap                            [Default: 20] \n"
           " -precision: local step arithmetic                          [Default: fp64]    (fp64, fp32 with fp64 fallback) \n"
           " -prox: stretch prox solver                                 [Default: lbfgs]   (lbfgs=reference, newton, learned) \n"
           " -fp64_switch: fp32 runs switch to fp64 below this residual [Default: 1e-8] \n"
           " -device: HIP device index                                  [Default: 0] \n"
           " -out: output directory                                     [Default: output] \n"
           "==========================================\n";
}

// ---- TetElements -----------------------------------------------------------------------
void TetElements::check(int st, const char* what) const {
    if (st != WRAPD_OK) throw DeviceError(st, std::string(what) + ": " + wrapd_last_error(m_ctx));
}

TetElements::TetElements(const std::vector<double>& rest, const std::vector<int>& tets, const std::vector<int>& pins,
                         const Settings& settings, int vertex_offset)
    : Elements(static_cast<int>(tets.size() / 4)), m_settings(settings) {
    m_tet.resize(tets.size() / 4);
    for (size_t e = 0; e < m_tet.size(); ++e)
        for (int q = 0; q < 4; ++q) m_tet[e].x[q] = tets[4 * e + q] + vertex_offset;
    {
        std::vector<char> pinned(rest.size() / 3, 0);
        for (int p : pins)
            if (p >= 0 && p < static_cast<int>(pinned.size())) pinned[p] = 1;
        for (size_t v = 0; v < pinned.size(); ++v)
            if (!pinned[v]) m_free_rows.push_back(static_cast<int>(v));
    }
    int st = wrapd_create(&m_ctx, settings.m_device);
    if (st != WRAPD_OK) throw DeviceError(st, "wrapd_create: no usable HIP device");
    const ModelSettings& ms = settings.m_model_settings;
    wrapd_material mat;
    mat.model = ms.elastic_model() == ModelSettings::ElasticModel::NH ? 1 : 0;
    mat.mu = ms.elastic_lame().mu();
    mat.lambda = ms.elastic_lame().lambda();
    mat.beta_static = ms.elastic_beta_static();
    mat.beta_min = ms.elastic_beta_min();
    mat.beta_max = ms.elastic_beta_max();
    const int n_v = static_cast<int>(rest.size() / 3);
    m_n_e = static_cast<int>(tets.size() / 4);
    const double* mu_e = settings.m_mu_e.empty() ? nullptr : settings.m_mu_e.data();
    const double* la_e = settings.m_lambda_e.empty() ? nullptr : settings.m_lambda_e.data();
    st = wrapd_upload_mesh(m_ctx, n_v, rest.data(), m_n_e, tets.data(), static_cast<int>(pins.size()), pins.data(),
                           &mat, mu_e, la_e);
    if (st == WRAPD_ERR_INVERTED_REST) {
        wrapd_destroy(m_ctx);
        m_ctx = nullptr;
        throw DeviceError(st, "**Error: Inverted tet in TetElements constructor.");
    }
    check(st, "wrapd_upload_mesh");
    check(wrapd_sizes(m_ctx, nullptr, nullptr, &m_n_f, nullptr), "wrapd_sizes");
    wrapd_lbfgs_cfg cfg;
    cfg.memory = settings.m_gstep.memory();
    cfg.max_iters = settings.m_gstep.max_iters();
    cfg.kappa = settings.m_gstep.kappa();
    cfg.linesearch = settings.m_gstep.linesearch() == Settings::GStep::LineSearch::Backtracking ? 1 : 0;
    check(wrapd_lbfgs_configure(m_ctx, &cfg), "wrapd_lbfgs_configure");
    check(wrapd_set_rot_aware(m_ctx, settings.m_rot_awareness == Settings::RotAwareness::ENABLED ? 1 : 0),
          "wrapd_set_rot_aware");
}

TetElements::~TetElements() {
    if (m_ctx) wrapd_destroy(m_ctx);
}

void TetElements::update_fix_cache(const math::MatX3& x_fix) { check(wrapd_set_x_fix(m_ctx, x_fix.data()), "update_fix_cache"); }
void TetElements::set_x_free(const math::MatX3& x) { check(wrapd_set_x_free(m_ctx, x.data()), "set_x_free"); }
void TetElements::get_x_free(math::MatX3& x) {
    x = math::MatX3(m_n_f);
    check(wrapd_get_x_free(m_ctx, x.data()), "get_x_free");
}
void TetElements::update_defo_cache() { check(wrapd_defo_energy(m_ctx, &m_energy, &m_flips), "update_defo_cache"); }
void TetElements::initialize_dual_vars() {
    const bool ra = m_settings.m_rot_awareness == Settings::RotAwareness::ENABLED;
    double pe = 0.0;
    check(wrapd_init_duals(m_ctx, ra ? 1 : 0, &pe), "initialize_dual_vars");
    if (ra) m_curr_pe = pe;
}

double TetElements::polar_update(bool update_cand) {
    polar_update_launch(update_cand);
    return polar_update_finish();
}

void TetElements::polar_update_launch(bool update_cand) {
    // src/TetElements.cpp:256-356; the candidate-weight counter of :404-428
    wrapd_local_args a;
    a.precision = m_fp64_tail ? WRAPD_FP64 : m_settings.m_precision;
    a.prox_mode = m_settings.m_prox_mode;
    a.prox_iters = m_settings.m_model_settings.elastic_prox_iters();
    a.update_cand = update_cand ? 1 : 0;
    a.cand_use_hessian = (update_cand && m_cand_count > m_settings.m_reweighting_delay) ? 1 : 0;
    // Newton tolerance relative to the prox Hessian scale; 0 selects the precision's default
    a.newton_rtol = m_settings.m_newton_rtol > 0 ? m_settings.m_newton_rtol
                                                 : (a.precision == WRAPD_FP32 ? 1e-6 : 1e-12);
    a.fallback_sigma = m_settings.m_fallback_sigma;
    if (a.prox_mode != WRAPD_PROX_LBFGS && a.prox_iters < 30) a.prox_iters = 30;
    check(wrapd_local_step_launch(m_ctx, &a), "polar_update");
    if (update_cand) m_cand_count++;
}

double TetElements::polar_update_finish() {
    check(wrapd_local_step_result(m_ctx, m_curr_pe, &m_last_local), "polar_update");
    m_curr_pe = m_last_local.pe;
    return m_last_local.delta_auglag;
}

void TetElements::update(bool update_cand) {
    wrapd_local_args a{};
    a.precision = WRAPD_FP64;
    a.prox_mode = WRAPD_PROX_LBFGS;
    a.prox_iters = m_settings.m_model_settings.elastic_prox_iters();
    a.update_cand = update_cand ? 1 : 0;
    a.cand_use_hessian = (update_cand && m_cand_count > m_settings.m_reweighting_delay) ? 1 : 0;
    check(wrapd_local_step_unaware(m_ctx, &a), "update");
    if (update_cand) {
        m_cand_count++;
        check(wrapd_max_wsq_ratio(m_ctx, &m_last_local.max_wsq_ratio), "max_wsq_ratio");
    }
}

void TetElements::midupdate_weights(bool direct) {
    // src/TetElements.cpp:360-388
    if (m_settings.m_reweighting != Settings::Reweighting::ENABLED) return;
    if (m_settings.m_model_settings.elastic_model() != ModelSettings::ElasticModel::NH)
        throw DeviceError(WRAPD_ERR_STATE, "Error: Adaptive weights are only supported for the NH model.");
    const int use_h = m_cand_count > m_settings.m_reweighting_delay ? 1 : 0;
    check(wrapd_candidate_weights(m_ctx, use_h, direct ? 1 : 0), "midupdate_weights");
    m_cand_count++;
    update_actual_weights();
}

void TetElements::update_actual_weights() { check(wrapd_update_actual_weights(m_ctx), "update_actual_weights"); }
double TetElements::max_wsq_ratio() { return m_last_local.max_wsq_ratio; }

// ---- the rest of the Elements interface (src/TetElements.cpp), for callers outside ADMM ----
std::vector<double> TetElements::elem(const char* name, int width) const {
    std::vector<double> v(static_cast<size_t>(m_n_e) * width);
    check(wrapd_get_elem(m_ctx, name, v.data()), name);
    return v;
}

math::MatX3 TetElements::all_to_free(const math::MatX3& X) const {
    math::MatX3 x(static_cast<int>(m_free_rows.size()));
    for (size_t i = 0; i < m_free_rows.size(); ++i)
        for (int c = 0; c < 3; ++c) x(static_cast<int>(i), c) = X(m_free_rows[i], c);
    return x;
}

void TetElements::update_defo_cache(const math::MatX3& x_free) {
    set_x_free(x_free);
    update_defo_cache();
}

void TetElements::initialize_dual_vars(const math::MatX3& X) {
    // src/TetElements.cpp:145-183 reads F(X) without moving x: evaluate at X, then return to x
    math::MatX3 x;
    get_x_free(x);
    set_x_free(all_to_free(X));
    initialize_dual_vars();
    update_defo_cache(x);
}

void TetElements::midupdate_weights(const math::MatX3& X) {
    // src/TetElements.cpp:360-379: the candidate weights from F(X), X all vertices
    math::MatX3 x;
    get_x_free(x);
    set_x_free(all_to_free(X));
    midupdate_weights(true);
    update_defo_cache(x);
}

math::MatX3 TetElements::add_to_b() {
    math::MatX3 b(m_n_f);
    check(wrapd_add_to_b(m_ctx, b.data()), "add_to_b");
    return b;
}

void TetElements::get_A_triplets(math::Triplets& triplets) {
    // src/TetElements.cpp:483-496 gives one triplet per element entry; the sums per (i, j)
    // are the same matrix after SpMat::setFromTriplets
    const int nnz = wrapd_get_A(m_ctx, nullptr, nullptr, nullptr);
    if (nnz < 0) check(-nnz, "get_A_triplets");
    std::vector<int> rp(m_n_f + 1), cols(nnz);
    std::vector<double> vals(nnz);
    const int r = wrapd_get_A(m_ctx, rp.data(), cols.data(), vals.data());
    if (r < 0) check(-r, "get_A_triplets");
    triplets.reserve(triplets.size() + static_cast<size_t>(nnz));
    for (int i = 0; i < m_n_f; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p) triplets.emplace_back(i, cols[p], vals[p]);
}

void TetElements::update_A_coeffs(math::SpMat& A) {
    // src/TetElements.cpp:498-510 adds every entry into A (inserting it into an empty A)
    math::Triplets t;
    get_A_triplets(t);
    if (A.nonZeros() == 0) {
        if (A.rows() != m_n_f || A.cols() != m_n_f) A.resize(m_n_f, m_n_f);
        A.setFromTriplets(t.begin(), t.end());
        return;
    }
    for (const math::Triplet& e : t) A.coeffRef(e.row(), e.col()) += e.value();
}

double TetElements::penalty_energy(bool lagged_u) const {
    // src/TetElements.cpp:512-525
    if (lagged_u) throw std::runtime_error("penalty_energy(lagged_u): the previous duals are not kept on the device");
    const std::vector<double> S = elem("symF", 9), P = elem("P", 9), w = elem("wsq", 1);
    double se = 0.0;
    for (int e = 0; e < m_n_e; ++e) {
        double n2 = 0.0;
        for (int k = 0; k < 9; ++k) {
            const double d = S[9ull * e + k] - P[9ull * e + k];
            n2 += d * d;
        }
        se += 0.5 * w[e] * n2;
    }
    return se;
}

double TetElements::primal_residual_sq(Settings::RotAwareness ra) const {
    // src/TetElements.cpp:527-544: sum w^2 |S - Z|^2, S the cached symmetric factor (rotation
    // aware) or F^T (unaware), which is what the device's cache holds in each mode
    if (ra != m_settings.m_rot_awareness)
        throw std::runtime_error("primal_residual_sq: the deformation cache holds the other rotation mode's factor");
    const std::vector<double> S = elem("symF", 9), P = elem("P", 9), U = elem("U", 9), w = elem("wsq", 1);
    double r = 0.0;
    for (int e = 0; e < m_n_e; ++e) {
        double n2 = 0.0;
        for (int k = 0; k < 9; ++k) {
            const double d = S[9ull * e + k] - (P[9ull * e + k] + U[9ull * e + k]);
            n2 += d * d;
        }
        r += w[e] * n2;
    }
    return r;
}

double TetElements::polar_potential_energy() const {
    // src/TetElements.cpp:638-675: the energy of Z.  sigma(Z^T) enters only as |sigma|^2 =
    // |Z|_F^2 and J = prod sigma = det Z (the SVD's sign convention carries det's sign)
    const std::vector<double> P = elem("P", 9), U = elem("U", 9), vol = elem("vol", 1);
    const ModelSettings& ms = m_settings.m_model_settings;
    const bool arap = ms.elastic_model() == ModelSettings::ElasticModel::ARAP;
    double pe = 0.0;
    for (int e = 0; e < m_n_e; ++e) {
        double Z[9];
        for (int k = 0; k < 9; ++k) Z[k] = P[9ull * e + k] + U[9ull * e + k];
        if (arap) {
            double n2 = 0.0;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    const double d = Z[3 * r + c] - (r == c ? 1.0 : 0.0);
                    n2 += d * d;
                }
            pe += 0.5 * vol[e] * n2;
            continue;
        }
        const double mu = m_settings.m_mu_e.empty() ? ms.elastic_lame().mu() : m_settings.m_mu_e[e];
        const double lambda = m_settings.m_lambda_e.empty() ? ms.elastic_lame().lambda() : m_settings.m_lambda_e[e];
        double I1 = 0.0;
        for (int k = 0; k < 9; ++k) I1 += Z[k] * Z[k];
        const double J = Z[0] * (Z[4] * Z[8] - Z[5] * Z[7]) - Z[1] * (Z[3] * Z[8] - Z[5] * Z[6]) +
                         Z[2] * (Z[3] * Z[7] - Z[4] * Z[6]);
        const double J0 = 0.1;
        double val = 0.5 * mu * (I1 - 3.0);
        if (J > J0) {
            const double lj = std::log(J);
            val += -mu * lj + 0.5 * lambda * lj * lj;
        } else {
            const double fJ = std::log(J0) + (J - J0) / J0 - 0.5 * std::pow(J / J0 - 1.0, 2.0);
            val += -mu * fJ + 0.5 * lambda * fJ * fJ;
        }
        pe += vol[e] * val;
    }
    return pe;
}

double TetElements::global_obj_value(const math::MatX3& x_free) {
    // src/TetElements.cpp:678-688; the deformation cache ends at x_free, as there
    set_x_free(x_free);
    double v = 0.0;
    check(wrapd_global_value_grad(m_ctx, &v, nullptr), "global_obj_value");
    update_defo_cache();
    return v;
}

double TetElements::global_obj_grad(const math::MatX3& x_free, math::MatX3& grad) {
    set_x_free(x_free);
    grad = math::MatX3(m_n_f);
    double v = 0.0;
    check(wrapd_global_value_grad(m_ctx, &v, grad.data()), "global_obj_grad");
    update_defo_cache();
    return v;
}

math::MatX3 TetElements::global_obj_grad() {
    math::MatX3 g(m_n_f);
    double v = 0.0;
    check(wrapd_global_value_grad(m_ctx, &v, g.data()), "global_obj_grad");
    return g;
}
void TetElements::refactor() {
    check(wrapd_refactor(m_ctx), "factorize");
    check(wrapd_synchronize(m_ctx), "factorize");
}

void TetElements::save(ByteWriter& w) {
    w.pod(m_curr_pe);
    w.pod(m_cand_count);
    w.pod(m_energy);
    w.pod(m_flips);
    w.pod(m_last_local);
    w.pod(m_fp64_tail);
    size_t n = 0;
    check(wrapd_state_save(m_ctx, nullptr, 0, &n), "state size");
    std::vector<char> dev(n);
    check(wrapd_state_save(m_ctx, dev.data(), n, &n), "state save");
    w.vec(dev);
}

void TetElements::load(ByteReader& r) {
    r.pod(m_curr_pe);
    r.pod(m_cand_count);
    r.pod(m_energy);
    r.pod(m_flips);
    r.pod(m_last_local);
    r.pod(m_fp64_tail);
    std::vector<char> dev;
    r.vec(dev);
    check(wrapd_state_load(m_ctx, dev.data(), dev.size()), "state load");
}
void TetElements::reset_history() { check(wrapd_lbfgs_reset(m_ctx), "clear_history"); }
void TetElements::global_step_begin() { check(wrapd_global_step_begin(m_ctx), "global_step"); }

void TetElements::global_step(double tol, wrapd_global_out& out) {
    check(wrapd_global_step(m_ctx, tol, &out), "global_step");
    m_energy = out.energy;
    m_flips = out.flips;
}
void TetElements::global_step_unaware(wrapd_global_out& out) {
    check(wrapd_global_step_unaware(m_ctx, &out), "global_step");
    m_energy = out.energy;
    m_flips = out.flips;
}
void TetElements::set_x_ref(const math::MatX3& x) { check(wrapd_set_x_ref(m_ctx, x.data()), "set_x_ref"); }
void TetElements::x_distance_launch() { check(wrapd_x_distance_launch(m_ctx), "x_distance"); }
double TetElements::x_distance_result() {
    double d = 0.0;
    check(wrapd_x_distance_result(m_ctx, &d), "x_distance");
    return d;
}
double TetElements::x_distance() {
    double d = 0.0;
    check(wrapd_x_distance(m_ctx, &d), "x_distance");
    return d;
}

// ---- System (src/System.hpp:75-92, src/System.cpp:33-63) -------------------------------
int System::add_nodes(const math::MatX3& X, const math::MatX3& X_init, const math::MatX3& X_soln) {
    const int prev = m_X.rows(), n = X.rows();
    auto app = [&](math::MatX3& dst, const math::MatX3& src) {
        dst.rows_ = prev + n;
        dst.v.insert(dst.v.end(), src.v.begin(), src.v.end());
    };
    app(m_X, X);
    app(m_X_init, X_init);
    app(m_X_soln, X_soln);
    return prev + n;
}

void System::set_pins(const std::vector<int>& inds, const std::vector<math::Vec3>& points) {
    const bool pin_in_place = points.size() != inds.size();
    if ((m_X.rows() == 0 && pin_in_place) || (pin_in_place && !points.empty()))
        throw DeviceError(WRAPD_ERR_ARG, "**Solver::set_pins Error: Bad input.");
    m_constraints->m_pin_inds = inds;
}

void System::get_A(math::SpMat& A) {
    // src/System.cpp:78-90
    math::Triplets t;
    if (num_tet_elements() > 0) m_tet_elements->get_A_triplets(t);
    const int n = m_tet_elements ? m_tet_elements->num_free() : 0;
    A.resize(n, n);
    A.setFromTriplets(t.begin(), t.end());
}

void System::update_A(math::SpMat& A) {
    // src/System.cpp:93-100
    if (num_tet_elements() > 0) m_tet_elements->update_A_coeffs(A);
}

bool System::possibly_update_weights(double gamma) {
    if (num_tet_elements() == 0) return false;
    const bool rw = m_tet_elements->max_wsq_ratio() > gamma;
    if (rw) m_tet_elements->update_actual_weights();
    return rw;
}

// ---- AlgorithmData (src/AlgorithmData.cpp:28-121) --------------------------------------
AlgorithmData::AlgorithmData(const Settings& s, std::shared_ptr<System> sys) : m_settings(s), m_system(sys) {
    const int n = sys->num_all_verts();
    std::vector<char> pinned(n, 0);
    for (int p : sys->constraint_set()->m_pin_inds) { pinned[p] = 1; m_fix_list.push_back(p); }
    for (int i = 0; i < n; ++i) if (!pinned[i]) m_free_list.push_back(i);
    m_n_free = static_cast<int>(m_free_list.size());
    const int it = std::max(0, s.m_max_admm_iters);
    m_objectives.reserve(it + 1);
    m_per_iter_inner_iters.reserve(it);
}

void AlgorithmData::log_initial_data() {
    m_init_objective = m_system->problem_objective();
    m_objectives.assign(1, m_init_objective);
    m_state_x_errors.assign(1, 1.0);
    m_reweighted.assign(1, 1);
    m_accumulated_time_s.assign(1, 0.0);
    m_wall_time_s.assign(1, 0.0);
    m_flips_count.assign(1, m_system->inversion_count());
    m_init_state_x_error = m_system->tet_elements()->x_distance();
}

void AlgorithmData::resolve_x_error() const {
    if (m_xerr_pending < 0) return;
    m_state_x_errors[m_xerr_pending] = m_system->tet_elements()->x_distance_result() / m_init_state_x_error;
    m_xerr_pending = -1;
}

void AlgorithmData::update_per_iter_data() {
    const double obj = m_system->problem_objective();
    m_flips_count.push_back(m_system->inversion_count());
    resolve_x_error();
    if (m_settings.m_io.soln_provided()) {   // the convergence test reads it now
        m_state_x_errors.push_back(m_system->tet_elements()->x_distance() / m_init_state_x_error);
    } else {
        m_system->tet_elements()->x_distance_launch();
        m_xerr_pending = static_cast<int>(m_state_x_errors.size());
        m_state_x_errors.push_back(std::numeric_limits<double>::quiet_NaN());
    }
    m_objectives.push_back(obj);
    const double rel = std::fabs(m_objectives[m_iter + 1] - m_objectives[m_iter]) / (std::fabs(m_objectives[m_iter + 1]) + 1.0);
    m_norm_obj_residuals.push_back(rel);
}

void AlgorithmData::update_per_iter_runtime() {
    m_accumulated_time_s.push_back((m_runtime.local_ms + m_runtime.global_ms + m_runtime.refactor_ms) / 1000.);
    m_wall_time_s.push_back(m_wall.elapsed_ms() / 1000.);
}

void AlgorithmData::save(ByteWriter& w) const {
    resolve_x_error();
    w.pod(m_runtime);
    w.pod(m_iter);
    w.pod(m_init_state_x_error);
    w.pod(m_init_objective);
    w.pod(m_init_auglag_obj);
    w.pod(m_last_localstep_delta_auglag_obj);
    w.vec(m_norm_obj_residuals);
    w.vec(m_per_iter_inner_iters);
    w.vec(m_per_iter_ls_delta_auglag);
    w.vec(m_per_iter_gs_delta_auglag);
    w.vec(m_objectives);
    w.vec(m_state_x_errors);
    w.vec(m_accumulated_time_s);
    w.vec(m_wall_time_s);
    w.vec(m_reweighted);
    w.vec(m_flips_count);
    w.vec(m_curr_x_free.v);
    w.vec(m_curr_x_fix.v);
}

void AlgorithmData::load(ByteReader& r) {
    r.pod(m_runtime);
    r.pod(m_iter);
    r.pod(m_init_state_x_error);
    r.pod(m_init_objective);
    r.pod(m_init_auglag_obj);
    r.pod(m_last_localstep_delta_auglag_obj);
    r.vec(m_norm_obj_residuals);
    r.vec(m_per_iter_inner_iters);
    r.vec(m_per_iter_ls_delta_auglag);
    r.vec(m_per_iter_gs_delta_auglag);
    r.vec(m_objectives);
    r.vec(m_state_x_errors);
    r.vec(m_accumulated_time_s);
    r.vec(m_wall_time_s);
    r.vec(m_reweighted);
    r.vec(m_flips_count);
    std::vector<double> xf, xx;
    r.vec(xf);
    r.vec(xx);
    if (xf.size() != m_curr_x_free.v.size() || xx.size() != m_curr_x_fix.v.size())
        throw std::runtime_error("checkpoint: the state belongs to another mesh");
    m_curr_x_free.v = xf;
    m_curr_x_fix.v = xx;
}

void AlgorithmData::print_initial_data() const {
    std::cout << "\nADMM it |  time  |     objective    |obj.resid.|";
    if (m_settings.m_io.soln_provided()) std::cout << " X error  |";
    std::cout << "flips| RW? |";
    if (m_settings.m_rot_awareness == Settings::RotAwareness::ENABLED) std::cout << "LS:DeltaAL|GS:DeltaAL| L-BFGS ";
    std::cout << "\n   init |" << std::fixed << std::setprecision(3) << std::setw(7) << m_accumulated_time_s[0] << " | "
              << std::scientific << std::setprecision(10) << m_objectives[0] << " | -------- | ";
    if (m_settings.m_io.soln_provided()) std::cout << std::scientific << std::setprecision(2) << m_state_x_errors[0] << " | ";
    std::cout << std::setw(3) << m_flips_count[0] << " | " << std::setw(3) << m_reweighted[0] << " | ";
    if (m_settings.m_rot_awareness == Settings::RotAwareness::ENABLED) std::cout << "-------- | -------- | -----";
    std::cout << std::endl << std::fixed;
}

static void print_row(std::ostream& os, const Settings& s, int i, double t, double obj, double res, double xe, int fl,
                      int rw, double ls, double gs, int inner) {
    os << std::setw(7) << i << " |" << std::fixed << std::setprecision(3) << std::setw(7) << t << " | " << std::scientific
       << std::setprecision(10) << obj << " | " << std::scientific << std::setprecision(2) << res << " | ";
    if (s.m_io.soln_provided()) os << std::scientific << std::setprecision(2) << xe << " | ";
    os << std::setw(3) << fl << " | " << std::setw(3) << rw << " | ";
    if (s.m_rot_awareness == Settings::RotAwareness::ENABLED)
        os << std::setprecision(2) << ls << " | " << gs << " |" << std::setw(4) << inner;
    os << std::endl << std::fixed;
}

void AlgorithmData::print_curr_iter_data() const {
    const int i = m_iter;
    print_row(std::cout, m_settings, i, m_accumulated_time_s[i + 1], m_objectives[i + 1], m_norm_obj_residuals[i],
              m_state_x_errors[i + 1], m_flips_count[i + 1], m_reweighted[i + 1], m_per_iter_ls_delta_auglag[i],
              m_per_iter_gs_delta_auglag[i], m_per_iter_inner_iters[i]);
}

void AlgorithmData::print_final_iter_data() const {
    const int i = static_cast<int>(m_objectives.size()) - 2;
    if (i < 0) return;
    print_row(std::cout, m_settings, i, m_accumulated_time_s[i + 1], m_objectives[i + 1], m_norm_obj_residuals[i],
              m_state_x_errors[i + 1], m_flips_count[i + 1], m_reweighted[i + 1], m_per_iter_ls_delta_auglag[i],
              m_per_iter_gs_delta_auglag[i], m_per_iter_inner_iters[i]);
}

// ---- GlobalStepBehavior (src/GlobalStepBehavior.hpp:78-268) -----------------------------
GlobalStepBehavior::GlobalStepBehavior(const Settings& s, std::shared_ptr<System> sys, std::shared_ptr<AlgorithmData> ad)
    : m_settings(s), m_system(sys), m_algorithm_data(ad) {
    // midupdate_WtW(X_init) with X_init already on the device, then A and its factorization
    m_system->tet_elements()->midupdate_weights(true);
    m_system->tet_elements()->refactor();
}

void GlobalStepBehavior::initialize_weights() {
    m_system->tet_elements()->midupdate_weights(false);
    m_system->tet_elements()->refactor();
}

bool GlobalStepBehavior::update_weights() {
    const bool updated = m_system->possibly_update_weights(m_settings.m_gamma);
    if (updated) m_system->tet_elements()->refactor();
    return updated;
}

void UnawareGS::global_step() {
    MicroTimer t;
    wrapd_global_out o{};
    m_system->tet_elements()->global_step_unaware(o);
    m_algorithm_data->m_runtime.global_ms += t.elapsed_ms();
    m_algorithm_data->m_per_iter_inner_iters.push_back(1);
    m_algorithm_data->m_runtime.inner_iters += 1;
}

RotAwareGS::RotAwareGS(const Settings& s, std::shared_ptr<System> sys, std::shared_ptr<AlgorithmData> ad)
    : GlobalStepBehavior(s, sys, ad) {
    m_system->tet_elements()->reset_history();   // update_A0(A, true)
}

void RotAwareGS::initialize_weights() {
    GlobalStepBehavior::initialize_weights();
    m_system->tet_elements()->reset_history();   // update_A0(A, false)
}

bool RotAwareGS::update_weights() {
    const bool updated = GlobalStepBehavior::update_weights();
    if (updated) m_system->tet_elements()->reset_history();
    return updated;
}

void RotAwareGS::begin_global_step() { m_system->tet_elements()->global_step_begin(); }

void RotAwareGS::global_step() {
    MicroTimer t;
    wrapd_global_out o{};
    m_system->tet_elements()->global_step(std::fabs(m_algorithm_data->m_last_localstep_delta_auglag_obj), o);
    m_algorithm_data->m_runtime.global_ms += t.elapsed_ms();
    m_algorithm_data->m_per_iter_inner_iters.push_back(o.iters);
    m_algorithm_data->m_runtime.inner_iters += o.iters;
    m_obj_bef = o.obj_bef;
    m_obj_aft = o.obj_aft;
}

// ---- ADMM (src/ADMM.cpp:34-182) -------------------------------------------------------
bool ADMM::initialize(std::shared_ptr<System> sys, const Settings& settings) {
    m_settings = settings;
    m_system = sys;
    m_algorithm_data = std::make_shared<AlgorithmData>(m_settings, m_system);
    m_algorithm_data->m_runtime = RuntimeData();
    m_micro_timer.reset();
    AlgorithmData& ad = *m_algorithm_data;
    const math::MatX3& Xi = sys->X_init();
    ad.m_init_x_fix = math::MatX3(static_cast<int>(ad.m_fix_list.size()));
    for (size_t j = 0; j < ad.m_fix_list.size(); ++j)
        for (int c = 0; c < 3; ++c) ad.m_init_x_fix(static_cast<int>(j), c) = Xi(ad.m_fix_list[j], c);
    ad.m_curr_x_fix = ad.m_init_x_fix;
    auto& te = sys->tet_elements();
    if (!te) throw DeviceError(WRAPD_ERR_STATE, "no elements");
    te->update_fix_cache(ad.m_curr_x_fix);
    // the global-step constructors evaluate F(X_init): put X_init's free rows on the device
    math::MatX3 xf(ad.num_free_verts());
    for (int i = 0; i < ad.num_free_verts(); ++i)
        for (int c = 0; c < 3; ++c) xf(i, c) = Xi(ad.m_free_list[i], c);
    te->set_x_free(xf);
    if (m_settings.m_rot_awareness == Settings::RotAwareness::DISABLED)
        m_global_step_behavior = std::make_shared<UnawareGS>(m_settings, m_system, m_algorithm_data);
    else
        m_global_step_behavior = std::make_shared<RotAwareGS>(m_settings, m_system, m_algorithm_data);
    ad.m_runtime.setup_solve_ms += m_micro_timer.elapsed_ms();
    m_initialized = true;
    return true;
}

void ADMM::setup_solve(math::MatX3& X) {
    m_micro_timer.reset();
    AlgorithmData& ad = *m_algorithm_data;
    auto& te = m_system->tet_elements();
    const int nf = ad.num_free_verts();
    ad.m_curr_x_free = math::MatX3(nf);
    ad.m_soln_x_free = math::MatX3(nf);
    const math::MatX3& Xs = m_system->X_soln();
    for (int i = 0; i < nf; ++i)
        for (int c = 0; c < 3; ++c) {
            ad.m_curr_x_free(i, c) = X(ad.m_free_list[i], c);
            ad.m_soln_x_free(i, c) = Xs(ad.m_free_list[i], c);
        }
    te->update_fix_cache(ad.m_curr_x_fix);
    te->set_x_free(ad.m_curr_x_free);
    te->set_x_ref(ad.m_soln_x_free);
    te->update_defo_cache();
    te->initialize_dual_vars();
    ad.m_iter = 0;
    ad.log_initial_data();
    if (m_settings.m_io.verbose()) ad.print_initial_data();
    ad.m_runtime.setup_solve_ms += m_micro_timer.elapsed_ms();
    if (m_settings.m_reweighting == Settings::Reweighting::ENABLED) m_global_step_behavior->initialize_weights();
    ad.m_wall.reset();
}

void ADMM::iterate() {
    AlgorithmData& ad = *m_algorithm_data;
    const bool upd = m_settings.m_reweighting == Settings::Reweighting::ENABLED;
    // mixed precision: once the objective residual is below fp64_switch, the remaining
    // local steps run in FP64, so the run ends at the FP64 fixed point
    if (m_settings.m_precision == WRAPD_FP32 && ad.m_runtime.fp64_switch_iter < 0 && ad.m_iter > 0 &&
        ad.current_normalized_objective_residual() < m_settings.m_fp64_switch) {
        m_system->tet_elements()->set_fp64_tail(true);
        ad.m_runtime.fp64_switch_iter = ad.m_iter;
    }
    m_micro_timer.reset();
    if (m_settings.m_rot_awareness == Settings::RotAwareness::DISABLED) {
        m_system->local_update(upd);
    } else {   // the global step's first launches are queued before the host reads the local step's results
        auto& te = m_system->tet_elements();
        te->polar_update_launch(upd);
        m_global_step_behavior->begin_global_step();
        ad.m_last_localstep_delta_auglag_obj = te->polar_update_finish();
    }
    ad.m_runtime.local_ms += m_micro_timer.elapsed_ms();
    const wrapd_local_out& lo = m_system->tet_elements()->last_local();
    ad.m_runtime.prox_iters += lo.prox_iters_total;
    ad.m_runtime.fallback_elems += lo.fallback_count;
    ad.m_runtime.learned_accepted += lo.learned_accepted;
    ad.m_runtime.learned_used += lo.learned_used;
    m_global_step_behavior->global_step();
    const double gs_delta = m_global_step_behavior->last_step_delta_auglag();
    m_micro_timer.reset();
    bool rw = false;
    if (upd) rw = m_global_step_behavior->update_weights();
    ad.m_runtime.refactor_ms += m_micro_timer.elapsed_ms();
    ad.update_per_iter_reweighted(rw);
    ad.update_per_iter_delta_auglag(ad.m_last_localstep_delta_auglag_obj, gs_delta);
    ad.update_per_iter_data();
    ad.update_per_iter_runtime();
    if (m_settings.m_io.verbose() > 0 && ad.m_iter % 50 == 0) ad.print_curr_iter_data();
    ad.m_iter += 1;
}

bool ADMM::converged() {
    const AlgorithmData& ad = *m_algorithm_data;
    if (ad.no_recent_flips() && ad.current_normalized_objective_residual() < m_settings.m_tol) return true;
    if (m_settings.m_io.soln_provided() && ad.current_normalized_x_error() < 1.e-4) return true;
    return false;
}

void ADMM::finish_solve(math::MatX3& X) {
    AlgorithmData& ad = *m_algorithm_data;
    m_system->tet_elements()->get_x_free(ad.m_curr_x_free);
    for (int i = 0; i < ad.num_free_verts(); ++i)
        for (int c = 0; c < 3; ++c) X(ad.m_free_list[i], c) = ad.m_curr_x_free(i, c);
    for (int j = 0; j < ad.num_fix_verts(); ++j)
        for (int c = 0; c < 3; ++c) X(ad.m_fix_list[j], c) = ad.m_curr_x_fix(j, c);
    if (m_settings.m_io.verbose() > 0 && (ad.m_iter - 1) % 50 != 0) {
        ad.print_final_iter_data();
        std::cout << std::endl;
    }
}

// File: "WRPDCKP1", the element/iteration counts, then AlgorithmData, the global-step
// behaviour and the element set (whose device part is wrapd_state_save's blob).
void ADMM::checkpoint(const std::string& path) {
    if (!m_initialized) throw std::runtime_error("checkpoint: ADMM not initialized");
    ByteWriter w;
    w.buf.insert(w.buf.end(), {'W', 'R', 'P', 'D', 'C', 'K', 'P', '1'});
    const int ne = m_system->num_tet_elements(), nf = m_algorithm_data->num_free_verts();
    w.pod(ne);
    w.pod(nf);
    m_algorithm_data->save(w);
    m_global_step_behavior->save(w);
    m_system->tet_elements()->save(w);
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("checkpoint: cannot write " + path);
    const size_t n = std::fwrite(w.buf.data(), 1, w.buf.size(), f);
    const bool ok = std::fclose(f) == 0 && n == w.buf.size();
    if (!ok) throw std::runtime_error("checkpoint: write failed for " + path);
}

void ADMM::resume(const std::string& path) {
    if (!m_initialized) throw std::runtime_error("resume: ADMM not initialized");
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("resume: cannot read " + path);
    std::vector<char> buf;
    char tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    ByteReader r{buf.data(), buf.data() + buf.size()};
    r.need(8);
    if (std::memcmp(r.p, "WRPDCKP1", 8) != 0) throw std::runtime_error("resume: not a checkpoint: " + path);
    r.p += 8;
    int ne = 0, nf = 0;
    r.pod(ne);
    r.pod(nf);
    if (ne != m_system->num_tet_elements() || nf != m_algorithm_data->num_free_verts())
        throw std::runtime_error("resume: the checkpoint belongs to another mesh");
    m_algorithm_data->load(r);
    m_global_step_behavior->load(r);
    m_system->tet_elements()->load(r);
    if (r.p != r.end) throw std::runtime_error("resume: trailing data in " + path);
    m_system->tet_elements()->refactor();   // the factorization of the restored weights
}

void ADMM::solve(math::MatX3& X) {
    setup_solve(X);
    for (int s = 0; s < m_settings.m_max_admm_iters; ++s) {
        iterate();
        if (converged()) break;
    }
    m_micro_timer.reset();
    finish_solve(X);
    m_algorithm_data->m_runtime.finish_solve_ms += m_micro_timer.elapsed_ms();
}

// ---- Solver (src/Solver.cpp:32-65) ------------------------------------------------------
void Solver::set_pins(const std::vector<int>& inds, const std::vector<math::Vec3>& points) {
    m_system->set_pins(inds, points);
}

bool Solver::initialize(const Settings& settings) {
    m_settings = settings;
    m_system->initialize();
    m_admm.initialize(m_system, m_settings);
    m_system->set_initial_pose();
    initialized = true;
    return true;
}

void Solver::solve() {
    math::MatX3 X = m_system->X();
    m_admm.solve(X);
    m_system->X(X);
}

}  // namespace wrapd
