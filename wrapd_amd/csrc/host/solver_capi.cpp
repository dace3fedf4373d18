// C ABI of the host solver (include/wrapd_solver.h).
#include <cstring>
#include <string>

#include "../../../include/wrapd_solver.h"
#include "wrapd_host.hpp"

struct wrapd_solver {
    std::unique_ptr<wrapd::Solver> solver;
    wrapd::Settings settings;
    std::string err;
    bool initialized = false, setup = false, finished = false;
};

namespace {

template <typename F>
int guarded(wrapd_solver* s, F&& f) {
    if (!s) return WRAPD_ERR_ARG;
    try {
        f();
        return WRAPD_OK;
    } catch (const wrapd::DeviceError& e) {
        s->err = e.what();
        return e.status;
    } catch (const std::exception& e) {
        s->err = e.what();
        return WRAPD_ERR_STATE;
    }
}

}  // namespace

extern "C" {

void wrapd_solver_default_settings(wrapd_solver_settings* s) {
    // Settings / GStep / ModelSettings defaults (src/Settings.hpp:123-162, src/ModelSettings.hpp:33-41)
    std::memset(s, 0, sizeof(*s));
    s->model = 1;
    s->youngs = 1e7;
    s->poisson = 0.45;
    s->beta_static = 1.0;
    s->beta_min = 0.1;
    s->beta_max = 10.0;
    s->rot_aware = 1;
    s->reweighting = 1;
    s->rw_delay = 50;
    s->gamma = 1.5;
    s->tol = 1e-11;
    s->max_admm_iters = 5000;
    s->global_max_iters = 100;
    s->global_ls = 0;
    s->kappa = 1.0;
    s->lbfgs_m = 6;
    s->prox_iters = 20;
    s->precision = WRAPD_FP64;
    s->prox_mode = WRAPD_PROX_LBFGS;
    s->newton_rtol = 0.0;   // precision default
    s->fallback_sigma = 0.2;
    s->fp64_switch = 1e-8;
    s->soln_provided = 0;
    s->verbose = 0;
    s->device = 0;
}

int wrapd_solver_create(wrapd_solver** out, const wrapd_solver_settings* cs, int n_v, const double* rest,
                        const double* init, const double* soln, int n_e, const int* tets, int n_pins, const int* pins,
                        const double* mu_e, const double* lambda_e) {
    if (!out || !cs || !rest || !init || !tets || n_v <= 0 || n_e <= 0) return WRAPD_ERR_ARG;
    wrapd_solver* s = new wrapd_solver();
    *out = s;
    return guarded(s, [&] {
        wrapd::Settings& st = s->settings;
        st.m_model_settings.elastic_model(cs->model == 1 ? "nh" : "arap");
        st.m_model_settings.youngs(cs->youngs);
        st.m_model_settings.poisson(cs->poisson);
        st.m_model_settings.elastic_beta_static(cs->beta_static);
        st.m_model_settings.elastic_beta_min(cs->beta_min);
        st.m_model_settings.elastic_beta_max(cs->beta_max);
        st.m_model_settings.elastic_prox_iters(cs->prox_iters);
        st.m_rot_awareness = cs->rot_aware ? wrapd::Settings::RotAwareness::ENABLED : wrapd::Settings::RotAwareness::DISABLED;
        st.m_reweighting = cs->reweighting ? wrapd::Settings::Reweighting::ENABLED : wrapd::Settings::Reweighting::DISABLED;
        st.m_reweighting_delay = cs->rw_delay;
        st.m_gamma = cs->gamma;
        st.m_tol = cs->tol;
        st.m_max_admm_iters = cs->max_admm_iters;
        st.m_gstep.max_iters(cs->global_max_iters);
        st.m_gstep.linesearch(cs->global_ls ? wrapd::Settings::GStep::LineSearch::Backtracking
                                            : wrapd::Settings::GStep::LineSearch::None);
        st.m_gstep.kappa(cs->kappa);
        st.m_gstep.memory(cs->lbfgs_m);
        st.m_precision = cs->precision;
        st.m_prox_mode = cs->prox_mode;
        st.m_newton_rtol = cs->newton_rtol;
        st.m_fallback_sigma = cs->fallback_sigma;
        st.m_fp64_switch = cs->fp64_switch;
        st.m_io.verbose(cs->verbose);
        if (cs->soln_provided) st.m_io.soln_mesh("given");
        st.m_device = cs->device;
        if (mu_e) st.m_mu_e.assign(mu_e, mu_e + n_e);
        if (lambda_e) st.m_lambda_e.assign(lambda_e, lambda_e + n_e);
        s->solver = std::make_unique<wrapd::Solver>(st);
        wrapd::math::MatX3 X(n_v), Xi(n_v), Xs(n_v);
        std::memcpy(X.data(), rest, sizeof(double) * 3 * n_v);
        std::memcpy(Xi.data(), init, sizeof(double) * 3 * n_v);
        std::memcpy(Xs.data(), soln ? soln : rest, sizeof(double) * 3 * n_v);
        s->solver->m_system->add_nodes(X, Xi, Xs);
        // binding::add_tetmesh: sorted pins, free/fixed maps, create_tets_from_mesh, set_pins
        std::vector<int> sorted(pins, pins + n_pins);
        std::sort(sorted.begin(), sorted.end());
        std::vector<int> free_list(n_v), fix_list(n_v, -1);
        int j = 0, k = 0;
        for (int i = 0; i < n_v; ++i) {
            if (j < n_pins && i == sorted[j]) { free_list[i] = -1; fix_list[i] = j; ++j; }
            else { free_list[i] = k; fix_list[i] = -1; ++k; }
        }
        wrapd::create_tets_from_mesh<double, wrapd::TetElements>(s->solver->m_system->tet_elements(), rest, tets, n_e,
                                                                  st, free_list, fix_list, 0);
        s->solver->set_pins(sorted);
    });
}

int wrapd_solver_destroy(wrapd_solver* s) {
    delete s;
    return WRAPD_OK;
}

const char* wrapd_solver_last_error(wrapd_solver* s) { return s ? s->err.c_str() : "null solver"; }

int wrapd_solver_initialize(wrapd_solver* s) {
    return guarded(s, [&] {
        s->solver->initialize(s->settings);
        s->initialized = true;
    });
}

int wrapd_solver_setup(wrapd_solver* s) {
    return guarded(s, [&] {
        if (!s->initialized) throw wrapd::DeviceError(WRAPD_ERR_STATE, "initialize first");
        wrapd::math::MatX3 X = s->solver->m_system->X();
        s->solver->admm().setup_solve(X);
        s->setup = true;
    });
}

int wrapd_solver_iterate(wrapd_solver* s, int* conv) {
    return guarded(s, [&] {
        if (!s->setup) throw wrapd::DeviceError(WRAPD_ERR_STATE, "setup first");
        s->solver->admm().iterate();
        const bool c = s->solver->admm().converged();
        if (conv) *conv = c ? 1 : 0;
    });
}

int wrapd_solver_checkpoint(wrapd_solver* s, const char* path) {
    return guarded(s, [&] {
        if (!s->setup) throw wrapd::DeviceError(WRAPD_ERR_STATE, "setup first");
        if (!path) throw wrapd::DeviceError(WRAPD_ERR_ARG, "checkpoint: no path");
        s->solver->admm().checkpoint(path);
    });
}

int wrapd_solver_resume(wrapd_solver* s, const char* path) {
    return guarded(s, [&] {
        if (!s->setup) throw wrapd::DeviceError(WRAPD_ERR_STATE, "setup first");
        if (!path) throw wrapd::DeviceError(WRAPD_ERR_ARG, "resume: no path");
        s->solver->admm().resume(path);
    });
}

int wrapd_solver_finish(wrapd_solver* s) {
    return guarded(s, [&] {
        wrapd::math::MatX3 X = s->solver->m_system->X();
        s->solver->admm().finish_solve(X);
        s->solver->m_system->X(X);
        s->finished = true;
    });
}

int wrapd_solver_solve(wrapd_solver* s) {
    return guarded(s, [&] {
        if (!s->initialized) throw wrapd::DeviceError(WRAPD_ERR_STATE, "initialize first");
        s->solver->solve();
        s->finished = true;
    });
}

int wrapd_solver_num_iters(wrapd_solver* s) {
    if (!s || !s->solver || !s->solver->algorithm_data()) return -1;
    return s->solver->algorithm_data()->m_iter;
}

int wrapd_solver_get_log(wrapd_solver* s, const char* name, double* out, int max_len) {
    if (!s || !s->solver || !s->solver->algorithm_data() || !name) return -1;
    const wrapd::AlgorithmData* ad = s->solver->algorithm_data();
    const std::string n(name);
    auto cp = [&](const auto& v) {
        const int m = std::min(max_len, static_cast<int>(v.size()));
        for (int i = 0; i < m && out; ++i) out[i] = static_cast<double>(v[i]);
        return static_cast<int>(v.size());
    };
    if (n == "objectives") return cp(ad->per_iter_objectives());
    if (n == "reweighted") return cp(ad->per_iter_reweighted());
    if (n == "inner_iters") return cp(ad->per_iter_inner_iters());
    if (n == "ls_dal") return cp(ad->m_per_iter_ls_delta_auglag);
    if (n == "gs_dal") return cp(ad->m_per_iter_gs_delta_auglag);
    if (n == "flips") return cp(ad->per_iter_flips());
    if (n == "x_errors") return cp(ad->per_iter_state_x_errors());
    if (n == "residuals") return cp(ad->m_norm_obj_residuals);
    if (n == "accumulated_time_s") return cp(ad->per_iter_accumulated_time_s());
    if (n == "wall_time_s") return cp(ad->per_iter_wall_time_s());
    return -1;
}

int wrapd_solver_get_x(wrapd_solver* s, double* X) {
    if (!s || !s->solver || !X) return WRAPD_ERR_ARG;
    const wrapd::math::MatX3& x = s->solver->m_system->X();
    std::memcpy(X, x.data(), sizeof(double) * 3 * x.rows());
    return WRAPD_OK;
}

int wrapd_solver_get_x_free(wrapd_solver* s, double* x) {
    return guarded(s, [&] {
        wrapd::math::MatX3 xf;
        s->solver->m_system->tet_elements()->get_x_free(xf);
        std::memcpy(x, xf.data(), sizeof(double) * 3 * xf.rows());
    });
}

int wrapd_solver_get_runtime(wrapd_solver* s, double* o) {
    if (!s || !s->solver || !s->solver->algorithm_data() || !o) return WRAPD_ERR_ARG;
    const wrapd::RuntimeData& r = s->solver->algorithm_data()->m_runtime;
    o[0] = r.setup_solve_ms;
    o[1] = r.local_ms;
    o[2] = r.global_ms;
    o[3] = r.refactor_ms;
    o[4] = r.inner_iters;
    o[5] = r.finish_solve_ms;
    o[6] = r.prox_iters;
    o[7] = r.fallback_elems;
    o[8] = r.learned_accepted;
    o[9] = r.fp64_switch_iter;
    o[10] = r.learned_used;
    return WRAPD_OK;
}

wrapd_ctx* wrapd_solver_ctx(wrapd_solver* s) {
    if (!s || !s->solver) return nullptr;
    auto& te = s->solver->m_system->tet_elements();
    return te ? te->ctx() : nullptr;
}

}  // extern "C"
