/* wrapd_solver.h - C ABI of the host-side WRAPD solver (the reference's Solver/ADMM
 * driven as samples/tests/quasi.cpp:63-97 drives them), over the kernels of wrapd_gpu.h.
 * Used by the Python tests and the benchmark; the `quasi` executable links the same code.
 */
#ifndef WRAPD_SOLVER_H_
#define WRAPD_SOLVER_H_

#include "wrapd_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wrapd_solver wrapd_solver;

typedef struct {
    int model;                /* 0 ARAP, 1 NH                                   (-model)      */
    double youngs, poisson;   /* Young's modulus (reference: fixed 1e7), Poisson (-poisson)   */
    double beta_static, beta_min, beta_max;
    int rot_aware;            /* -ra */
    int reweighting;          /* -rw */
    double rw_delay;          /* -rw_delay */
    double gamma;             /* -gamma */
    double tol;               /* -tol */
    int max_admm_iters;       /* -it */
    int global_max_iters;     /* -global_it */
    int global_ls;            /* -global_ls */
    double kappa;             /* -kappa */
    int lbfgs_m;              /* SplitLBFGS memory (reference: 6) */
    int prox_iters;           /* prox iteration cap (reference: 20) */
    int precision;            /* WRAPD_FP64 / WRAPD_FP32 */
    int prox_mode;            /* WRAPD_PROX_* */
    double newton_rtol;
    double fallback_sigma;
    double fp64_switch;       /* WRAPD_FP32: FP64 local steps once |dE|/(|E|+1) < this (sticky) */
    int soln_provided;
    int verbose;
    int device;
} wrapd_solver_settings;

void wrapd_solver_default_settings(wrapd_solver_settings* s);
/* Solver(settings) + binding::add_tetmesh (samples/utils/AddMeshes.hpp:61-132); soln may be
 * NULL (the rest mesh is used, as quasi does); mu_e/lambda_e may be NULL. */
int wrapd_solver_create(wrapd_solver** out, const wrapd_solver_settings* s, int n_verts, const double* rest,
                        const double* init, const double* soln, int n_elems, const int* tets, int n_pins,
                        const int* pins, const double* mu_e, const double* lambda_e);
int wrapd_solver_destroy(wrapd_solver* s);
const char* wrapd_solver_last_error(wrapd_solver* s);
int wrapd_solver_initialize(wrapd_solver* s);                 /* Solver::initialize          */
int wrapd_solver_setup(wrapd_solver* s);                      /* ADMM::solve up to the loop  */
int wrapd_solver_iterate(wrapd_solver* s, int* converged);    /* ADMM::iterate + converged   */
int wrapd_solver_finish(wrapd_solver* s);                     /* ADMM::finish_solve          */
int wrapd_solver_solve(wrapd_solver* s);                      /* Solver::solve (whole loop)  */
int wrapd_solver_num_iters(wrapd_solver* s);
/* Checkpoint/resume (no counterpart in the reference, which lists it as an extension):
 * checkpoint writes the whole ADMM state after the last iterate to a file; resume, on a
 * solver created and set up with the same mesh and settings, loads it (and refactors) so
 * that the following iterations continue the checkpointed run bit for bit. */
int wrapd_solver_checkpoint(wrapd_solver* s, const char* path);
int wrapd_solver_resume(wrapd_solver* s, const char* path);
/* logs: objectives, reweighted, inner_iters, ls_dal, gs_dal, flips, x_errors, residuals,
 * accumulated_time_s, wall_time_s.  Returns the length (writes at most max_len). */
int wrapd_solver_get_log(wrapd_solver* s, const char* name, double* out, int max_len);
int wrapd_solver_get_x(wrapd_solver* s, double* X);           /* (n_verts,3) after finish     */
int wrapd_solver_get_x_free(wrapd_solver* s, double* x);      /* current free positions       */
/* [setup_ms, local_ms, global_ms, refactor_ms, inner_iters, finish_ms, prox_iters, fallback,
 *  learned_accepted, fp64_switch_iter (-1: none), learned_used] */
int wrapd_solver_get_runtime(wrapd_solver* s, double* out11);
wrapd_ctx* wrapd_solver_ctx(wrapd_solver* s);

#ifdef __cplusplus
}
#endif
#endif /* WRAPD_SOLVER_H_ */
