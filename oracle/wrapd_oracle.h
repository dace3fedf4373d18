/* Oracle: CPU FP64 restatement of the reference WRAPD ADMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or the CPU
 * baseline.  The product (wrapd_amd/, include/wrapd_gpu.h) never links it.
 *
 * Conventions: 3x3 matrices row-major (M[3*r+c]); positions (n,3) row-major;
 * element state arrays are (n_e, 9) / (n_e, 3) / (n_e) row-major.
 */
#ifndef WRAPD_ORACLE_H_
#define WRAPD_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int model;            /* 0 = ARAP, 1 = NH            (src/ModelSettings.hpp:33-41) */
    double mu, lambda;    /* global Lame parameters      (src/Lame.hpp:50-53)          */
    double beta_static, beta_min, beta_max;
    int rot_aware;        /* -ra                          (src/Settings.hpp:154-162)   */
    int reweighting;      /* -rw                                                        */
    double rw_delay;      /* -rw_delay                                                  */
    double gamma;         /* -gamma                                                     */
    double tol;           /* -tol                                                       */
    int max_admm_iters;   /* -it                                                        */
    int global_max_iters; /* -global_it                   (src/Settings.hpp:123-127)   */
    int global_ls;        /* -global_ls: 0 none, 1 Armijo backtracking                  */
    double kappa;         /* -kappa                                                     */
    int lbfgs_m;          /* SplitLBFGS memory (reference: compile-time 6)              */
    int prox_iters;       /* prox L-BFGS iteration cap (reference: 20)                  */
    int soln_provided;    /* -soln given: enables the x-error early exit                */
} oracle_settings_t;

void oracle_default_settings(oracle_settings_t* s);

/* ---- unit functions ---------------------------------------------------------- */
void oracle_svd3(const double* F, double* sigma, double* U, double* V);
double oracle_nh_psi(const double* s, double mu, double lambda);          /* energy_density */
void oracle_nh_hess(const double* s, double mu, double lambda, double* H); /* after makePD */
double oracle_hess_lambda_max(const double* s, double mu, double lambda);
int oracle_prox_lbfgs(const double* sig_hat, double wsq_over_vol, double mu, double lambda,
                      int max_iters, double* x_inout);
void oracle_sym_eig3(const double* A, double* evals /* ascending */);

/* ---- solver -------------------------------------------------------------------- */
void* oracle_create(const oracle_settings_t* s, int n_v, const double* rest, const double* init,
                    const double* soln, int n_e, const int* tets, int n_pins, const int* pins,
                    const double* mu_e, const double* lambda_e);
void oracle_destroy(void* h);
int oracle_initialize(void* h);          /* ADMM::initialize                               */
int oracle_setup_elements(void* h);   /* element state at X_init, no linear system */
int oracle_setup_solve(void* h);         /* ADMM::solve up to the iteration loop           */
int oracle_iterate(void* h);             /* one ADMM::iterate; returns 1 when converged    */
int oracle_finish(void* h);              /* ADMM::finish_solve                             */
int oracle_solve(void* h);               /* setup + loop + finish; returns iterations done */
int oracle_num_iters(void* h);
int oracle_n_free(void* h);
/* logs: "objectives","reweighted","inner_iters","ls_dal","gs_dal","flips","x_errors",
 *       "residuals","accumulated_time_s" */
int oracle_get_log(void* h, const char* name, double* out, int max_len);
void oracle_get_x(void* h, double* X);                   /* (n_v,3) current positions    */
void oracle_get_x_free(void* h, double* x);              /* (n_f,3)                      */
void oracle_set_x_free(void* h, const double* x);
/* element state: "U","Z","wsq","cand_wsq","sigma_last","vol","Di"(12, row-major 3x4),
 * "F","symF","sigma","mu","lambda" */
int oracle_get_elem(void* h, const char* name, double* out);
int oracle_set_elem(void* h, const char* name, const double* in);
double oracle_get_scalar(void* h, const char* name);     /* "curr_pe", "last_ls_dal", ... */
void oracle_set_scalar(void* h, const char* name, double v);
double oracle_local_step(void* h, int update_cand);      /* TetElements::polar_update    */
void oracle_update_defo_cache(void* h);                  /* at the current x_free        */
double oracle_global_value(void* h, const double* x_free);
double oracle_global_grad(void* h, const double* x_free, double* grad);
double oracle_potential_energy(void* h);                 /* potential_energy_x (cache)   */
int oracle_inversion_count(void* h);
/* A (upper triangle, free vertices) as COO sorted by (row, col); returns nnz */
int oracle_get_A(void* h, int* rows, int* cols, double* vals, int max_nnz);
/* solves A x = b for 3 columns stored (n_f,3) row-major with the current factor */
int oracle_solve3(void* h, const double* b, double* x);

#ifdef __cplusplus
}
#endif
#endif /* WRAPD_ORACLE_H_ */
