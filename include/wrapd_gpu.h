/* wrapd_gpu.h - C ABI of the MI355X WRAPD kernels (one GPU, one HIP stream per context).
 *
 * Every call returns a status code (WRAPD_OK = 0); nothing throws or exits across this
 * boundary.  A context owns all device memory of one solve and is confined to the
 * host thread that created it.  Arrays are plain host pointers:
 *   positions (n,3) row-major doubles, tets (n_e,4) int32, 3x3 matrices row-major.
 * Free/fixed vertex numbering follows samples/utils/AddMeshes.hpp:89-110: pins are
 * sorted, free vertices are numbered in vertex order, fixed ones in pin order.
 */
#ifndef WRAPD_GPU_H_
#define WRAPD_GPU_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    WRAPD_OK = 0,
    WRAPD_ERR_ARG = 1,            /* bad argument / size                                   */
    WRAPD_ERR_HIP = 2,            /* HIP runtime error                                      */
    WRAPD_ERR_INVERTED_REST = 3,  /* rest element with negative volume (src/TetElements.cpp:99-101) */
    WRAPD_ERR_NOT_SPD = 4,        /* non-positive pivot in the factorization (singular A)   */
    WRAPD_ERR_NAN = 5,            /* NaN L-BFGS decrement (SplitLBFGS.hpp:181-184)          */
    WRAPD_ERR_NEG_SIGMA = 6,      /* negative prox singular value (src/TetElements.cpp:320-323) */
    WRAPD_ERR_STATE = 7,          /* call out of order                                      */
    WRAPD_ERR_NO_DEVICE = 8,      /* no HIP device                                          */
    WRAPD_ERR_LINESEARCH = 9      /* line search failure (SplitLBFGS.hpp:200-203)            */
};

typedef struct wrapd_ctx wrapd_ctx;

typedef struct {
    int model;                  /* 0 = ARAP, 1 = Neo-Hookean (src/ModelSettings.hpp:43-46)   */
    double mu, lambda;          /* Lame parameters used when no per-element arrays are given */
    double beta_static;         /* w^2 = beta_static k vol (src/TetElements.cpp:139)          */
    double beta_min, beta_max;  /* reweighting clamp (src/TetElements.cpp:399-402)            */
} wrapd_material;

enum { WRAPD_FP64 = 0, WRAPD_FP32 = 1 };                               /* local-step precision */
enum { WRAPD_PROX_LBFGS = 0, WRAPD_PROX_NEWTON = 1, WRAPD_PROX_LEARNED = 2 };

typedef struct {
    int precision;              /* WRAPD_FP64 | WRAPD_FP32 (FP32 with per-element FP64 fallback) */
    int prox_mode;              /* LBFGS = reference solver (src/TetElements.cpp:302-318)        */
    int prox_iters;             /* L-BFGS cap (reference: 20) or Newton cap                      */
    int update_cand;            /* compute candidate weights (src/TetElements.cpp:351-353)       */
    int cand_use_hessian;       /* reweighting counter past the delay (src/TetElements.cpp:405)  */
    double newton_rtol;         /* Newton tolerance, relative to (mu + w^2/vol) max(1,|s|)       */
    double fallback_sigma;      /* FP32: elements with s_hat[2] below this redo the step in FP64 */
} wrapd_local_args;

typedef struct {
    double delta_auglag;        /* aft - bef + pe - pe_prev (src/TetElements.cpp:355)        */
    double bef, aft, pe;
    double max_wsq_ratio;       /* Elements::max_wsq_ratio (src/Elements.hpp:56-69)          */
    double prox_iters_total;    /* prox iterations summed over elements                      */
    double fallback_count;      /* FP32 elements redone in FP64                               */
    double neg_sigma_count;     /* elements whose prox returned a negative singular value    */
    double learned_accepted;    /* learned warm starts accepted without a Newton iteration   */
    double evals_total;         /* prox objective evaluations (line-search trials / Newton steps) */
    double evals_max;           /* largest per-element count                                 */
    double learned_used;        /* elements whose Newton prox started from the learned guess */
} wrapd_local_out;

typedef struct {
    int memory;                 /* history M (reference: 6, SplitLBFGS.hpp:41)               */
    int max_iters;              /* -global_it (src/Settings.hpp:124)                          */
    double kappa;               /* -kappa                                                     */
    int linesearch;             /* 0 none, 1 Armijo backtracking (-global_ls)                 */
} wrapd_lbfgs_cfg;

typedef struct {
    int iters;                  /* value returned by SplitLBFGS::minimize                     */
    double obj_bef, obj_aft;    /* GlobalProblem::value before/after (GlobalStepBehavior.hpp:247,256) */
    double energy;              /* potential_energy_x at the final x (src/TetElements.cpp:598-635) */
    int flips;                  /* inversion_count (src/TetElements.cpp:444-468)              */
    int exit_reason;            /* 0 loop end, 1 kappa early exit, 2 zero denominator, 3 no change */
} wrapd_global_out;

/* ---- lifecycle -------------------------------------------------------------------- */
int wrapd_create(wrapd_ctx** ctx, int device);
int wrapd_destroy(wrapd_ctx* ctx);
const char* wrapd_last_error(wrapd_ctx* ctx);

/* TetElements constructor (src/TetElements.cpp:36-143) fed as create_tets_from_mesh
 * (src/TetElements.hpp:250-288) and binding::add_tetmesh (samples/utils/AddMeshes.hpp:61-132)
 * would: precomputes Dm^-1, volumes, D_local, w^2 on the device, builds the
 * vertex->element gather lists and the pattern of A.  mu_e/lambda_e may be NULL. */
int wrapd_upload_mesh(wrapd_ctx* ctx, int n_verts, const double* rest, int n_elems, const int* tets,
                      int n_pins, const int* pins, const wrapd_material* mat, const double* mu_e,
                      const double* lambda_e);
int wrapd_sizes(wrapd_ctx* ctx, int* n_verts, int* n_elems, int* n_free, int* n_fix);

/* TetElements::update_fix_cache (src/TetElements.cpp:547-562): x_fix (n_fix,3) */
int wrapd_set_x_fix(wrapd_ctx* ctx, const double* x_fix);
/* free positions (n_free,3) in and out (ADMM m_curr_x_free) */
int wrapd_set_x_free(wrapd_ctx* ctx, const double* x_free);
int wrapd_get_x_free(wrapd_ctx* ctx, double* x_free);

/* |x_free - x_ref| for the x-error log (src/AlgorithmData.cpp:63, 70); x_ref (n_free,3) */
int wrapd_set_x_ref(wrapd_ctx* ctx, const double* x_ref);
int wrapd_x_distance(wrapd_ctx* ctx, double* dist);
/* The same in two halves (no reference counterpart): the launch returns at once, the result
 * waits for that launch's reduction.  The ADMM driver logs the distance this way so that the
 * next iteration's kernels are queued before the host waits for it. */
int wrapd_x_distance_launch(wrapd_ctx* ctx);
int wrapd_x_distance_result(wrapd_ctx* ctx, double* dist);

/* Settings::m_rot_awareness (src/Settings.hpp:107-111): selects what the deformation cache
 * holds (the symmetric polar factor, or F^T for the rotation-unaware local step). */
int wrapd_set_rot_aware(wrapd_ctx* ctx, int rot_aware);

/* TetElements::initialize_dual_vars (src/TetElements.cpp:145-183) at the current positions;
 * returns polar_potential_energy() for rot_aware = 1 (src/TetElements.cpp:638-675). */
int wrapd_init_duals(wrapd_ctx* ctx, int rot_aware, double* polar_pe);

/* update_defo_cache + potential_energy_x + inversion_count at the current x
 * (src/TetElements.cpp:564-635, 444-468); also refreshes the cached symmetric factor. */
int wrapd_defo_energy(wrapd_ctx* ctx, double* energy, int* flips);

/* update_all_candidate_weights (src/TetElements.cpp:391-432) from sigma(F(x))
 * (midupdate_weights, :360-388); use_hessian = reweighting counter past the delay;
 * direct = 1 forms F from all four corners at once as midupdate_weights(X) does
 * (:366-373), direct = 0 as the deformation cache does (:567-587). */
int wrapd_candidate_weights(wrapd_ctx* ctx, int use_hessian, int direct);
int wrapd_max_wsq_ratio(wrapd_ctx* ctx, double* ratio);
/* update_actual_weights (src/TetElements.cpp:435-442) */
int wrapd_update_actual_weights(wrapd_ctx* ctx);

/* TetElements::polar_update (src/TetElements.cpp:256-356): local step on all elements. */
int wrapd_local_step(wrapd_ctx* ctx, const wrapd_local_args* args, double pe_prev, wrapd_local_out* out);
/* The same in two calls: the launch, then its results (which wait for it).  Between them the
 * caller may queue wrapd_global_step_begin, so the GPU works while the host reads the local
 * step's results.  Results without a launch before them: WRAPD_ERR_STATE. */
int wrapd_local_step_launch(wrapd_ctx* ctx, const wrapd_local_args* args);
int wrapd_local_step_result(wrapd_ctx* ctx, double pe_prev, wrapd_local_out* out);
/* TetElements::update (src/TetElements.cpp:193-254): rotation-unaware local step (-ra 0). */
int wrapd_local_step_unaware(wrapd_ctx* ctx, const wrapd_local_args* args);

/* global_obj_value / global_obj_grad (src/TetElements.cpp:678-787) at the current x;
 * grad (n_free,3) may be NULL. */
int wrapd_global_value_grad(wrapd_ctx* ctx, double* value, double* grad);

/* System::get_A / update_A (src/System.cpp:78-100) + LinearSolver::compute/factorize
 * (deps/mcloptlib/include/MCL/LinearSolver.hpp:288-292): assembles A on the device,
 * runs the host symbolic analysis once, then the numeric factorization. */
int wrapd_refactor(wrapd_ctx* ctx);
/* LinearSolver::solve for the three coordinate columns (SplitLBFGS.hpp:159-171):
 * b, x (n_free,3) host arrays. */
int wrapd_solve3(wrapd_ctx* ctx, const double* b, double* x);
/* LinearSolver::compute / factorize (deps/mcloptlib/include/MCL/LinearSolver.hpp:279-298) for
 * a matrix the caller gives, on a context without a mesh.  analyze: the pattern of the upper
 * triangle as CSR (row i holds sorted columns j >= i, the diagonal first), n <= 2^31; xyz
 * (n,3) coordinates of the unknowns for a geometric nested dissection, or NULL for one by
 * breadth-first level structures of the matrix graph.  factorize: the values in that
 * pattern's order; WRAPD_ERR_NOT_SPD if A is not positive definite.  wrapd_solve3 then
 * solves with three right-hand sides (n,3). */
int wrapd_linsolve_analyze(wrapd_ctx* ctx, int n, const int* row_ptr, const int* cols, const double* xyz);
int wrapd_linsolve_factorize(wrapd_ctx* ctx, const double* vals);
/* A values (upper triangle, CSR over free vertices), for checking. Returns nnz. */
int wrapd_get_A(wrapd_ctx* ctx, int* row_ptr, int* cols, double* vals);
/* factorization statistics: [nsuper, nnz(L), Cholesky flops, max front, solve leaf tasks,
 * solve levels, G entries, factor levels, factor launches, factor tile flops (as issued,
 * padding included), front pool doubles, G formation flops (L11^{-1} and -L21 L11^{-1})] */
int wrapd_chol_stats(wrapd_ctx* ctx, double* stats12);

/* SplitLBFGS (deps/mcloptlib/include/MCL/SplitLBFGS.hpp:41-257) */
int wrapd_lbfgs_configure(wrapd_ctx* ctx, const wrapd_lbfgs_cfg* cfg);
int wrapd_lbfgs_reset(wrapd_ctx* ctx);              /* update_A0 / clear_history: m_k = 0 */
/* RotAwareGS::global_step (src/GlobalStepBehavior.hpp:246-258): value, minimize with the
 * early-exit tolerance |dAL_local|, value; energy and flips at the final x. */
int wrapd_global_step(wrapd_ctx* ctx, double early_exit_tol, wrapd_global_out* out);
/* Queues the global step's first value and gradient (which do not depend on the
 * tolerance); the next wrapd_global_step continues from them instead of launching them. */
int wrapd_global_step_begin(wrapd_ctx* ctx);
/* UnawareGS::global_step (src/GlobalStepBehavior.hpp:164-184): x = A^-1 b with
 * b from add_to_b (src/TetElements.cpp:470-481). */
int wrapd_global_step_unaware(wrapd_ctx* ctx, wrapd_global_out* out);

/* TetElements::add_to_b (src/TetElements.cpp:470-481): b = sum_e w_e^2 D_e^T (Z_e - U_e - Dfix x_fix)
 * over the free vertices, (n_free,3), from the duals as they are (the right-hand side of the
 * rotation-unaware global step). */
int wrapd_add_to_b(wrapd_ctx* ctx, double* b);

/* element state, for checking: "U","Z","wsq","cand_wsq","sigma_last","vol","Di","symF"
 * with (n_e,9)/(n_e)/(n_e,3)/(n_e,12) row-major shapes. */
int wrapd_get_elem(wrapd_ctx* ctx, const char* name, double* out);

/* Checkpoint/resume of the device-resident ADMM state: element duals and weights (U, Z-U,
 * w^2, candidate w^2, cached F, last prox solutions), positions, the L-BFGS history with
 * its Gram entries and gradients, and the L-BFGS bookkeeping.  No counterpart in the
 * reference (checkpoint/resume is one of its listed extensions); for callers that stop and
 * later continue a run bit for bit.  The mesh, material and settings are not included:
 * load into a context set up with the same ones.  The factorization is not included
 * either: call wrapd_refactor after a load (it depends only on w^2).
 * save: with buf == NULL only *size (bytes) is set; otherwise cap must be >= that size. */
int wrapd_state_save(wrapd_ctx* ctx, void* buf, size_t cap, size_t* size);
int wrapd_state_load(wrapd_ctx* ctx, const void* buf, size_t size);
int wrapd_set_elem(wrapd_ctx* ctx, const char* name, const double* in);
/* math::svd (src/Math.hpp:140-161) as the device computes it, for checking: n row-major
 * 3x3 matrices F in, sigma (n,3) sorted by magnitude with det's sign on sigma[2], U and V
 * (n,9) row-major rotations out; precision WRAPD_FP64 or WRAPD_FP32 (the FP32 local step's
 * arithmetic, inputs rounded to float). */
int wrapd_svd3_batch(wrapd_ctx* ctx, const double* F, int n, int precision, double* sigma, double* U, double* V);

/* timing of the device work, measured with HIP events on the context's stream */
int wrapd_profile_enable(wrapd_ctx* ctx, int on);
/* only = NULL or "": every scope; otherwise the one scope named (each timed launch costs two
 * event records in the stream, which slow a host-driven loop down noticeably) */
int wrapd_profile_only(wrapd_ctx* ctx, const char* only);
int wrapd_profile_get(wrapd_ctx* ctx, const char* kernel, double* total_ms, double* launches);
int wrapd_synchronize(wrapd_ctx* ctx);

/* ---- learned prox warm start ----------------------------------------------------------
 * An extension of the stretch prox (src/TetElements.cpp:302-318, solved there by
 * LBFGS<double,3,8>): no counterpart in the reference.  An MLP 6 -> 64 -> 64 -> 3 (GELU,
 * BF16 matrix-core products, FP32 accumulation) maps (sorted s_hat, w^2/vol, mu, lambda)
 * to a guess of the prox solution; with WRAPD_PROX_LEARNED the local step starts its
 * Newton prox from that guess when its optimality residual beats the previous solution's.
 * Parameters: 4803 floats W1[64][6] b1[64] W2[64][64] b2[64] W3[3][64] b3[3];
 * normalisation: 18 floats in_mean[6] in_std[6] out_mean[3] out_std[3] (features s0 s1 s2
 * log(w^2/vol/mu) log(lambda/mu) log(mu); outputs x - s_hat).
 * Tuples: 9 floats s0 s1 s2 w^2/vol mu lambda | x0 x1 x2. */
int wrapd_mlp_init(wrapd_ctx* ctx, unsigned long long seed);            /* Glorot weights, identity norm */
int wrapd_mlp_set_params(wrapd_ctx* ctx, const float* params4803, const float* norm18);
int wrapd_mlp_get_params(wrapd_ctx* ctx, float* params4803, float* norm18);
/* in: n x 6 (s0 s1 s2 w^2/vol mu lambda), out: n x 3 predicted prox solutions (host arrays) */
int wrapd_mlp_infer(wrapd_ctx* ctx, const float* in, int n, float* out);
/* mean squared error (standardised targets) and its gradient on the n given tuples */
int wrapd_mlp_grad(wrapd_ctx* ctx, const float* tuples, int n, float* grad4803, float* loss);
/* wrapd_mlp_grad without device allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same result, bit for bit, as wrapd_mlp_grad on the
 * same tuples; the tuples, sample list, per-wave partial gradients and sums are staged in a
 * static 8 MiB device array of the library instead of memory allocated per call, so at most
 * 65536 tuples per call (WRAPD_ERR_ARG beyond).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (from several contexts or threads) race. */
int wrapd_mlp_grad_noalloc(wrapd_ctx* ctx, const float* tuples, int n, float* grad4803, float* loss);
/* tuple log: while started, every NH local step appends its n_e (input, solution) tuples */
int wrapd_mlp_log_start(wrapd_ctx* ctx, long long capacity);
int wrapd_mlp_log_stop(wrapd_ctx* ctx);
long long wrapd_mlp_log_count(wrapd_ctx* ctx);
int wrapd_mlp_log_get(wrapd_ctx* ctx, long long offset, long long n, float* out);
int wrapd_mlp_log_set(wrapd_ctx* ctx, const float* tuples, long long n);
/* normalisation from the logged tuples (a strided sample of at most 1e6) */
int wrapd_mlp_fit_norm(wrapd_ctx* ctx);
/* w^2/vol redrawn log-uniformly over [w_lo, w_hi] (<= 0: the logged range) and the
 * targets re-solved by the FP64 Newton prox */
int wrapd_mlp_augment(wrapd_ctx* ctx, double w_lo, double w_hi, unsigned long long seed);
/* full-batch Adam steps on the n given tuples, in order (the Adam state continues from the
 * previous steps; wrapd_mlp_init / set_params restart it); loss_hist (steps floats) may be NULL */
int wrapd_mlp_train_full(wrapd_ctx* ctx, const float* tuples, int n, int steps, float lr, float* loss_hist);
/* Adam steps on minibatches drawn from the log; loss_hist (steps floats) may be NULL */
int wrapd_mlp_train(wrapd_ctx* ctx, int steps, int batch, float lr, unsigned long long seed, float* loss_hist);

#ifdef __cplusplus
}
#endif
#endif /* WRAPD_GPU_H_ */
