// Internal: the device context behind wrapd_ctx and the launchers shared by the .hip files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/wrapd_gpu.h"
#include "chol_plan.hpp"
#include "device/wd_mlp.hpp"
#include "device/wd_reduce.hpp"
#include "chol_symbolic.hpp"

namespace wd {

constexpr int kElemThreads = 256;     // element kernels
constexpr int kLocalThreads = 256;    // local step (register heavy: one wave per SIMD either way; 256 keeps C2's grid under the in-launch reduction limit)
constexpr int kRedStride = wd::kTailSlots;   // reduction slots per block in `partials`
constexpr int kMaxHist = 8;           // largest L-BFGS memory supported on the device
constexpr int kMaxBlocks = wd::kTailBlocks;   // capacity of the partials buffer (blocks)

// device scalar slots (ctx->scal)
enum ScalarSlot {
    S_RED = 0,          // 0..31: result of the last reduction
    S_ALPHA = 32,       // 8: two-loop alpha (logical order)
    S_COEF = 40,        // 8: alpha - beta
    S_DIR = 48, S_DENOM = 49, S_DG = 50, S_VAL = 51, S_EN = 52, S_FLIPS = 53, S_GTP = 54, S_GMAX = 55,
    S_GS = 56,          // 8: g . s_j of the current two-loop (logical order), kept by its alpha step
    S_GRAM = 64,        // 64: s_i . y_j by slot (8 x 8)
    S_XD = 128,         // |x - x_ref|^2 of the last x-distance launch
    S_LOCAL = 136,      // 11: the last local step's sums and maxima
    S_LOCAL_FB = 152,   // 11: the same over its FP64 fallback elements (FP32 local steps)
    S_NUM = 168
};

// defo-pass mode bits
enum DefoMode {
    DM_WRITE_SYMF = 1, DM_VALUE = 2, DM_GRAD = 4, DM_ENERGY = 8, DM_CAND = 16, DM_DIRECT = 32,
    DM_FT_CACHE = 64, DM_BVEC = 128
};

struct Prof {
    bool on = false;
    std::string only;   // the one scope timed, or empty: all
    struct Pending { std::string name; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<double, double>> acc;   // ms, launches
};

}  // namespace wd

struct wrapd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // the factorization's second stream (FINAL launches of supernodes whose panels end before
    // their level's last panel step) and the events that order it with `stream`
    hipStream_t fstream2 = nullptr;
    hipEvent_t fev_fork = nullptr, fev_join = nullptr;
    std::vector<hipEvent_t> fev_final;   // after a level's FINAL launch k on fstream2
    std::string last_error;
    wd::Prof prof;

    // sizes
    int n_v = 0, n_e = 0, n_ep = 0, n_f = 0, n_fix = 0;
    bool mesh_ready = false;
    bool linsolve = false;      // a matrix-only context (wrapd_linsolve_analyze): A given by the caller
    bool dead = false;          // a device wait timed out: every entry point but destroy refuses
    // host copies
    std::vector<int> tets, pins, free_idx, fix_idx;
    std::vector<double> rest;
    wrapd_material mat{};
    bool hetero = false;

    // element arrays (structure of arrays, stride n_ep)
    int* vidx = nullptr;        // int4 per element: index into xall (free: [0,n_f), fixed: n_f + j)
    double* Di = nullptr;       // 12 x n_ep: Di[(r*4+k)*n_ep + e] = D_local(r,k)
    double* vol = nullptr;
    double* kst = nullptr;      // material stiffness k_e
    double* mu_e = nullptr;     // optional
    double* lam_e = nullptr;    // optional
    double* wsq = nullptr;
    double* cand = nullptr;
    double* U = nullptr;        // 9 x n_ep
    double* P = nullptr;        // 9 x n_ep, P = Z - U
    double* symF = nullptr;     // 9 x n_ep (Ft for the rotation-unaware mode)
    double* sig = nullptr;      // 3 x n_ep last prox solution
    double* gstage = nullptr;   // corner gradients (3 doubles) of the free corners, vertex-major: in v2e order

    // vertex vectors (n_f x 3 AoS)
    double* xall = nullptr;     // (n_f + n_fix) x 3
    double* grad = nullptr;
    double* grad_old = nullptr;
    double* qv = nullptr;
    double* rv = nullptr;
    double* bv = nullptr;
    double* xref = nullptr;     // reference positions for the x-error log
    double* hS[wd::kMaxHist] = {};
    double* hY[wd::kMaxHist] = {};
    int* v2e_ptr = nullptr;     // free vertex v's corners are gstage rows [v2e_ptr[v], v2e_ptr[v+1]), element order
    int* cpos = nullptr;        // int4 per element: each corner's gstage row (-1: a fixed vertex)

    // A
    std::vector<int> Ap, Aj;    // upper CSR over free vertices (host)
    int nnzA = 0;
    int* a_cptr = nullptr;      // A slot -> contributions (e*16 + a*4 + b)
    int* a_cidx = nullptr;
    double* Ax = nullptr;

    // Cholesky
    wrapd::CholSymbolic sym;
    bool analyzed = false, factored = false;
    double* G = nullptr;
    double* ysol = nullptr;     // n_f x 3 forward result (new numbering)
    double* xsol = nullptr;     // n_f x 3 backward result (new numbering); pending between solves above the leaves
    double* ubuf = nullptr;     // forward update rows x 3, each at its slot (d_udest); pending between solves
    int* d_perm = nullptr;
    int* d_iperm = nullptr;
    int* d_col0 = nullptr;
    int* d_fsize = nullptr;
    int64_t* d_rowptr = nullptr;
    int* d_rows = nullptr;
    int64_t* d_goff = nullptr;
    int64_t* d_uoff = nullptr;
    int64_t* d_asm_ptr = nullptr;
    int64_t* d_asm_idx = nullptr;
    int64_t* d_udest = nullptr;
    int* d_leaf_ptr = nullptr;
    int* d_leaf_sn = nullptr;
    size_t leaf_lds_fwd = 0, leaf_lds_bwd = 0;   // LDS of a leaf-subtree workgroup
    // dataflow schedule of the supernodes above the leaf subtrees: one launch per
    // direction with the leaf walks, ordered so that each workgroup waits only on earlier ones
    void* d_flow_f = nullptr;                    // forward tasks
    void* d_flow_b = nullptr;                    // backward tasks
    int* d_flow_abort = nullptr;                 // sticky: a bounded wait ran out
    int* d_bcopy = nullptr;                      // (row, caller's row) pairs of the dataflow supernodes' columns
    int n_bcopy = 0;
    double* bperm = nullptr;                     // their right-hand-side rows, copied by the leaf launch
    int n_flow_f = 0, n_flow_b = 0;
    // numeric factorization (plan from chol_plan, fronts in one pool)
    wrapd::FactorPlan fplan;
    double* pool = nullptr;
    int64_t* d_foff = nullptr;
    int64_t* d_cbase = nullptr;
    int64_t* d_acol_ptr = nullptr;
    int* d_acol = nullptr;
    int64_t* d_ccol_ptr = nullptr;
    int* d_ccol = nullptr;
    int64_t* d_relptr = nullptr;
    int* d_rel = nullptr;
    int64_t* d_cchunk_ptr = nullptr;
    int* d_cchunk = nullptr;
    int* d_ftasks = nullptr;
    int* d_ffail = nullptr;

    // learned prox warm start (wd_mlp.hip)
    float* mlp_master = nullptr;   // FP32 parameters
    float* mlp_m1 = nullptr;       // Adam moments
    float* mlp_m2 = nullptr;
    float* mlp_gsum = nullptr;     // last summed gradient
    float* mlp_norm = nullptr;
    uint16_t* mlp_bf = nullptr;    // BF16 matrix-core tiles
    float* mlp_loss = nullptr;
    float* mlp_partial = nullptr;  // per-wave partial gradients
    float* mlp_lpart = nullptr;
    int mlp_partial_waves = 0;
    bool mlp_ready = false;
    int mlp_t = 0;                 // Adam step
    float* mlp_log = nullptr;      // training tuples, 9 floats each
    long long mlp_log_cap = 0, mlp_log_count = 0;
    bool mlp_logging = false;

    // reductions
    double* partials = nullptr;
    double* scal = nullptr;      // device scalars
    int* d_red_abort = nullptr;        // sticky: a reduction's wait ran out
    unsigned* d_fin_ticket = nullptr;  // k_finalize's count of finished slots (0 between launches)
    unsigned red_epoch = 0;
    unsigned xd_epoch = 0;             // the epoch of the last x-distance reduction (its host copy: S_XD)
    int alpha_n = -1;                  // S_ALPHA holds the alphas of this list for the gradient in grad (-1: none)
    int alpha_slots[wd::kMaxHist] = {};
    unsigned host_epoch = 0;           // the epoch of the last reduction with a host copy (what the host waits for)
    unsigned local_epoch = 0;          // the epoch of the last local step launched
    bool local_pending = false;        // a local step launched and its results not read yet
    bool local_fb = false;             // ... and it had an FP64 fallback pass (S_LOCAL_FB)
    // FP32 local steps defer their FP64 fallbacks: the element kernel marks them in fb_info
    // (0: none, else 1 + the FP32 attempt's evaluations), a stable compaction lists them in
    // fb_list (fb_count on the device) and a second kernel redoes them in FP64, 64 per wave
    float* fb_info = nullptr;
    int* fb_list = nullptr;
    int* fb_count = nullptr;
    void* fb_tmp = nullptr;
    size_t fb_tmp_bytes = 0;
    bool gs_begun = false;             // the global step's first launches are queued (wrapd_global_step_begin)
    wd::RedTail pending_tail;          // armed by arm_tail, taken by the next reducing launch
    struct PendingFinalize {
        int nslots = 0;
        unsigned maxmask = 0;
        int dst = 0, nblk = 0;   // nblk 0: none
        bool host = true;        // with the host copy
    } pending_fin;                     // a taken tail whose launch is too wide to finish it
    void* scratch = nullptr;     // this device's address of the library's static scratch array (wd_capi.hip)
    // launches wider than this many blocks finish their reductions in a separate one-block
    // launch (kTailMaxBlocks; WRAPD_TAIL_MAX_BLOCKS in the environment at wrapd_create
    // lowers it, so small meshes can exercise that path in tests)
    int tail_max_blocks = 1024;
    // the factorization plan's choices: WRAPD_L21_RIGHT_MAX and WRAPD_ASM_CHUNK in the
    // environment at wrapd_create override the defaults, so tests can exercise every path
    // (all give the same factor)
    wrapd::PlanOptions plan_opt;
    // the learned warm start runs the network for elements whose previous prox solution has
    // a relative optimality residual above this (WRAPD_LEARNED_GATE at wrapd_create).  At
    // 1e-3 the network ran for most 16-element blocks in steady state and cost more than the
    // Newton iterations it saved (C2: 108 vs 75 us per local step for 21% fewer iterations;
    // C5: 0.85 vs 0.58 ms); from 1e-1 on it runs where a start is far off (a first step,
    // a jump after reweighting) and costs ~5% elsewhere
    double learned_gate = 1e-1;
    double* scal_host = nullptr; // pinned host mirror
    unsigned* host_seq = nullptr; // pinned: the epoch of the last reduction whose host copy is stored

    // SplitLBFGS host-side state (deps/mcloptlib/include/MCL/SplitLBFGS.hpp:48-62)
    wrapd_lbfgs_cfg lcfg{6, 100, 1.0, 0};
    int m_k = 0;
    int pos2slot[wd::kMaxHist] = {0, 1, 2, 3, 4, 5, 6, 7};   // reference position -> device slot
    bool grad_old_valid = false;
    int rot_aware = 1;
    double last_energy = 0.0;
    int last_flips = 0;

    ~wrapd_ctx();
};

namespace wd {

// profiling scope around a launch
struct ProfScope {
    wrapd_ctx* c;
    int idx = -1;
    ProfScope(wrapd_ctx* ctx, const char* name);
    ~ProfScope();
};

inline int blocks_for(long long n, int nt) { return static_cast<int>((n + nt - 1) / nt); }

// (wd_setup.hip) vidx, A's pattern and contribution lists, the vertex -> corner lists
int build_mesh_maps(wrapd_ctx* c, const int* d_tets, const int* d_comb, long long* n_corners);
// launchers (wd_elem.hip)
void launch_setup(wrapd_ctx* c, const int* d_tv, const double* d_rest, int* d_bad);
void launch_init_duals(wrapd_ctx* c, int rot_aware);
void launch_defo(wrapd_ctx* c, int mode, int use_hessian);   // partial slots: 0 value, 1 energy, 2 flips
int prepare_local_fallback(wrapd_ctx* c);   // the deferred-fallback list's compaction storage
void launch_local(wrapd_ctx* c, const wrapd_local_args* a);  // slots: 0 bef,1 aft,2 pe,3 maxratio,4 iters,5 fallback,6 neg,7 accepted
void launch_local_unaware(wrapd_ctx* c, const wrapd_local_args* a);
void launch_update_weights(wrapd_ctx* c);
void launch_max_ratio(wrapd_ctx* c);
void launch_assemble_A(wrapd_ctx* c);
void launch_svd_batch(wrapd_ctx* c, int n, int precision, const double* F, double* sig, double* U, double* V);
// (wd_vec.hip)
// The next reducing launch finishes its own reduction into scal[dst, dst + nslots) in its
// last block, summing the blocks' partials in a fixed order.  host: also into the pinned
// host copy, followed by the reduction's epoch (false for results only later launches read).
void arm_tail(wrapd_ctx* c, int nslots, unsigned maxmask, int dst, bool host = true);
// The armed tail for a launch of nblk blocks.  With may_split and more than
// c->tail_max_blocks (kTailMaxBlocks) blocks the reduction is instead left to finish_reduction, to be called
// right after that launch (a one-block finalize launch: one block reading that many
// partials is slower than a launch).
wd::RedTail take_tail(wrapd_ctx* c, int nblk, bool may_split);
void finish_reduction(wrapd_ctx* c);
constexpr int kTailMaxBlocks = 1024;
// mode: 0 plain, 1 first (dg), 2 step (y + Gram).  alpha_slots: the next two-loop's list
// (mode 1, or mode 2 with the list itself); its alphas are then left in S_ALPHA, recorded in
// alpha_n / alpha_slots
void launch_gather_grad(wrapd_ctx* c, double* out, int mode, int new_slot, const int* list_slots, int nlist,
                        const int* alpha_slots = nullptr, int nalpha = 0);
// dots of base with vecs into S_RED; post 1 then runs the alpha recurrence, 2 the beta one
// (SlotList of `slots`), in the launch's last block.  post 3: the beta recurrence and then
// g . r = base . g + sum coef_j g . s_j (extra = g) into S_DIR for the host, without forming r
void launch_multidot(wrapd_ctx* c, const double* base, double* const* vecs, int nvec, int post = 0,
                     const int* slots = nullptr, const double* extra = nullptr);
void launch_axpy_q(wrapd_ctx* c, const double* g, const int* slots, int n);   // qv = g - sum alpha y
void launch_axpy_r(wrapd_ctx* c, const double* g, const int* slots, int n);   // rv += sum coef s, dir = g.r
void launch_dot(wrapd_ctx* c, const double* a, const double* b); // partial slot 0
void launch_xstep(wrapd_ctx* c, double rate, const double* dirv, double* s_out);
// the same with the direction r = rv + sum coef_j s_j formed per element (as launch_axpy_r would)
void launch_xstep_r(wrapd_ctx* c, double rate, const int* slots, int n, double* s_out);
void launch_diffnorm(wrapd_ctx* c, const double* a, const double* b, int dst = S_RED);   // |a - b|^2 -> dst
void launch_copy(wrapd_ctx* c, double* dst, const double* src, long long n);
void launch_fill(wrapd_ctx* c, double* dst, double v, long long n);
void launch_armijo_x(wrapd_ctx* c, double alpha, const double* p, double* xbuf); // xbuf = x + alpha p
// (wd_solve.hip)
int chol_prepare_device(wrapd_ctx* c);
// x = A^{-1} b (n_f x 3 device arrays); with ny > 0 the right-hand side is
// b - sum_j alpha_j y_{yslots[j]} (alpha in S_ALPHA), SplitLBFGS's q, formed inside the solve
void launch_solve(wrapd_ctx* c, const double* b, double* x, const int* yslots = nullptr, int ny = 0);
int solve_aborted(wrapd_ctx* c);   // nonzero once a solve's dependency wait has timed out (syncs)
// (wd_mlp.hip)
MlpDevice mlp_device(wrapd_ctx* c);
int mlp_upload(wrapd_ctx* c, const float* params, const float* norm);
int mlp_infer(wrapd_ctx* c, const float* d_in, int n, float* d_out);
int mlp_step(wrapd_ctx* c, const float* d_tuples, long long n_data, const int* d_idx, long long n, uint64_t seed,
             float lr, int t);   // loss to c->mlp_loss[0], summed gradient to c->mlp_gsum
void launch_mlp_augment(wrapd_ctx* c, float* d_tuples, long long n, double lo, double hi, uint64_t seed);
inline int mlp_grad_waves(long long n) { return static_cast<int>((n + 255) / 256); }   // 4 rounds of 64 per wave
void launch_mlp_grad(wrapd_ctx* c, const float* d_tuples, const int* d_idx, long long n, float* partial, float* lpart,
                     float* gsum, float* loss);
int mlp_sample(wrapd_ctx* c, long long stride, long long m, float* host);   // every stride-th logged tuple
// (wd_factor.hip)
int factor_prepare_device(wrapd_ctx* c);
void launch_factor(wrapd_ctx* c);   // status flag in c->d_ffail

int check_hip(wrapd_ctx* c, hipError_t e, const char* what);

}  // namespace wd
