// Supernodal triangular solves with three right-hand sides (the x, y, z columns of
// SplitLBFGS's H0 application, deps/mcloptlib/include/MCL/SplitLBFGS.hpp:159-171).
//
// Each supernode s stores G_s = [L11^{-1}; -L21 L11^{-1}] (f_s x ns_s, column-major), so
//   forward  (L y = b):   [y_J ; u_s] = [0 ; f_R] + G_s c_s,  c_s = b_J + child updates
//   backward (L^T x = y): x_J = G_s^T [y_J ; x_R]
// Both are one dense product per supernode: rows (forward) or columns (backward) are
// independent, so a wide supernode is a single step of the schedule.  Child updates are
// gathered by the parent from precomputed lists (deterministic, no atomics).
//
// Schedule: small supernodes are grouped into leaf subtrees, each walked by one
// workgroup in postorder; the remaining supernodes run level by level.  A level with
// a few wide supernodes must still fill 256 CUs, so each level picks its task size:
//   forward  tasks (supernode, block of R rows), R in {256, 128, 64, 32}: the
//            256 threads are 256/R column groups, reduced in LDS in a fixed order;
//   backward tasks (supernode, block of 4 CPW columns), CPW columns per wave in {4, 2, 1}.
// No task depends on another task of its launch (no cross-workgroup reduction: on this
// 8-XCD part a device-scope fence costs an L2 writeback).
#include "wd_ctx.hpp"
#include "device/wd_reduce.hpp"

namespace wd {

constexpr int kSolveThreads = 256;
constexpr int kBwdChunk = 2048;       // rows of [y_J; x_R] staged in LDS at a time
constexpr int kLevelTasks = 512;      // task count a level aims for
constexpr int kNsMax = 1024;          // supernode width limit (symbolic ns_max)
constexpr int kSmallFront = 256;      // leaf-subtree front limit (symbolic small_front)

struct SolveMaps {
    const int* col0;
    const int* fsize;
    const int64_t* rowptr;
    const int* rows;
    const int64_t* goff;
    const int64_t* uoff;
    const int64_t* asm_ptr;
    const int64_t* asm_idx;
    const int* perm;
};

// t += sum_k g[k * gstride] * v[k * vstride + 0..2], k < n.  Loads are issued in batches
// of 8 independent reads before they are used, so a thread has 8 misses in flight
// instead of one: the solves' dense products are latency-bound otherwise.
__device__ __forceinline__ void strided_dot3(const double* __restrict__ g, int64_t gstride,
                                             const double* __restrict__ v, int vstride, int n, double& t0,
                                             double& t1, double& t2) {
    constexpr int B = 8;
    int k = 0;
    for (; k + B <= n; k += B) {
        double gv[B];
#pragma unroll
        for (int u = 0; u < B; ++u) gv[u] = g[static_cast<int64_t>(k + u) * gstride];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const double* vp = v + (k + u) * vstride;
            t0 += gv[u] * vp[0];
            t1 += gv[u] * vp[1];
            t2 += gv[u] * vp[2];
        }
    }
    for (; k < n; ++k) {
        const double gk = g[static_cast<int64_t>(k) * gstride];
        const double* vp = v + k * vstride;
        t0 += gk * vp[0];
        t1 += gk * vp[1];
        t2 += gk * vp[2];
    }
}

__device__ __forceinline__ void assemble_row(const SolveMaps& M, int64_t prow, const double* __restrict__ ubuf,
                                             double a[3]) {
    for (int64_t q = M.asm_ptr[prow]; q < M.asm_ptr[prow + 1]; ++q) {
        const double* u = ubuf + 3 * M.asm_idx[q];
        a[0] += u[0];
        a[1] += u[1];
        a[2] += u[2];
    }
}

// Leaf subtrees: one workgroup walks a whole subtree in postorder.  A subtree is a
// contiguous range of supernodes (postorder), and every update row it produces except its
// root's is consumed inside it, so those rows stay in LDS (dynamic, sized by the largest
// subtree); only the root's rows go to the global buffer for the level kernels.
__global__ void __launch_bounds__(kSolveThreads)
k_fwd_leaf(const int* __restrict__ leaf_ptr, const int* __restrict__ leaf_sn, SolveMaps M,
           const double* __restrict__ b, const double* __restrict__ G, double* __restrict__ ysol,
           double* __restrict__ ubuf) {
    extern __shared__ double dyn[];
    double* cv = dyn;                     // kSmallFront x 3
    double* ul = dyn + kSmallFront * 3;   // the subtree's update rows x 3
    const int t = blockIdx.x;
    const int q0 = leaf_ptr[t], q1 = leaf_ptr[t + 1];
    const int root = leaf_sn[q1 - 1];
    const int64_t ubase = M.uoff[leaf_sn[q0]];
    for (int q = q0; q < q1; ++q) {
        const int s = leaf_sn[q];
        const int c0 = M.col0[s];
        const int ns = M.col0[s + 1] - c0;
        const int f = M.fsize[s];
        const int64_t rp = M.rowptr[s];
        for (int r = threadIdx.x; r < f; r += kSolveThreads) {
            double a[3] = {0.0, 0.0, 0.0};
            if (r < ns) {
                const int old = M.perm[c0 + r];
                a[0] = b[3 * old + 0];
                a[1] = b[3 * old + 1];
                a[2] = b[3 * old + 2];
            }
            for (int64_t k = M.asm_ptr[rp + r]; k < M.asm_ptr[rp + r + 1]; ++k) {
                const double* u = ul + 3 * (M.asm_idx[k] - ubase);
                a[0] += u[0];
                a[1] += u[1];
                a[2] += u[2];
            }
            cv[3 * r + 0] = a[0];
            cv[3 * r + 1] = a[1];
            cv[3 * r + 2] = a[2];
        }
        __syncthreads();
        const double* Gs = G + M.goff[s];
        for (int r = threadIdx.x; r < f; r += kSolveThreads) {
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
            const int jend = r < ns ? r + 1 : ns;   // L11^{-1} is lower triangular
            strided_dot3(Gs + r, f, cv, 3, jend, t0, t1, t2);
            if (r < ns) {
                double* y = ysol + 3ll * (c0 + r);
                y[0] = t0; y[1] = t1; y[2] = t2;
            } else {
                const int64_t ui = M.uoff[s] + (r - ns);
                const double u0 = cv[3 * r + 0] + t0, u1 = cv[3 * r + 1] + t1, u2 = cv[3 * r + 2] + t2;
                double* u = ul + 3 * (ui - ubase);
                u[0] = u0; u[1] = u1; u[2] = u2;
                if (s == root) {
                    double* ug = ubuf + 3 * ui;
                    ug[0] = u0; ug[1] = u1; ug[2] = u2;
                }
            }
        }
        __syncthreads();
    }
}

// forward: task (s, row block rb) with R rows; thread (g, r) = (tid / R, tid % R) sums
// columns j = g, g + 256/R, ... of row r
__global__ void __launch_bounds__(kSolveThreads)
k_fwd_level(const int2* __restrict__ tasks, int R, SolveMaps M, const double* __restrict__ b,
            const double* __restrict__ G, double* __restrict__ ysol, double* __restrict__ ubuf) {
    __shared__ double cv[kNsMax * 3];
    __shared__ double red[kSolveThreads * 3];
    const int2 task = tasks[blockIdx.x];
    const int s = task.x, rb = task.y;
    const int c0 = M.col0[s];
    const int ns = M.col0[s + 1] - c0;
    const int f = M.fsize[s];
    const int64_t rp = M.rowptr[s];
    for (int j = threadIdx.x; j < ns; j += kSolveThreads) {
        const int old = M.perm[c0 + j];
        double a[3] = {b[3 * old + 0], b[3 * old + 1], b[3 * old + 2]};
        assemble_row(M, rp + j, ubuf, a);
        cv[3 * j + 0] = a[0];
        cv[3 * j + 1] = a[1];
        cv[3 * j + 2] = a[2];
    }
    __syncthreads();
    const int ng = kSolveThreads / R;
    const int g = threadIdx.x / R, rl = threadIdx.x - g * R;
    const int r = rb * R + rl;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (r < f) {
        const double* Gs = G + M.goff[s] + r;
        const int jend = r < ns ? r + 1 : ns;   // L11^{-1} is lower triangular
        if (g < jend)   // columns g, g + ng, ... below jend
            strided_dot3(Gs + static_cast<int64_t>(g) * f, static_cast<int64_t>(ng) * f, cv + 3 * g, 3 * ng,
                         (jend - g + ng - 1) / ng, t0, t1, t2);
    }
    if (ng > 1) {
        red[3 * threadIdx.x + 0] = t0;
        red[3 * threadIdx.x + 1] = t1;
        red[3 * threadIdx.x + 2] = t2;
        __syncthreads();
        if (g != 0) return;
        for (int k = 1; k < ng; ++k) {
            const double* o = red + 3 * (k * R + rl);
            t0 += o[0];
            t1 += o[1];
            t2 += o[2];
        }
    }
    if (r >= f) return;
    if (r < ns) {
        double* y = ysol + 3ll * (c0 + r);
        y[0] = t0; y[1] = t1; y[2] = t2;
    } else {
        double a[3] = {0.0, 0.0, 0.0};
        assemble_row(M, rp + r, ubuf, a);
        double* u = ubuf + 3 * (M.uoff[s] + (r - ns));
        u[0] = a[0] + t0;
        u[1] = a[1] + t1;
        u[2] = a[2] + t2;
    }
}

// backward: task (s, column block cb) of 4 * CPW columns, CPW per wave: x_J = G^T [y_J; x_R]
template <int CPW>
__global__ void __launch_bounds__(kSolveThreads)
k_bwd_level(const int2* __restrict__ tasks, SolveMaps M, const double* __restrict__ G,
            const double* __restrict__ ysol, double* __restrict__ xsol, double* __restrict__ out) {
    __shared__ double vv[kBwdChunk * 3];
    const int2 task = tasks[blockIdx.x];
    const int s = task.x, cb = task.y;
    const int c0 = M.col0[s];
    const int ns = M.col0[s + 1] - c0;
    const int f = M.fsize[s];
    const int64_t rp = M.rowptr[s];
    const double* Gs = G + M.goff[s];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jfirst = cb * 4 * CPW;
    double acc[CPW][3];
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) acc[jj][0] = acc[jj][1] = acc[jj][2] = 0.0;
    for (int r0 = jfirst - jfirst % 64; r0 < f; r0 += kBwdChunk) {   // rows above jfirst are zero
        const int rc = min(kBwdChunk, f - r0);
        for (int r = threadIdx.x; r < rc; r += kSolveThreads) {
            const int row = r0 + r;
            const double* src = row < ns ? ysol + 3ll * (c0 + row) : xsol + 3ll * M.rows[rp + row];
            vv[3 * r + 0] = src[0];
            vv[3 * r + 1] = src[1];
            vv[3 * r + 2] = src[2];
        }
        __syncthreads();
#pragma unroll
        for (int jj = 0; jj < CPW; ++jj) {
            const int j = jfirst + w * CPW + jj;
            if (j < ns) {
                const double* gc = Gs + static_cast<int64_t>(j) * f + r0;
                const int rstart = j > r0 ? j - r0 : 0;   // rows above the diagonal are zero
                const int r1 = rstart + lane;
                if (r1 < rc)
                    strided_dot3(gc + r1, 64, vv + 3 * r1, 3 * 64, (rc - r1 + 63) / 64, acc[jj][0], acc[jj][1],
                                 acc[jj][2]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
        const int j = jfirst + w * CPW + jj;
        const double a0 = wave_sum(acc[jj][0]);
        const double a1 = wave_sum(acc[jj][1]);
        const double a2 = wave_sum(acc[jj][2]);
        if (lane == 0 && j < ns) {
            double* x = xsol + 3ll * (c0 + j);
            x[0] = a0; x[1] = a1; x[2] = a2;
            double* xo = out + 3ll * M.perm[c0 + j];
            xo[0] = a0; xo[1] = a1; xo[2] = a2;
        }
    }
}

// The subtree's own solution rows stay in LDS too (its supernodes' columns are the
// contiguous range [col0[first], col0[root+1])); rows above the subtree come from xsol.
// Results go to xsol and, in the caller's numbering, to out.
__global__ void __launch_bounds__(kSolveThreads)
k_bwd_leaf(const int* __restrict__ leaf_ptr, const int* __restrict__ leaf_sn, SolveMaps M,
           const double* __restrict__ G, const double* __restrict__ ysol, double* __restrict__ xsol,
           double* __restrict__ out) {
    extern __shared__ double dyn[];
    double* vv = dyn;                     // kSmallFront x 3
    double* xl = dyn + kSmallFront * 3;   // the subtree's columns x 3
    const int t = blockIdx.x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q0 = leaf_ptr[t], q1 = leaf_ptr[t + 1];
    const int cfirst = M.col0[leaf_sn[q0]], cend = M.col0[leaf_sn[q1 - 1] + 1];
    for (int q = q1 - 1; q >= q0; --q) {
        const int s = leaf_sn[q];
        const int c0 = M.col0[s];
        const int ns = M.col0[s + 1] - c0;
        const int f = M.fsize[s];
        const int64_t rp = M.rowptr[s];
        for (int r = threadIdx.x; r < f; r += kSolveThreads) {
            const double* src;
            if (r < ns) {
                src = ysol + 3ll * (c0 + r);
            } else {
                const int row = M.rows[rp + r];
                src = (row >= cfirst && row < cend) ? xl + 3 * (row - cfirst) : xsol + 3ll * row;
            }
            vv[3 * r + 0] = src[0];
            vv[3 * r + 1] = src[1];
            vv[3 * r + 2] = src[2];
        }
        __syncthreads();
        const double* Gs = G + M.goff[s];
        for (int j = w; j < ns; j += kSolveThreads / 64) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            const double* gc = Gs + static_cast<int64_t>(j) * f;
            const int r1 = j + lane;
            if (r1 < f) strided_dot3(gc + r1, 64, vv + 3 * r1, 3 * 64, (f - r1 + 63) / 64, a0, a1, a2);
            a0 = wave_sum(a0);
            a1 = wave_sum(a1);
            a2 = wave_sum(a2);
            if (lane == 0) {
                double* x = xl + 3 * (c0 + j - cfirst);
                x[0] = a0; x[1] = a1; x[2] = a2;
                double* xo = out + 3ll * M.perm[c0 + j];
                xo[0] = a0; xo[1] = a1; xo[2] = a2;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
template <typename T>
static int up(wrapd_ctx* c, T*& dptr, const std::vector<T>& h) {
    if (dptr) { (void)hipFree(dptr); dptr = nullptr; }
    size_t bytes = std::max<size_t>(sizeof(T), h.size() * sizeof(T));
    if (check_hip(c, hipMalloc(&dptr, bytes), "hipMalloc(chol)")) return 1;
    if (!h.empty() && check_hip(c, hipMemcpy(dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy(chol)"))
        return 1;
    return 0;
}

int chol_prepare_device(wrapd_ctx* c) {
    const wrapd::CholSymbolic& S = c->sym;
    if (S.max_ns > kNsMax) { c->last_error = "supernode wider than the solve kernel supports"; return WRAPD_ERR_ARG; }
    int rc = 0;
    rc |= up(c, c->d_perm, S.perm);
    rc |= up(c, c->d_iperm, S.iperm);
    rc |= up(c, c->d_col0, S.col0);
    rc |= up(c, c->d_fsize, S.fsize);
    rc |= up(c, c->d_rowptr, S.rowptr);
    rc |= up(c, c->d_rows, S.rows);
    rc |= up(c, c->d_goff, S.goff);
    rc |= up(c, c->d_uoff, S.uoff);
    rc |= up(c, c->d_asm_ptr, S.asm_ptr);
    rc |= up(c, c->d_asm_idx, S.asm_idx);
    rc |= up(c, c->d_leaf_ptr, S.leaf_ptr);
    // LDS of the leaf kernels: the largest subtree's update rows (forward) and columns
    // (backward); a subtree must be a contiguous supernode range
    int64_t urows = 0, cols = 0;
    for (size_t t = 0; t + 1 < S.leaf_ptr.size(); ++t) {
        const int first = S.leaf_sn[S.leaf_ptr[t]], root = S.leaf_sn[S.leaf_ptr[t + 1] - 1];
        for (int q = S.leaf_ptr[t]; q < S.leaf_ptr[t + 1]; ++q)
            if (S.leaf_sn[q] != first + (q - S.leaf_ptr[t])) {
                c->last_error = "leaf subtree is not a contiguous supernode range";
                return WRAPD_ERR_STATE;
            }
        urows = std::max<int64_t>(urows, S.uoff[root + 1] - S.uoff[first]);
        cols = std::max<int64_t>(cols, S.col0[root + 1] - S.col0[first]);
    }
    c->leaf_lds_fwd = static_cast<size_t>(kSmallFront + urows) * 3 * sizeof(double);
    c->leaf_lds_bwd = static_cast<size_t>(kSmallFront + cols) * 3 * sizeof(double);
    if (c->leaf_lds_fwd > 65536 || c->leaf_lds_bwd > 65536) {
        c->last_error = "leaf subtree too large for the solve kernels' LDS";
        return WRAPD_ERR_ARG;
    }
    rc |= up(c, c->d_leaf_sn, S.leaf_sn);
    // level task lists: forward (s, row block) with R rows per task, backward
    // (s, column block) with 4 CPW columns per task, both chosen per level
    std::vector<int> ft, bt;
    c->fwd_task_ptr.assign(1, 0);
    c->bwd_task_ptr.assign(1, 0);
    c->fwd_rows.clear();
    c->bwd_cpw.clear();
    const int nlev = static_cast<int>(S.level_ptr.size()) - 1;
    for (int l = 0; l < nlev; ++l) {
        int R = 256, cpw = 4;
        for (; R > 32; R /= 2) {
            int64_t n = 0;
            for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) n += (S.fsize[S.level_sn[i]] + R - 1) / R;
            if (n >= kLevelTasks) break;
        }
        for (; cpw > 1; cpw /= 2) {
            int64_t n = 0;
            for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) {
                const int sn = S.level_sn[i];
                n += (S.col0[sn + 1] - S.col0[sn] + 4 * cpw - 1) / (4 * cpw);
            }
            if (n >= kLevelTasks) break;
        }
        for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) {
            const int sn = S.level_sn[i];
            const int ns = S.col0[sn + 1] - S.col0[sn];
            for (int rb = 0; rb * R < S.fsize[sn]; ++rb) { ft.push_back(sn); ft.push_back(rb); }
            for (int cb = 0; cb * 4 * cpw < ns; ++cb) { bt.push_back(sn); bt.push_back(cb); }
        }
        c->fwd_rows.push_back(R);
        c->bwd_cpw.push_back(cpw);
        c->fwd_task_ptr.push_back(static_cast<int>(ft.size() / 2));
        c->bwd_task_ptr.push_back(static_cast<int>(bt.size() / 2));
    }
    rc |= up(c, c->d_fwd_tasks, ft);
    rc |= up(c, c->d_bwd_tasks, bt);
    std::vector<double> z3(3ull * std::max(1, c->n_f), 0.0);
    rc |= up(c, c->ysol, z3);
    rc |= up(c, c->xsol, z3);
    std::vector<double> zu(3ull * std::max<int64_t>(1, S.uoff[S.nsuper]), 0.0);
    rc |= up(c, c->ubuf, zu);
    if (c->G) { (void)hipFree(c->G); c->G = nullptr; }
    rc |= check_hip(c, hipMalloc(&c->G, std::max<size_t>(8, S.goff[S.nsuper] * sizeof(double))), "hipMalloc(G)");
    return rc ? WRAPD_ERR_HIP : WRAPD_OK;
}

void launch_solve(wrapd_ctx* c, const double* b, double* x) {
    const wrapd::CholSymbolic& S = c->sym;
    SolveMaps M{c->d_col0, c->d_fsize, c->d_rowptr, c->d_rows, c->d_goff, c->d_uoff, c->d_asm_ptr, c->d_asm_idx,
                c->d_perm};
    const int nleaf = static_cast<int>(S.leaf_ptr.size()) - 1;
    const int nlev = static_cast<int>(c->fwd_task_ptr.size()) - 1;
    const int2* ft = reinterpret_cast<const int2*>(c->d_fwd_tasks);
    const int2* bt = reinterpret_cast<const int2*>(c->d_bwd_tasks);
    {
        ProfScope ps(c, "solve_fwd");
        if (nleaf > 0)
            hipLaunchKernelGGL(k_fwd_leaf, dim3(nleaf), dim3(kSolveThreads), c->leaf_lds_fwd, c->stream, c->d_leaf_ptr,
                               c->d_leaf_sn, M, b, c->G, c->ysol, c->ubuf);
        for (int l = 0; l < nlev; ++l) {
            const int n = c->fwd_task_ptr[l + 1] - c->fwd_task_ptr[l];
            if (n > 0)
                hipLaunchKernelGGL(k_fwd_level, dim3(n), dim3(kSolveThreads), 0, c->stream, ft + c->fwd_task_ptr[l],
                                   c->fwd_rows[l], M, b, c->G, c->ysol, c->ubuf);
        }
    }
    {
        ProfScope ps(c, "solve_bwd");
        for (int l = nlev - 1; l >= 0; --l) {
            const int n = c->bwd_task_ptr[l + 1] - c->bwd_task_ptr[l];
            if (n == 0) continue;
            const int2* t = bt + c->bwd_task_ptr[l];
            if (c->bwd_cpw[l] == 4)
                hipLaunchKernelGGL(k_bwd_level<4>, dim3(n), dim3(kSolveThreads), 0, c->stream, t, M, c->G, c->ysol, c->xsol, x);
            else if (c->bwd_cpw[l] == 2)
                hipLaunchKernelGGL(k_bwd_level<2>, dim3(n), dim3(kSolveThreads), 0, c->stream, t, M, c->G, c->ysol, c->xsol, x);
            else
                hipLaunchKernelGGL(k_bwd_level<1>, dim3(n), dim3(kSolveThreads), 0, c->stream, t, M, c->G, c->ysol, c->xsol, x);
        }
        if (nleaf > 0)
            hipLaunchKernelGGL(k_bwd_leaf, dim3(nleaf), dim3(kSolveThreads), c->leaf_lds_bwd, c->stream, c->d_leaf_ptr,
                               c->d_leaf_sn, M, c->G, c->ysol, c->xsol, x);
    }
}

}  // namespace wd
