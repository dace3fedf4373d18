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
// workgroup in postorder.  The remaining supernodes are split into tasks whose size is
// chosen per level of the supernodal tree, so that a level with a few wide supernodes
// still fills 256 CUs:
//   forward  tasks (supernode, block of R rows), R in {256, ..., 16}: the 256 threads
//            are 256/R column groups, reduced in LDS in a fixed order;
//   backward tasks (supernode, block of 4 cpw columns), cpw columns per wave in {4, 2, 1}.
// Each direction is one launch (k_fwd_solve, k_bwd_solve) holding the leaf subtrees and
// the tasks: a workgroup waits for the rows it reads, not for the end of its level or of
// the leaf walks.  Sums are taken in a fixed order, so results do not depend on timing.
#include "wd_ctx.hpp"
#include "device/wd_reduce.hpp"

namespace wd {

constexpr int kSolveThreads = 256;
constexpr int kBwdChunk = 2048;       // rows of [y_J; x_R] staged in LDS at a time
constexpr int kLevelTasks = 512;      // task count a level aims for
constexpr int kNsMax = 1024;          // supernode width limit (symbolic ns_max)
constexpr int kSmallFront = 256;      // leaf-subtree front limit (symbolic small_front)
constexpr int kFwdPrefetch = 16;      // G terms per thread loaded before a forward task's wait
constexpr int kBwdPrefetch = 16;      // G terms per thread loaded before a backward task's wait (all its columns)

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
    const int64_t* udest;   // update row -> its slot (its position in asm_idx)
};

// t += sum_k g[k * gstride] * v[k * vstride + 0..2], k < n, in order of k.  Loads are
// issued in batches of 12 independent reads before they are used (the last batch
// predicated), so a thread has up to 12 misses in flight instead of one: the solves'
// dense products are latency-bound otherwise (16 costs the forward kernel occupancy).
__device__ __forceinline__ void strided_dot3(const double* __restrict__ g, int64_t gstride,
                                             const double* __restrict__ v, int vstride, int n, double& t0,
                                             double& t1, double& t2) {
    constexpr int B = 12;
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
    if (k < n) {
        double gv[B];
#pragma unroll
        for (int u = 0; u < B; ++u) gv[u] = k + u < n ? g[static_cast<int64_t>(k + u) * gstride] : 0.0;
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (k + u < n) {
                const double* vp = v + (k + u) * vstride;
                t0 += gv[u] * vp[0];
                t1 += gv[u] * vp[1];
                t2 += gv[u] * vp[2];
            }
    }
}

constexpr int kSpinLimit = 1 << 20;   // polls (one to two seconds) before a wait gives up
// ld_agent / st_agent and the pending bit pattern are device/wd_reduce.hpp's.

// v already holds one load of p[0..2]; loads the pending words again until none is left.
// False when this wait or another one has given up (v may then hold pending words).  After
// 16 polls a wait sleeps ~1 us between polls: hundreds of waiting workgroups poll the same
// few rows near the root, and their loads otherwise slow down the stores they wait for.
__device__ bool poll3(const double* p, double v[3], int* abortf) {
    for (int polls = 1; pending(v[0]) || pending(v[1]) || pending(v[2]); ++polls) {
        if ((polls & 63) == 0) {
            if (__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
            if (polls >= kSpinLimit) {
                __hip_atomic_store(abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
        }
        if (polls < 16)
            __builtin_amdgcn_s_sleep(2);
        else
            __builtin_amdgcn_s_sleep(40);
        if (pending(v[0])) v[0] = ld_agent(p + 0);
        if (pending(v[1])) v[1] = ld_agent(p + 1);
        if (pending(v[2])) v[2] = ld_agent(p + 2);
    }
    return true;
}

// poll3 for one word
__device__ bool poll1(const double* p, double& v, int* abortf) {
    for (int polls = 1; pending(v); ++polls) {
        if ((polls & 63) == 0) {
            if (__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
            if (polls >= kSpinLimit) {
                __hip_atomic_store(abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
        }
        if (polls < 16)
            __builtin_amdgcn_s_sleep(2);
        else
            __builtin_amdgcn_s_sleep(40);
        v = ld_agent(p);
    }
    return true;
}

// Leaf subtrees: one workgroup walks a whole subtree in postorder.  A subtree is a
// contiguous range of supernodes [first, root] (postorder, checked at set-up), and every
// update row it produces except its root's is consumed inside it, so those rows stay in
// LDS (dynamic, sized by the largest subtree); only the root's rows go to global memory,
// to their slots.  Fronts have at most kSmallFront = kSolveThreads rows, one per thread.
// Per supernode the loads that do not depend on the previous step (G, perm, the maps,
// the next supernode's sizes) are issued before those that do.
constexpr int kLeafPrefetch = 16;   // G terms per row loaded ahead in the leaf kernels
constexpr int kLeafMaxSn = 64;      // supernodes per leaf subtree (checked at set-up)
static_assert(kSmallFront == kSolveThreads, "leaf kernels: one front row per thread");

// A leaf subtree's supernode sizes, loaded into LDS once (vector loads: scalar loads
// would share a wait counter with the LDS traffic of every step).
struct LeafMeta {
    int col0[kLeafMaxSn + 1];
    int fsize[kLeafMaxSn];
    long long rowptr[kLeafMaxSn], goff[kLeafMaxSn], uoff[kLeafMaxSn];
};
__device__ __forceinline__ void load_leaf_meta(const SolveMaps& M, int first, int n, LeafMeta& L) {
    for (int i = threadIdx.x; i <= n; i += kSolveThreads) {
        L.col0[i] = M.col0[first + i];
        if (i < n) {
            L.fsize[i] = M.fsize[first + i];
            L.rowptr[i] = M.rowptr[first + i];
            L.goff[i] = M.goff[first + i];
            L.uoff[i] = M.uoff[first + i];
        }
    }
    __syncthreads();
}
// Each workgroup also copies its share of the dataflow supernodes' right-hand-side rows
// into bperm (b in the solve's numbering; bcopy holds (destination, source) row pairs),
// read by the dataflow tasks of the same launch.  The first row's loads go out with the
// subtree's first loads.
__device__ __forceinline__ void copy_row3(const double* __restrict__ src, double* __restrict__ dst) {
    dst[0] = src[0];
    dst[1] = src[1];
    dst[2] = src[2];
}
// The right-hand side as read by the forward solve: b - sum_j alpha_j y_j (n terms, alpha in
// device memory), formed per row as SplitLBFGS's first loop leaves q (j from n - 1 down),
// so the two-loop's q never goes through memory.  n = 0: b itself.
struct SolveRhs {
    const double* b;
    const double* alpha;
    int n;
    const double* y[kMaxHist];
};
__device__ __forceinline__ void rhs_row(const SolveRhs& R, long long row, double out[3]) {
    const double* bp = R.b + 3 * row;
    out[0] = bp[0];
    out[1] = bp[1];
    out[2] = bp[2];
    for (int j = R.n - 1; j >= 0; --j) {
        const double a = R.alpha[j];
        const double* yp = R.y[j] + 3 * row;
        out[0] -= a * yp[0];
        out[1] -= a * yp[1];
        out[2] -= a * yp[2];
    }
}
// LDS of a leaf-subtree workgroup: its LeafMeta, then the rows (kLeafRowsOff doubles in).
constexpr int kLeafRowsOff = static_cast<int>((sizeof(LeafMeta) + 15) / 16 * 2);
__device__ __forceinline__ void fwd_leaf_body(int t, int nleaf, const int* __restrict__ leaf_ptr,
                                              const int* __restrict__ leaf_sn, const SolveMaps& M,
                                              const SolveRhs& rhs, const double* __restrict__ G,
                                              double* __restrict__ ysol, double* __restrict__ slots,
                                              const int* __restrict__ bcopy, int nbcopy, double* __restrict__ bperm,
                                              double* dyn) {
    constexpr int B = kLeafPrefetch;
    constexpr int KA = 4;   // update rows per front row read ahead (more are read in a loop)
    LeafMeta& L = *reinterpret_cast<LeafMeta*>(dyn);
    double* cv = dyn + kLeafRowsOff;      // kSmallFront x 3
    double* ul = cv + kSmallFront * 3;    // the subtree's update rows x 3
    const int per = (nbcopy + nleaf - 1) / nleaf;
    const int cq0 = t * per, cq1 = min(nbcopy, cq0 + per);
    int cdst = -1;
    double cb[3] = {0.0, 0.0, 0.0};
    if (cq0 + static_cast<int>(threadIdx.x) < cq1) {
        const int q = cq0 + threadIdx.x;
        cdst = bcopy[2 * q];
        rhs_row(rhs, bcopy[2 * q + 1], cb);
    }
    const int q0 = leaf_ptr[t], q1 = leaf_ptr[t + 1];
    const int first = leaf_sn[q0], nsn = q1 - q0;
    load_leaf_meta(M, first, nsn, L);
    if (cdst >= 0) {   // read by the dataflow tasks of this launch
        double* dst = bperm + 3ll * cdst;
        st_agent(dst + 0, cb[0]);
        st_agent(dst + 1, cb[1]);
        st_agent(dst + 2, cb[2]);
    }
    for (int q = cq0 + kSolveThreads + threadIdx.x; q < cq1; q += kSolveThreads) {
        double t3[3];
        rhs_row(rhs, bcopy[2 * q + 1], t3);
        double* dst = bperm + 3ll * bcopy[2 * q];
        st_agent(dst + 0, t3[0]);
        st_agent(dst + 1, t3[1]);
        st_agent(dst + 2, t3[2]);
    }
    const int64_t ubase = L.uoff[0];
    const int r = threadIdx.x;
    // this row's permutation entry and assembly range of step i: loaded one step ahead, so a
    // step's second-level loads (the right-hand side, the update indices) go out at its start
    int old_n = 0;
    int64_t k0_n = 0, k1_n = 0;
    auto maps_of = [&](int i) {
        const int c0 = L.col0[i], ns = L.col0[i + 1] - c0, f = L.fsize[i];
        const int64_t rp = L.rowptr[i];
        old_n = r < ns ? M.perm[c0 + r] : 0;
        k0_n = r < f ? M.asm_ptr[rp + r] : 0;
        k1_n = r < f ? M.asm_ptr[rp + r + 1] : 0;
    };
    if (nsn > 0) maps_of(0);
    for (int i = 0; i < nsn; ++i) {
        const int c0 = L.col0[i], ns = L.col0[i + 1] - c0, f = L.fsize[i];
        const int64_t go = L.goff[i], uo = L.uoff[i];
        const int jend = r < ns ? r + 1 : (r < f ? ns : 0);   // L11^{-1} is lower triangular
        const double* gp = G + go + r;
        double gv[B];
        const int lg = wrapd::g_ld(f);
#pragma unroll
        for (int u = 0; u < B; ++u) gv[u] = u < jend ? gp[static_cast<int64_t>(u) * lg] : 0.0;
        const int old = old_n;
        const int64_t k0 = k0_n, k1 = k1_n;
        int64_t ai[KA];
#pragma unroll
        for (int i = 0; i < KA; ++i) ai[i] = k0 + i < k1 ? M.asm_idx[k0 + i] : 0;
        if (i + 1 < nsn) maps_of(i + 1);
        if (r < f) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            if (r < ns) {
                double t3[3];
                rhs_row(rhs, old, t3);
                a0 = t3[0];
                a1 = t3[1];
                a2 = t3[2];
            }
#pragma unroll
            for (int i = 0; i < KA; ++i)
                if (k0 + i < k1) {
                    const double* u = ul + 3 * (ai[i] - ubase);
                    a0 += u[0];
                    a1 += u[1];
                    a2 += u[2];
                }
            for (int64_t k = k0 + KA; k < k1; ++k) {
                const double* u = ul + 3 * (M.asm_idx[k] - ubase);
                a0 += u[0];
                a1 += u[1];
                a2 += u[2];
            }
            cv[3 * r + 0] = a0;
            cv[3 * r + 1] = a1;
            cv[3 * r + 2] = a2;
        }
        __syncthreads();
        if (r < f) {
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
            for (int u = 0; u < B; ++u)
                if (u < jend) {
                    t0 += gv[u] * cv[3 * u + 0];
                    t1 += gv[u] * cv[3 * u + 1];
                    t2 += gv[u] * cv[3 * u + 2];
                }
            if (jend > B) strided_dot3(gp + static_cast<int64_t>(B) * lg, lg, cv + 3 * B, 3, jend - B, t0, t1, t2);
            if (r < ns) {
                double* y = ysol + 3ll * (c0 + r);
                y[0] = t0; y[1] = t1; y[2] = t2;
            } else {
                const int64_t ui = uo + (r - ns);
                const double u0 = cv[3 * r + 0] + t0, u1 = cv[3 * r + 1] + t1, u2 = cv[3 * r + 2] + t2;
                double* u = ul + 3 * (ui - ubase);
                u[0] = u0; u[1] = u1; u[2] = u2;
                if (i == nsn - 1) {   // the subtree's root: read by the dataflow tasks of this launch
                    double* ug = slots + 3 * M.udest[ui];
                    st_agent(ug + 0, u0);
                    st_agent(ug + 1, u1);
                    st_agent(ug + 2, u2);
                }
            }
        }
        __syncthreads();
    }
}

// The same copy when there are no leaf subtrees.
__global__ void k_bperm(const int* __restrict__ bcopy, int nbcopy, SolveRhs rhs, double* __restrict__ bperm) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nbcopy) {
        double t3[3];
        rhs_row(rhs, bcopy[2 * q + 1], t3);
        copy_row3(t3, bperm + 3ll * bcopy[2 * q]);
    }
}

// ---------------------------------------------------------------------------------
// Dataflow schedule above the leaf subtrees: the forward launch holds the leaf walks and
// then the forward tasks, the backward launch the backward tasks and then the leaf walks.
// A workgroup waits for the rows it reads, not for the end of a level: a level boundary
// costs a launch (~2.3 us) plus the dependent loads of its first tasks (~1.2 us each on
// data the previous launch wrote).
//  - Forward tasks are ordered bottom-up after the leaf walks and read update rows of
//    their children; backward tasks are ordered top-down and read solution rows of their
//    ancestors, and the leaf walks after them read those of theirs.  So a workgroup only
//    waits on lower-numbered workgroups, which are dispatched before it.
//  - The rows handed between tasks (update rows in `slots`, backward solution rows in
//    `xsol`) hold kPending until written.  A task polls the words it needs until none is
//    pending: one round trip after the producer's store, where a separate done-flag
//    would add the producer's store completion and a dependent load after the flag.
//    Every pending word is written once per solve, and NaN results stay quiet NaNs, so a
//    word that is not pending is final.
//  - Whoever reads a word last sets it back to kPending for the next solve: a forward
//    task its update rows' slots (it is their only reader) and its pivot rows' xsol; a
//    backward task its pivot rows' slots (read by every forward task of its supernode,
//    all finished in the previous launch).  The set-up fills both arrays with kPending.
//  - No shared counter or ticket: atomics on one address serialize at ~25 ns each,
//    which on thousands of tasks costs more than the levels themselves.
//  - Rows handed from one task to another in the same launch are stored and loaded at
//    agent scope (written through to memory, read past this XCD's L2).
//  - Every wait is bounded.  When one expires it sets a sticky abort flag, and from then
//    on tasks write NaN instead of waiting; the host reports the flag.
// Operands that no task of the launch writes (G, b, the maps) are loaded before the
// wait, so their latency overlaps the children's work.
struct SolveTask {
    int s, first, width, ns;   // forward: rows [first, first + width); backward: columns, width = 4 x cpw
    int f, c0, np, nu;         // forward: np pivot-row slots from qa, nu update-row slots from uq0
    long long goff, rowptr;
    long long uoff, qa;
    long long uq0, pad;
};
static_assert(sizeof(SolveTask) == 80, "SolveTask layout");
static_assert(kBwdChunk % kSolveThreads == 0 && kNsMax % kSolveThreads == 0, "staging layout");

// Forward task (s, rows [first, first + R)): thread (g, r) = (tid / R, tid % R) sums
// columns j = g, g + 256/R, ... of row r; the 256/R partial sums are added in LDS in a
// fixed order.  Update rows travel in `slots`: row i of u_s is written once, to
// slots[udest[i]], inside its parent front row's range [asm_ptr[p], asm_ptr[p + 1]).  A
// task's inputs are therefore two contiguous slot ranges (those of the supernode's pivot
// rows and those of the task's update rows, both in the task record), polled in parallel
// rounds of kFwdStage into LDS and then summed per row in slot order (a round's loads go
// out before the previous round is summed).  The pivot rows'
// right-hand side comes from bperm, b in the solve's numbering (pending until the leaf
// workgroups of the launch have copied it), so every load of the task depends on the task
// record alone.  Slots and b rows are loaded and polled word by word, consecutive threads
// on consecutive words: these loads go past the L2, so a wave loading word c of 64 rows
// 24 bytes apart would fetch every line the rows span once per word.
// B: G terms per thread loaded before the wait.  LDS (dyn): cv, then sred.
constexpr int kFwdStage = 2 * kSolveThreads;
constexpr int kFwdFlowLds = (kNsMax + kFwdStage) * 3 * static_cast<int>(sizeof(double));
template <int B>
__device__ __forceinline__ void fwd_flow_body(int ti, const SolveTask* __restrict__ tasks, int* abortf,
                                              const SolveMaps& M, const double* __restrict__ bperm,
                                              const double* __restrict__ G, double* __restrict__ ysol,
                                              double* __restrict__ xsol, double* __restrict__ slots, double* dyn,
                                              int& sh_ok) {
    constexpr int kRowsPerThread = kNsMax / kSolveThreads;
    constexpr int kWordsPerThread = 3 * kFwdStage / kSolveThreads;   // slot words per window
    constexpr int kBWordsPerThread = 3 * kRowsPerThread;              // b words
    double* cv = dyn;                    // kNsMax x 3
    double* sred = dyn + kNsMax * 3;     // staged slots, then the column-group partial sums
    static_assert(kFwdStage >= kSolveThreads, "the partial sums reuse the staging buffer");
    if (threadIdx.x == 0) sh_ok = __hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 0 : 1;
    const SolveTask T = tasks[ti];
    const int ns = T.ns, f = T.f, c0 = T.c0;
    const int R = T.width, ng = kSolveThreads / R;
    const int g = threadIdx.x / R, rl = threadIdx.x - g * R;
    const int r = T.first + rl;
    const int np = T.np, ntot = T.np + T.nu;
    // word wi of the staged sequence: 3 words per slot, pivot slots first
    auto word_ptr = [&](int wi) {
        return wi < 3 * np ? slots + 3 * T.qa + wi : slots + 3 * T.uq0 + (wi - 3 * np);
    };
    double v[kWordsPerThread];
    auto load_window = [&](int base) {   // the words of slots [base, base + kFwdStage)
#pragma unroll
        for (int k = 0; k < kWordsPerThread; ++k) {
            const int wi = 3 * base + threadIdx.x + k * kSolveThreads;
            v[k] = wi < 3 * ntot ? ld_agent(word_ptr(wi)) : 0.0;
        }
    };
    // Before the wait: the first window of inputs, this thread's first G terms, b of the
    // pivot rows into LDS, the rows' slot ranges and, for an update row, where its result goes.
    load_window(0);
    const int jend = r < ns ? r + 1 : ns;   // L11^{-1} is lower triangular
    const int nterm = (r < f && g < jend) ? (jend - g + ng - 1) / ng : 0;
    const int lg = wrapd::g_ld(f);
    const double* gp = G + T.goff + r + static_cast<int64_t>(g) * lg;
    const int64_t gstride = static_cast<int64_t>(ng) * lg;
    double gv[B];
#pragma unroll
    for (int u = 0; u < B; ++u) gv[u] = u < nterm ? gp[u * gstride] : 0.0;
    int qlo[kRowsPerThread], qhi[kRowsPerThread];   // relative to the pivot slots' start
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
        const int j = threadIdx.x + k * kSolveThreads;
        const bool in = j < ns;
        qlo[k] = in ? static_cast<int>(M.asm_ptr[T.rowptr + j] - T.qa) : 0;
        qhi[k] = in ? static_cast<int>(M.asm_ptr[T.rowptr + j + 1] - T.qa) : 0;
    }
    const double* bp = bperm + 3ll * c0;
    unsigned bpend = 0;   // this thread's b words still pending
#pragma unroll
    for (int k = 0; k < kBWordsPerThread; ++k) {
        const int wi = threadIdx.x + k * kSolveThreads;
        if (wi < 3 * ns) {
            const double x = ld_agent(bp + wi);
            cv[wi] = x;
            if (pending(x)) bpend |= 1u << k;
        }
    }
    const bool urow = g == 0 && r >= ns && r < f;
    int ulo = 0, uhi = 0;
    int64_t udst = 0;
    if (urow) {   // as indices into the staged sequence: pivot slots first, then update-row slots
        ulo = np + static_cast<int>(M.asm_ptr[T.rowptr + r] - T.uq0);
        uhi = np + static_cast<int>(M.asm_ptr[T.rowptr + r + 1] - T.uq0);
        udst = M.udest[T.uoff + (r - ns)];
    }
    __syncthreads();   // sh_ok is set
    const bool live = sh_ok != 0;
#pragma unroll
    for (int k = 0; k < kBWordsPerThread; ++k)
        if ((bpend >> k) & 1u) {
            const int wi = threadIdx.x + k * kSolveThreads;
            double x = cv[wi];
            if (live && !poll1(bp + wi, x, abortf)) sh_ok = 0;
            cv[wi] = x;
        }
    __syncthreads();   // cv holds b
    // The children's updates, polled, staged and summed in slot order.
    double ua0 = 0.0, ua1 = 0.0, ua2 = 0.0;
    for (int base = 0; base < ntot; base += kFwdStage) {
#pragma unroll
        for (int k = 0; k < kWordsPerThread; ++k) {
            const int wl = threadIdx.x + k * kSolveThreads;   // word of the window
            const int wi = 3 * base + wl;
            if (wi < 3 * ntot) {
                double* wp = word_ptr(wi);
                if (live && !poll1(wp, v[k], abortf)) sh_ok = 0;
                if (wi >= 3 * np) *wp = pending_value();   // this task is the update rows' only reader
            }
            sred[wl] = v[k];
        }
        if (base + kFwdStage < ntot) load_window(base + kFwdStage);   // in flight during this window's sums
        __syncthreads();
        const int wend = base + kFwdStage;
#pragma unroll
        for (int k = 0; k < kRowsPerThread; ++k) {
            const int j = threadIdx.x + k * kSolveThreads;
            const int i0 = max(qlo[k], base), i1 = min(qhi[k], wend);
            if (j < ns && i0 < i1) {
                double a0 = cv[3 * j + 0], a1 = cv[3 * j + 1], a2 = cv[3 * j + 2];
                for (int i = i0; i < i1; ++i) {
                    const double* sv = sred + 3 * (i - base);
                    a0 += sv[0];
                    a1 += sv[1];
                    a2 += sv[2];
                }
                cv[3 * j + 0] = a0;
                cv[3 * j + 1] = a1;
                cv[3 * j + 2] = a2;
            }
        }
        for (int i = max(ulo, base); i < min(uhi, wend); ++i) {
            const double* sv = sred + 3 * (i - base);
            ua0 += sv[0];
            ua1 += sv[1];
            ua2 += sv[2];
        }
        __syncthreads();
    }
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
    for (int u = 0; u < B; ++u)
        if (u < nterm) {
            const double* vp = cv + 3 * (g + u * ng);
            t0 += gv[u] * vp[0];
            t1 += gv[u] * vp[1];
            t2 += gv[u] * vp[2];
        }
    if (nterm > B) strided_dot3(gp + B * gstride, gstride, cv + 3 * (g + B * ng), 3 * ng, nterm - B, t0, t1, t2);
    if (ng > 1) {
        sred[3 * threadIdx.x + 0] = t0;
        sred[3 * threadIdx.x + 1] = t1;
        sred[3 * threadIdx.x + 2] = t2;
        __syncthreads();
        if (g == 0)
            for (int k = 1; k < ng; ++k) {
                const double* o = sred + 3 * (k * R + rl);
                t0 += o[0];
                t1 += o[1];
                t2 += o[2];
            }
    }
    const double nan = __builtin_nan("");
    const bool ok = sh_ok != 0;
    if (urow) {   // read by the parent's tasks in this launch
        double* u = slots + 3 * udst;
        st_agent(u + 0, ok ? ua0 + t0 : nan);
        st_agent(u + 1, ok ? ua1 + t1 : nan);
        st_agent(u + 2, ok ? ua2 + t2 : nan);
    }
    if (g == 0 && r < ns) {   // read by the next launch only
        double* y = ysol + 3ll * (c0 + r);
        y[0] = ok ? t0 : nan;
        y[1] = ok ? t1 : nan;
        y[2] = ok ? t2 : nan;
        double* x = xsol + 3ll * (c0 + r);   // pending until the backward launch writes it
        const double pv = pending_value();
        x[0] = pv;
        x[1] = pv;
        x[2] = pv;
    }
}

// The forward solve in one launch: workgroups [0, nleaf) walk the leaf subtrees (and copy
// bperm), the rest run the dataflow tasks, which only wait on lower-numbered workgroups.
template <int B>
__global__ void __launch_bounds__(kSolveThreads)
k_fwd_solve(int nleaf, const int* __restrict__ leaf_ptr, const int* __restrict__ leaf_sn,
            const int* __restrict__ bcopy, int nbcopy, const SolveTask* __restrict__ tasks, int* abortf, SolveMaps M,
            SolveRhs rhs, const double* __restrict__ G, double* __restrict__ ysol,
            double* __restrict__ xsol, double* __restrict__ slots, double* __restrict__ bperm) {
    extern __shared__ double dyn[];
    __shared__ int sh_ok;
    if (static_cast<int>(blockIdx.x) < nleaf)
        fwd_leaf_body(blockIdx.x, nleaf, leaf_ptr, leaf_sn, M, rhs, G, ysol, slots, bcopy, nbcopy, bperm, dyn);
    else
        fwd_flow_body<B>(blockIdx.x - nleaf, tasks, abortf, M, bperm, G, ysol, xsol, slots, dyn, sh_ok);
}

// Backward task (s, columns [first, first + 4 CPW)): x_J = G^T [y_J; x_R], CPW columns
// per wave, rows staged through LDS in chunks.  BT / CPW G terms per column (the first
// chunk's) are loaded before the wait.  The task also sets its pivot rows' forward slots
// and bperm rows back to pending.  LDS (dyn): the staged rows.
constexpr int kBwdFlowLds = kBwdChunk * 3 * static_cast<int>(sizeof(double));
template <int CPW, int BT>
__device__ __forceinline__ void bwd_flow_body(int ti, const SolveTask* __restrict__ tasks, int* abortf,
                                              const SolveMaps& M, const double* __restrict__ G,
                                              const double* __restrict__ ysol, double* __restrict__ xsol,
                                              double* __restrict__ out, double* __restrict__ slots,
                                              double* __restrict__ bperm, double* vv, int& sh_ok) {
    if (threadIdx.x == 0) sh_ok = __hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 0 : 1;
    const SolveTask T = tasks[ti];
    const int ns = T.ns, f = T.f, c0 = T.c0;
    constexpr int B = BT / CPW;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jfirst = T.first;
    const int rbase = jfirst - jfirst % 64;   // rows above jfirst are zero
    const double* Gs = G + T.goff;
    const int lg = wrapd::g_ld(f);
    double gv[CPW][B];
    {
        const int rc = min(kBwdChunk, f - rbase);
#pragma unroll
        for (int jj = 0; jj < CPW; ++jj) {
            const int j = jfirst + w * CPW + jj;
            int n = 0;
            const double* gc = Gs;
            if (j < ns) {
                const int r1 = (j > rbase ? j - rbase : 0) + lane;
                if (r1 < rc) {
                    n = (rc - r1 + 63) / 64;
                    gc = Gs + static_cast<int64_t>(j) * lg + rbase + r1;
                }
            }
#pragma unroll
            for (int u = 0; u < B; ++u) gv[jj][u] = u < n ? gc[64 * u] : 0.0;
        }
    }
    // Where this wave's columns' results go, and the task's forward slots (one range) to reset.
    int pj[CPW];
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
        const int j = jfirst + w * CPW + jj;
        pj[jj] = j < ns ? M.perm[c0 + j] : 0;
    }
    const int64_t sq0 = M.asm_ptr[T.rowptr + jfirst], sq1 = M.asm_ptr[T.rowptr + min(jfirst + T.width, ns)];
    // Before the wait, the first chunk's rows of this supernode (its forward result,
    // written by the previous launch) go to LDS, and the ancestor rows' indices to registers.
    constexpr int kStage = kBwdChunk / kSolveThreads;
    int srow[kStage];
    {
        const int rc = min(kBwdChunk, f - rbase);
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int r = threadIdx.x + k * kSolveThreads, row = rbase + r;
            srow[k] = -1;
            if (r < rc) {
                if (row < ns) {
                    const double* src = ysol + 3ll * (c0 + row);
                    vv[3 * r + 0] = src[0];
                    vv[3 * r + 1] = src[1];
                    vv[3 * r + 2] = src[2];
                } else {
                    srow[k] = M.rows[T.rowptr + row];
                }
            }
        }
    }
    __syncthreads();   // sh_ok is set
    const bool live = sh_ok != 0;
    double acc[CPW][3];
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) acc[jj][0] = acc[jj][1] = acc[jj][2] = 0.0;
    for (int r0 = rbase; r0 < f; r0 += kBwdChunk) {
        const int rc = min(kBwdChunk, f - r0);
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int r = threadIdx.x + k * kSolveThreads, row = r0 + r;
            if (r < rc) {
                int sr = srow[k];
                if (r0 != rbase) {   // later chunks are staged here, both kinds of rows
                    if (row < ns) {
                        const double* src = ysol + 3ll * (c0 + row);
                        vv[3 * r + 0] = src[0];
                        vv[3 * r + 1] = src[1];
                        vv[3 * r + 2] = src[2];
                        sr = -1;
                    } else {
                        sr = M.rows[T.rowptr + row];
                    }
                }
                if (sr >= 0) {   // an ancestor's solution, written in this launch
                    const double* src = xsol + 3ll * sr;
                    double v[3] = {ld_agent(src + 0), ld_agent(src + 1), ld_agent(src + 2)};
                    if (live && !poll3(src, v, abortf)) sh_ok = 0;
                    vv[3 * r + 0] = v[0];
                    vv[3 * r + 1] = v[1];
                    vv[3 * r + 2] = v[2];
                }
            }
        }
        __syncthreads();
        // lane l takes rows r1 + 64 u of each column; the columns' loads go out together
        int r1[CPW], n[CPW], nmax = 0;
#pragma unroll
        for (int jj = 0; jj < CPW; ++jj) {
            const int j = jfirst + w * CPW + jj;
            r1[jj] = (j > r0 ? j - r0 : 0) + lane;   // rows above the diagonal are zero
            n[jj] = (j < ns && r1[jj] < rc) ? (rc - r1[jj] + 63) / 64 : 0;
            nmax = max(nmax, n[jj]);
        }
        int k0 = 0;
        if (r0 == rbase) {
#pragma unroll
            for (int jj = 0; jj < CPW; ++jj)
#pragma unroll
                for (int u = 0; u < B; ++u)
                    if (u < n[jj]) {
                        const double* vp = vv + 3 * (r1[jj] + 64 * u);
                        acc[jj][0] += gv[jj][u] * vp[0];
                        acc[jj][1] += gv[jj][u] * vp[1];
                        acc[jj][2] += gv[jj][u] * vp[2];
                    }
            k0 = B;
        }
        constexpr int BB = 16 / CPW;   // row terms per column per batch
        for (int u0 = k0; u0 < nmax; u0 += BB) {
            double gb[CPW][BB];
#pragma unroll
            for (int jj = 0; jj < CPW; ++jj) {
                const double* gc = Gs + static_cast<int64_t>(jfirst + w * CPW + jj) * lg + r0 + r1[jj];
#pragma unroll
                for (int u = 0; u < BB; ++u) gb[jj][u] = u0 + u < n[jj] ? gc[64 * (u0 + u)] : 0.0;
            }
#pragma unroll
            for (int jj = 0; jj < CPW; ++jj)
#pragma unroll
                for (int u = 0; u < BB; ++u)
                    if (u0 + u < n[jj]) {
                        const double* vp = vv + 3 * (r1[jj] + 64 * (u0 + u));
                        acc[jj][0] += gb[jj][u] * vp[0];
                        acc[jj][1] += gb[jj][u] * vp[1];
                        acc[jj][2] += gb[jj][u] * vp[2];
                    }
        }
        __syncthreads();
    }
    const double nan = __builtin_nan("");
    const bool ok = sh_ok != 0;
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
        const int j = jfirst + w * CPW + jj;
        const double a0 = wave_sum(acc[jj][0]);
        const double a1 = wave_sum(acc[jj][1]);
        const double a2 = wave_sum(acc[jj][2]);
        acc[jj][0] = a0;
        acc[jj][1] = a1;
        acc[jj][2] = a2;
        if (lane == 0 && j < ns) {   // read by the descendants' tasks in this launch
            double* x = xsol + 3ll * (c0 + j);
            st_agent(x + 0, ok ? a0 : nan);
            st_agent(x + 1, ok ? a1 : nan);
            st_agent(x + 2, ok ? a2 : nan);
        }
    }
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {   // the caller's copy, read after this launch only
        const int j = jfirst + w * CPW + jj;
        if (lane == 0 && j < ns) {
            double* xo = out + 3ll * pj[jj];
            xo[0] = ok ? acc[jj][0] : nan;
            xo[1] = ok ? acc[jj][1] : nan;
            xo[2] = ok ? acc[jj][2] : nan;
        }
    }
    {   // the task's columns' forward slots and b rows: every forward task finished in the previous launch
        const double pv = pending_value();
        for (int64_t q = sq0 + threadIdx.x; q < sq1; q += kSolveThreads) {
            slots[3 * q + 0] = pv;
            slots[3 * q + 1] = pv;
            slots[3 * q + 2] = pv;
        }
        const int j = jfirst + static_cast<int>(threadIdx.x);
        if (static_cast<int>(threadIdx.x) < T.width && j < ns) {
            double* bp = bperm + 3ll * (c0 + j);
            bp[0] = pv;
            bp[1] = pv;
            bp[2] = pv;
        }
    }
}

// The subtree's own solution rows stay in LDS too (its supernodes' columns are the
// contiguous range [col0[first], col0[root + 1])); rows above the subtree come from xsol.
// They are copied to out, in the caller's numbering, at the end.  Wave w takes
// columns w, w + 4, ...; lane l rows j + l + 64 m.  The G terms of KC columns of a wave
// are loaded together, the first KC before the step's rows are staged.
__device__ __forceinline__ void bwd_leaf_body(int t, const int* __restrict__ leaf_ptr,
                                              const int* __restrict__ leaf_sn, const SolveMaps& M,
                                              const double* __restrict__ G, const double* __restrict__ ysol,
                                              const double* __restrict__ xsol, double* __restrict__ out,
                                              int* abortf, double* dyn) {
    constexpr int KC = 4;
    constexpr int MR = kSmallFront / 64;
    LeafMeta& L = *reinterpret_cast<LeafMeta*>(dyn);
    double* vv = dyn + kLeafRowsOff;      // kSmallFront x 3
    double* xl = vv + kSmallFront * 3;    // the subtree's columns x 3
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = threadIdx.x;
    const int q0 = leaf_ptr[t], q1 = leaf_ptr[t + 1];
    const int first = leaf_sn[q0], nsn = q1 - q0;
    load_leaf_meta(M, first, nsn, L);
    const int cfirst = L.col0[0], cend = L.col0[nsn];
    int row = -1;   // this thread's staged row above the supernode's own columns
    {
        const int i = nsn - 1;
        if (r < L.fsize[i] && r >= L.col0[i + 1] - L.col0[i]) row = M.rows[L.rowptr[i] + r];
    }
    for (int i = nsn - 1; i >= 0; --i) {
        const int c0 = L.col0[i], ns = L.col0[i + 1] - c0, f = L.fsize[i];
        const int64_t go = L.goff[i];
        const double* Gs = G + go;
        const int lg = wrapd::g_ld(f);
        double gv[KC][MR];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int j = w + 4 * k;
#pragma unroll
            for (int m = 0; m < MR; ++m) {
                const int rr = j + lane + 64 * m;
                gv[k][m] = (j < ns && rr < f) ? Gs[static_cast<int64_t>(j) * lg + rr] : 0.0;
            }
        }
        int prow = -1;   // the next (lower) supernode's row source, loaded ahead
        if (r < f) {
            const bool above = r >= ns && !(row >= cfirst && row < cend);
            if (above) {   // an ancestor's solution, written by a dataflow task of this launch
                const double* src = xsol + 3ll * row;
                double v[3] = {ld_agent(src + 0), ld_agent(src + 1), ld_agent(src + 2)};
                poll3(src, v, abortf);   // a wait that gave up leaves pending words: NaN results
                vv[3 * r + 0] = v[0];
                vv[3 * r + 1] = v[1];
                vv[3 * r + 2] = v[2];
            } else {
                const double* src = r < ns ? ysol + 3ll * (c0 + r) : xl + 3 * (row - cfirst);
                vv[3 * r + 0] = src[0];
                vv[3 * r + 1] = src[1];
                vv[3 * r + 2] = src[2];
            }
        }
        if (i > 0 && r < L.fsize[i - 1] && r >= c0 - L.col0[i - 1]) prow = M.rows[L.rowptr[i - 1] + r];
        __syncthreads();
        for (int jc = w; jc < ns; jc += 4 * KC) {
            if (jc != w) {
#pragma unroll
                for (int k = 0; k < KC; ++k) {
                    const int j = jc + 4 * k;
#pragma unroll
                    for (int m = 0; m < MR; ++m) {
                        const int rr = j + lane + 64 * m;
                        gv[k][m] = (j < ns && rr < f) ? Gs[static_cast<int64_t>(j) * lg + rr] : 0.0;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < KC; ++k) {
                const int j = jc + 4 * k;
                if (j >= ns) break;   // uniform across the wave
                double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
                for (int m = 0; m < MR; ++m) {
                    const int rr = j + lane + 64 * m;
                    if (rr < f) {
                        const double* vp = vv + 3 * rr;
                        a0 += gv[k][m] * vp[0];
                        a1 += gv[k][m] * vp[1];
                        a2 += gv[k][m] * vp[2];
                    }
                }
                a0 = wave_sum(a0);
                a1 = wave_sum(a1);
                a2 = wave_sum(a2);
                if (lane == 0) {
                    double* x = xl + 3 * (c0 + j - cfirst);
                    x[0] = a0; x[1] = a1; x[2] = a2;
                }
            }
        }
        __syncthreads();
        row = prow;
    }
    for (int c = cfirst + threadIdx.x; c < cend; c += kSolveThreads) {
        const double* x = xl + 3 * (c - cfirst);
        double* xo = out + 3ll * M.perm[c];
        xo[0] = x[0]; xo[1] = x[1]; xo[2] = x[2];
    }
}

// The backward solve in one launch: workgroups [0, ntask) run the dataflow tasks (each
// with its level's columns per wave, 4 x cpw columns per task), the rest walk the leaf
// subtrees, each once the ancestor rows it reads are written.
template <int BT>
__global__ void __launch_bounds__(kSolveThreads)
k_bwd_solve(const SolveTask* __restrict__ tasks, int ntask, const int* __restrict__ leaf_ptr,
            const int* __restrict__ leaf_sn, int* abortf, SolveMaps M, const double* __restrict__ G,
            const double* __restrict__ ysol, double* __restrict__ xsol, double* __restrict__ out,
            double* __restrict__ slots, double* __restrict__ bperm) {
    extern __shared__ double dyn[];
    __shared__ int sh_ok;
    const int ti = blockIdx.x;
    if (ti < ntask) {
        const int width = tasks[ti].width;   // uniform across the workgroup
        if (width == 16) bwd_flow_body<4, BT>(ti, tasks, abortf, M, G, ysol, xsol, out, slots, bperm, dyn, sh_ok);
        else if (width == 8) bwd_flow_body<2, BT>(ti, tasks, abortf, M, G, ysol, xsol, out, slots, bperm, dyn, sh_ok);
        else bwd_flow_body<1, BT>(ti, tasks, abortf, M, G, ysol, xsol, out, slots, bperm, dyn, sh_ok);
    } else {
        bwd_leaf_body(ti - ntask, leaf_ptr, leaf_sn, M, G, ysol, xsol, out, abortf, dyn);
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
    // each update row adds into exactly one parent front row: its slot is its position in asm_idx
    {
        const int64_t nu = S.uoff[S.nsuper];
        std::vector<int64_t> udest(std::max<int64_t>(1, nu), -1);
        bool bad = static_cast<int64_t>(S.asm_idx.size()) != nu;
        for (size_t q = 0; q < S.asm_idx.size() && !bad; ++q) {
            const int64_t i = S.asm_idx[q];
            if (i < 0 || i >= nu || udest[i] >= 0) bad = true;
            else udest[i] = static_cast<int64_t>(q);
        }
        if (bad) {
            c->last_error = "supernode update rows: each must add into exactly one parent row";
            return WRAPD_ERR_STATE;
        }
        rc |= up(c, c->d_udest, udest);
    }
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
        if (S.leaf_ptr[t + 1] - S.leaf_ptr[t] > kLeafMaxSn) {
            c->last_error = "leaf subtree has more supernodes than the solve kernels support";
            return WRAPD_ERR_ARG;
        }
        for (int q = S.leaf_ptr[t]; q < S.leaf_ptr[t + 1]; ++q)
            if (S.fsize[S.leaf_sn[q]] > kSmallFront) {
                c->last_error = "leaf subtree front larger than the solve kernels support";
                return WRAPD_ERR_ARG;
            }
        urows = std::max<int64_t>(urows, S.uoff[root + 1] - S.uoff[first]);
        cols = std::max<int64_t>(cols, S.col0[root + 1] - S.col0[first]);
    }
    c->leaf_lds_fwd = (kLeafRowsOff + static_cast<size_t>(kSmallFront + urows) * 3) * sizeof(double);
    c->leaf_lds_bwd = (kLeafRowsOff + static_cast<size_t>(kSmallFront + cols) * 3) * sizeof(double);
    if (c->leaf_lds_fwd > 65536 || c->leaf_lds_bwd > 65536) {
        c->last_error = "leaf subtree too large for the solve kernels' LDS";
        return WRAPD_ERR_ARG;
    }
    rc |= up(c, c->d_leaf_sn, S.leaf_sn);
    // Task widths are the widest that still give a level about kLevelTasks tasks:
    // forward (s, block of R rows), R in {256, ..., 16}; backward (s, block of 4 cpw
    // columns), cpw in {4, 2, 1}.  Wide tasks keep G's reads long and the per-task costs
    // (the right-hand side is gathered by every task of a supernode) few; narrow ones split
    // the few supernodes near the root.  Forward tasks go bottom-up, backward tasks top-down.
    const int nlev = static_cast<int>(S.level_ptr.size()) - 1;
    std::vector<char> in_level(S.nsuper, 0);
    for (int i = S.level_ptr[0]; i < S.level_ptr[nlev]; ++i) in_level[S.level_sn[i]] = 1;
    std::vector<int> lev_cpw(nlev, 1), lev_rows(nlev, 256), nft(S.nsuper, 0), nbt(S.nsuper, 0);
    for (int l = 0; l < nlev; ++l) {
        int R = 256;
        for (; R > 16; R /= 2) {
            int64_t n = 0;
            for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) n += (S.fsize[S.level_sn[i]] + R - 1) / R;
            if (n >= kLevelTasks) break;
        }
        lev_rows[l] = R;
        int cpw = 4;
        for (; cpw > 1; cpw /= 2) {
            int64_t n = 0;
            for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) {
                const int sn = S.level_sn[i];
                n += (S.col0[sn + 1] - S.col0[sn] + 4 * cpw - 1) / (4 * cpw);
            }
            if (n >= kLevelTasks) break;
        }
        lev_cpw[l] = cpw;
    }
    for (int l = 0; l < nlev; ++l)
        for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) {
            const int sn = S.level_sn[i];
            const int ns = S.col0[sn + 1] - S.col0[sn];
            nft[sn] = (S.fsize[sn] + lev_rows[l] - 1) / lev_rows[l];
            nbt[sn] = (ns + 4 * lev_cpw[l] - 1) / (4 * lev_cpw[l]);
        }
    // a task may only wait on lower-numbered tasks: a level supernode's parent is on a later level
    std::vector<int> lev_of(S.nsuper, -1);
    for (int l = 0; l < nlev; ++l)
        for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) lev_of[S.level_sn[i]] = l;
    for (int i = S.level_ptr[0]; i < S.level_ptr[nlev]; ++i) {
        const int sn = S.level_sn[i], p = S.parent[sn];
        if (p >= 0 && (!in_level[p] || lev_of[p] <= lev_of[sn])) {
            c->last_error = "supernode schedule: a level supernode's parent is in a leaf subtree or not above it";
            return WRAPD_ERR_STATE;
        }
    }
    auto task_of = [&](int sn, int first, int width) {
        SolveTask t{};
        t.s = sn;
        t.first = first;
        t.width = width;
        t.ns = S.col0[sn + 1] - S.col0[sn];
        t.f = S.fsize[sn];
        t.c0 = S.col0[sn];
        t.goff = S.goff[sn];
        t.rowptr = S.rowptr[sn];
        t.uoff = S.uoff[sn];
        return t;
    };
    // forward tasks bottom-up (children before parents), backward tasks top-down
    // a forward task's inputs: the slots of its supernode's pivot rows and of its own update rows
    std::vector<SolveTask> ft, bt;
    std::vector<int> bcopy;   // (row, caller's row) of every dataflow supernode column
    for (int l = 0; l < nlev; ++l)
        for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) {
            const int sn = S.level_sn[i];
            for (int rb = 0; rb < nft[sn]; ++rb) {
                SolveTask t = task_of(sn, rb * lev_rows[l], lev_rows[l]);
                const int64_t* ap = S.asm_ptr.data() + t.rowptr;
                t.qa = ap[0];
                const int64_t np = ap[t.ns] - ap[0];
                const int ur0 = std::max(t.first, t.ns), ur1 = std::min(t.first + t.width, t.f);
                t.uq0 = ur0 < ur1 ? ap[ur0] : 0;
                const int64_t nu = ur0 < ur1 ? ap[ur1] - ap[ur0] : 0;
                if (np + nu > INT32_MAX / 4) {
                    c->last_error = "supernode with more child update rows than the solve kernels support";
                    return WRAPD_ERR_ARG;
                }
                t.np = static_cast<int>(np);
                t.nu = static_cast<int>(nu);
                ft.push_back(t);
            }
            for (int j = S.col0[sn]; j < S.col0[sn + 1]; ++j) {
                bcopy.push_back(j);
                bcopy.push_back(S.perm[j]);
            }
        }
    for (int l = nlev - 1; l >= 0; --l) {
        for (int i = S.level_ptr[l]; i < S.level_ptr[l + 1]; ++i) {
            const int sn = S.level_sn[i];
            for (int cb = 0; cb < nbt[sn]; ++cb) bt.push_back(task_of(sn, cb * 4 * lev_cpw[l], 4 * lev_cpw[l]));
        }
    }
    c->n_flow_f = static_cast<int>(ft.size());
    c->n_flow_b = static_cast<int>(bt.size());
    rc |= up(c, reinterpret_cast<SolveTask*&>(c->d_flow_f), ft);
    rc |= up(c, reinterpret_cast<SolveTask*&>(c->d_flow_b), bt);
    rc |= up(c, c->d_flow_abort, std::vector<int>(1, 0));
    c->n_bcopy = static_cast<int>(bcopy.size() / 2);
    rc |= up(c, c->d_bcopy, bcopy);
    rc |= up(c, reinterpret_cast<uint64_t*&>(c->bperm), std::vector<uint64_t>(3ull * std::max(1, c->n_f), kPendingBits));
    // the rows the dataflow tasks hand over start pending (their bit pattern, copied as integers)
    std::vector<uint64_t> p3(3ull * std::max(1, c->n_f), kPendingBits);
    rc |= up(c, c->ysol, std::vector<double>(p3.size(), 0.0));
    rc |= up(c, reinterpret_cast<uint64_t*&>(c->xsol), p3);
    rc |= up(c, reinterpret_cast<uint64_t*&>(c->ubuf),
             std::vector<uint64_t>(3ull * std::max<int64_t>(1, S.uoff[S.nsuper]), kPendingBits));
    if (c->G) { (void)hipFree(c->G); c->G = nullptr; }
    rc |= check_hip(c, hipMalloc(&c->G, std::max<size_t>(8, S.goff[S.nsuper] * sizeof(double))), "hipMalloc(G)");
    return rc ? WRAPD_ERR_HIP : WRAPD_OK;
}

void launch_solve(wrapd_ctx* c, const double* b, double* x, const int* yslots, int ny) {
    const wrapd::CholSymbolic& S = c->sym;
    SolveMaps M{c->d_col0, c->d_fsize, c->d_rowptr, c->d_rows, c->d_goff, c->d_uoff, c->d_asm_ptr, c->d_asm_idx,
                c->d_perm, c->d_udest};
    const int nleaf = static_cast<int>(S.leaf_ptr.size()) - 1;
    const SolveTask* ft = static_cast<const SolveTask*>(c->d_flow_f);
    const SolveTask* bt = static_cast<const SolveTask*>(c->d_flow_b);
    SolveRhs rhs{};
    rhs.b = b;
    rhs.alpha = c->scal + S_ALPHA;
    rhs.n = ny;
    for (int j = 0; j < ny; ++j) rhs.y[j] = c->hY[yslots[j]];
    {
        ProfScope ps(c, "solve_fwd");
        if (nleaf == 0 && c->n_bcopy > 0)
            hipLaunchKernelGGL(k_bperm, dim3((c->n_bcopy + 255) / 256), dim3(256), 0, c->stream, c->d_bcopy, c->n_bcopy,
                               rhs, c->bperm);
        const size_t lds = std::max<size_t>(c->n_flow_f > 0 ? kFwdFlowLds : 0, nleaf > 0 ? c->leaf_lds_fwd : 0);
        if (nleaf + c->n_flow_f > 0)
            hipLaunchKernelGGL(k_fwd_solve<kFwdPrefetch>, dim3(nleaf + c->n_flow_f), dim3(kSolveThreads), lds, c->stream,
                               nleaf, c->d_leaf_ptr, c->d_leaf_sn, c->d_bcopy, c->n_bcopy, ft, c->d_flow_abort, M, rhs,
                               c->G, c->ysol, c->xsol, c->ubuf, c->bperm);
    }
    {
        ProfScope ps(c, "solve_bwd");
        const size_t lds = std::max<size_t>(c->n_flow_b > 0 ? kBwdFlowLds : 0, nleaf > 0 ? c->leaf_lds_bwd : 0);
        if (nleaf + c->n_flow_b > 0)
            hipLaunchKernelGGL(k_bwd_solve<kBwdPrefetch>, dim3(c->n_flow_b + nleaf), dim3(kSolveThreads), lds, c->stream,
                               bt, c->n_flow_b, c->d_leaf_ptr, c->d_leaf_sn, c->d_flow_abort, M, c->G, c->ysol, c->xsol,
                               x, c->ubuf, c->bperm);
    }
}

int solve_aborted(wrapd_ctx* c) {
    int v = 0;
    if (!c->d_flow_abort) return 0;
    if (hipMemcpyAsync(&v, c->d_flow_abort, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        v = 1;
    if (v) c->dead = true;   // the solves' rows are in no defined state after an expired wait
    return v;
}

}  // namespace wd
