// Vertex-space kernels: the atomic-free D^T z gather and the SplitLBFGS vector work.
// Their reductions finish in the launch's last block (reduce_tail, device/wd_reduce.hpp).
//
// The two-loop recursion of SplitLBFGS::minimize (deps/mcloptlib/include/MCL/SplitLBFGS.hpp:152-177)
// is evaluated with a maintained Gram matrix G[i][j] = s_i . y_j:
//   loop 1:  alpha_i = rho_i (s_i.g - sum_{j>i} alpha_j G[i][j]),  q = g - sum alpha_j y_j
//   loop 2:  beta_i  = rho_i (y_i.r + sum_{j<i} (alpha_j - beta_j) G[j][i]),
//            r <- r + sum (alpha_j - beta_j) s_j
// which is the same recursion with every inner product of the intermediate vectors
// expanded; each loop then costs one multi-dot pass and one multi-axpy pass instead of
// 2M dependent passes.  New Gram entries are computed in the gradient gather pass that
// produces y.
#include "wd_ctx.hpp"
#include "device/wd_reduce.hpp"

namespace wd {

constexpr int kVecThreads = 256;
constexpr int kGatherBatch = 8;   // corners per batch of loads in the D^T z gather

struct SlotList {
    int n;
    int slot[kMaxHist];
};

struct VecPtrs {
    const double* v[kMaxHist];
};

// The reduction of a launch with more than kTailMaxBlocks blocks, as a launch of its own
// with one block per slot: the same order as reduce_tail (lane v of 256 takes blocks v,
// v + 256, ...), so the same bits as the one-block version it replaces, which took 103 us
// for the C5 local step's 11 slots over 12,000 blocks.
__global__ void __launch_bounds__(256)
k_finalize(double* __restrict__ partials, int nblk, int nslots, unsigned maxmask, double* __restrict__ out,
           double* __restrict__ hout, unsigned* __restrict__ hseq, unsigned epoch, unsigned* __restrict__ ticket) {
    // block k reduces slot k; lane v takes blocks v, v + 256, ... in that order (the
    // in-launch tail's order), 16 loads in flight per lane
    constexpr int BB = 16;
    __shared__ double sh[4];
    const int k = blockIdx.x;
    const bool ismax = (maxmask >> k) & 1u;
    double* pk = partials + static_cast<long long>(k) * kTailBlocks;
    double acc = ismax ? -1e308 : 0.0;
    for (int b0 = threadIdx.x; b0 < nblk; b0 += 256 * BB) {
        double v[BB];
#pragma unroll
        for (int u = 0; u < BB; ++u) v[u] = b0 + 256 * u < nblk ? pk[b0 + 256 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < BB; ++u)
            if (b0 + 256 * u < nblk) {
                acc = ismax ? fmax(acc, v[u]) : acc + v[u];
                pk[b0 + 256 * u] = pending_value();   // pending again for the next in-launch reduction
            }
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const double t = ismax ? wave_max(acc) : wave_sum(acc);
    if (l == 0) sh[w] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r;
        if (ismax) {
            r = sh[0];
            for (int i = 1; i < 4; ++i) r = fmax(r, sh[i]);
        } else {
            r = 0.0;
            for (int i = 0; i < 4; ++i) r += sh[i];
        }
        out[k] = r;
        if (hout) __hip_atomic_store(hout + k, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        // the block finishing last publishes the epoch, after every block's host copy
        const unsigned prev = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_SYSTEM);
        if (prev == static_cast<unsigned>(nslots) - 1u) {
            __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (hseq) __hip_atomic_store(hseq, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Store the reduced Gram entries of the new pair (red: the reduction's results in LDS,
// s_new.y_j and s_j.y_new per list entry j, then y.y) and the y.y denominator, which the
// host reads, so it also goes to the host copy hscal.  Thread j < n stores entry j.
__device__ void gram_store(double* __restrict__ scal, double* __restrict__ hscal, const double* red,
                           const SlotList& L, int ns_pos) {
    const int sn = L.slot[ns_pos];
    const int j = threadIdx.x;
    if (j < L.n) {
        const int sj = L.slot[j];
        scal[S_GRAM + sn * kMaxHist + sj] = red[2 * j];       // s_new . y_j
        scal[S_GRAM + sj * kMaxHist + sn] = red[2 * j + 1];   // s_j . y_new
    }
    if (j == 0) {
        scal[S_DENOM] = red[2 * L.n];
        __hip_atomic_store(hscal + S_DENOM, red[2 * L.n], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// The two-loop recurrences on one block's LDS: red holds the multi-dot results (s_i.g for
// loop 1, y_i.r for loop 2), gm the list's Gram entries gm[i * n + j] = G[slot i][slot j],
// ab the alphas (loaded for loop 2).  Thread 0 runs the recurrence; the results go to
// scal (S_ALPHA, or S_COEF = alpha - beta), read by the next launch.  Logical order 0 = oldest.
__device__ void lbfgs_alpha(double* __restrict__ scal, const double* red, const double* gm, double* ab, int n) {
    for (int i = n - 1; i >= 0; --i) {
        double a = red[i];
        for (int j = i + 1; j < n; ++j) a -= ab[j] * gm[i * n + j];
        const double rho = 1.0 / gm[i * n + i];
        ab[i] = rho * a;
        scal[S_ALPHA + i] = ab[i];
    }
}
__device__ void lbfgs_beta(double* __restrict__ scal, const double* red, const double* gm, const double* alpha,
                           double* coef, int n) {
    for (int i = 0; i < n; ++i) {
        double b = red[i];
        for (int j = 0; j < i; ++j) b += coef[j] * gm[j * n + i];
        const double rho = 1.0 / gm[i * n + i];
        const double beta = rho * b;
        coef[i] = alpha[i] - beta;
        scal[S_COEF + i] = coef[i];
    }
}

// grad[v] = sum of the staged element-corner gradients of vertex v, in element order.
// mode 1: also |grad - grad_old|^2 (SplitLBFGS.hpp:138).
// mode 2: y_new = grad - grad_old into slot `ns`, and the Gram entries of the new pair.
// With a list LA (LA.n > 0; in mode 2 the list L itself), also the dots grad . s_j of the
// next two-loop and its alpha recurrence (what a multi-dot launch with post 1 computes
// next), reduced in the slots after the mode's own (mode 1: from 1, mode 2: from 2n + 1)
// and kept out of scal.  KG: the most slots of the launch.
template <int MODE, int KG>
__global__ void __launch_bounds__(kVecThreads)
k_gather(int n_f, const int* __restrict__ ptr, const double* __restrict__ gstage,
         double* __restrict__ out, const double* __restrict__ gold, double* __restrict__ ynew,
         const double* __restrict__ snew, VecPtrs S, VecPtrs Y, SlotList L, int ns_pos, SlotList LA, VecPtrs SA,
         double* __restrict__ partials, RedTail tail, double* __restrict__ scal) {
    const int kd = MODE == 1 ? 1 : 2 * L.n + 1;   // the first alpha-dot slot
    __shared__ double shk[KG][kVecThreads / 64];
    __shared__ double sres[KG], gm[kMaxHist * kMaxHist], sa[kMaxHist];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int v = lb * blockDim.x + threadIdx.x;
    // the block that finishes the reduction loads the list's Gram entries (written by earlier
    // launches; mode 2's new row and column come from this launch's sums) now, so their round
    // trip overlaps its own gather
    if (MODE != 0 && LA.n > 0 && tail.armed && blockIdx.x == gridDim.x - 1) {
        const int n = LA.n;
        for (int q = threadIdx.x; q < n * n; q += kVecThreads) {
            const int i = q / n, j = q - (q / n) * n;
            if (MODE != 2 || (i != ns_pos && j != ns_pos)) gm[q] = scal[S_GRAM + LA.slot[i] * kMaxHist + LA.slot[j]];
        }
    }
    double g[3] = {0.0, 0.0, 0.0};
    if (v < n_f) {
        // the vertex's corners are the staging rows [ptr[v], ptr[v+1]) (a vertex has ~20),
        // loaded kGatherBatch at a time and summed in corner order
        const int q1 = ptr[v + 1];
        for (int qb = ptr[v]; qb < q1; qb += kGatherBatch) {
            double a[kGatherBatch][3];
#pragma unroll
            for (int u = 0; u < kGatherBatch; ++u) {
                a[u][0] = a[u][1] = a[u][2] = 0.0;
                if (qb + u < q1) {
                    const double* p = gstage + 3ll * (qb + u);
                    a[u][0] = p[0];
                    a[u][1] = p[1];
                    a[u][2] = p[2];
                }
            }
#pragma unroll
            for (int u = 0; u < kGatherBatch; ++u)
                if (qb + u < q1) {
                    g[0] += a[u][0];
                    g[1] += a[u][1];
                    g[2] += a[u][2];
                }
        }
        out[3 * v + 0] = g[0];
        out[3 * v + 1] = g[1];
        out[3 * v + 2] = g[2];
    }
    if (MODE == 1) {
        double gs[kMaxHist];   // grad . s_j of the next two-loop's list
#pragma unroll
        for (int j = 0; j < kMaxHist; ++j) {
            gs[j] = 0.0;
            if (j < LA.n && v < n_f) {
                const double* sj = SA.v[j] + 3 * v;
                gs[j] = g[0] * sj[0] + g[1] * sj[1] + g[2] * sj[2];
            }
        }
        double a = 0.0;
        if (v < n_f) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double d = g[c] - gold[3 * v + c];
                a += d * d;
            }
        }
        double part[KG];
        part[0] = a;
#pragma unroll
        for (int j = 0; j + 1 < KG; ++j) part[1 + j] = j < kMaxHist ? gs[j] : 0.0;
        const double t = block_sums<kVecThreads, KG>(part, 1 + LA.n, shk);
        if (static_cast<int>(threadIdx.x) < tail.nslots) put_partial(partials, lb, threadIdx.x, t);
    }
    if (MODE == 2) {
        double y[3] = {0, 0, 0}, s[3] = {0, 0, 0};
        if (v < n_f) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                y[c] = g[c] - gold[3 * v + c];
                ynew[3 * v + c] = y[c];
                s[c] = snew[3 * v + c];
            }
        }
        // slot 2j: s_new . y_j, slot 2j + 1: s_j . y_new, slot 2n: y_new . y_new;
        // slots 2n + 1 + j: grad . s_j (with LA, which is L)
        double ab[2 * kMaxHist], gs[kMaxHist];
#pragma unroll
        for (int j = 0; j < kMaxHist; ++j) {
            double a = 0.0, b = 0.0, e = 0.0;
            if (j < L.n && v < n_f) {
                const double* yj = (j == ns_pos) ? nullptr : Y.v[j] + 3 * v;
                const double* sj = (j == ns_pos) ? nullptr : S.v[j] + 3 * v;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double yy = yj ? yj[c] : y[c];
                    const double ss = sj ? sj[c] : s[c];
                    a += s[c] * yy;
                    b += ss * y[c];
                    e += g[c] * ss;
                }
            }
            ab[2 * j] = a;
            ab[2 * j + 1] = b;
            gs[j] = LA.n > 0 ? e : 0.0;
        }
        const double d = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
        double part[KG];
#pragma unroll
        for (int k = 0; k < KG; ++k) {
            double pk = k < 2 * kMaxHist ? ab[k < 2 * kMaxHist ? k : 0] : 0.0;
            if (k >= 2 * L.n) pk = 0.0;
            if (k == 2 * L.n) pk = d;
#pragma unroll
            for (int j = 0; j < kMaxHist; ++j)
                if (k == kd + j) pk = gs[j];
            part[k] = pk;
        }
        const int nk = kd + LA.n;
        const double t = block_sums<kVecThreads, KG>(part, nk, shk);
        if (static_cast<int>(threadIdx.x) < tail.nslots) put_partial(partials, lb, threadIdx.x, t);
    }
    if (MODE != 0 && reduce_tail<kVecThreads, KG, false>(tail, partials, sres)) {
        static_assert(KG <= kTailSlots, "gather: slot count");
        if (MODE == 2) gram_store(scal, tail.hout - S_RED, sres, L, ns_pos);
        if (LA.n > 0) {   // the next two-loop's alphas; mode 2's new Gram row and column come from sres
            const int n = LA.n;
            if (MODE == 2)   // the rest of gm was loaded at the start
                for (int q = threadIdx.x; q < n * n; q += kVecThreads) {
                    const int i = q / n, j = q - (q / n) * n;
                    if (i == ns_pos) gm[q] = sres[2 * j];
                    else if (j == ns_pos) gm[q] = sres[2 * i + 1];
                }
            if (static_cast<int>(threadIdx.x) < n) scal[S_GS + threadIdx.x] = sres[kd + threadIdx.x];
            __syncthreads();
            if (threadIdx.x == 0) lbfgs_alpha(scal, sres + kd, gm, sa, n);
        }
        __syncthreads();   // every thread's stores are issued before the signal
        tail_signal(tail);
    }
}

// partial dots of `base` with up to 8 vectors, and with `extra` (slot nvec) when given
__global__ void __launch_bounds__(kVecThreads)
k_multidot(long long n3, const double* __restrict__ base, VecPtrs V, int nvec, const double* __restrict__ extra,
           double* __restrict__ partials, RedTail tail, double* __restrict__ scal, SlotList L, int post) {
    constexpr int KM = kMaxHist + 1;
    __shared__ double shk[KM][kVecThreads / 64];
    __shared__ double sres[KM], gm[kMaxHist * kMaxHist], sa[kMaxHist], sc[kMaxHist], sg[kMaxHist];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long long i0 = static_cast<long long>(lb) * blockDim.x * 4 + threadIdx.x;
    // the block that finishes the reduction loads the recurrence's inputs (written by earlier
    // launches) now, so their round trip overlaps its own dots
    if (post && tail.armed && blockIdx.x == gridDim.x - 1) {
        const int n = L.n;
        for (int q = threadIdx.x; q < n * n; q += kVecThreads)
            gm[q] = scal[S_GRAM + L.slot[q / n] * kMaxHist + L.slot[q % n]];
        if (post >= 2 && static_cast<int>(threadIdx.x) < n) sa[threadIdx.x] = scal[S_ALPHA + threadIdx.x];
        if (post == 3 && static_cast<int>(threadIdx.x) < n) sg[threadIdx.x] = scal[S_GS + threadIdx.x];
    }
    double acc[KM];
#pragma unroll
    for (int j = 0; j < KM; ++j) acc[j] = 0.0;
    double ae = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long long i = i0 + static_cast<long long>(u) * blockDim.x;
        if (i < n3) {
            const double b = base[i];
#pragma unroll
            for (int j = 0; j < kMaxHist; ++j)
                if (j < nvec) acc[j] += b * V.v[j][i];
            if (extra) ae += b * extra[i];
        }
    }
#pragma unroll
    for (int j = 0; j < KM; ++j)
        if (j == nvec) acc[j] = ae;
    const int nk = nvec + (extra ? 1 : 0);
    const double t = block_sums<kVecThreads, KM>(acc, nk, shk);
    if (static_cast<int>(threadIdx.x) < tail.nslots) put_partial(partials, lb, threadIdx.x, t);
    // the last block also runs the recurrence that consumes the dots (post 1: alpha, 2: beta),
    // with the list's Gram entries (and for beta the alphas) loaded into LDS together
    if (reduce_tail<kVecThreads, KM, false>(tail, partials, sres)) {
        if (post) {
            const int n = L.n;
            if (post == 1 && static_cast<int>(threadIdx.x) < n) scal[S_GS + threadIdx.x] = sres[threadIdx.x];
            __syncthreads();
            if (threadIdx.x == 0) {
                if (post == 1) lbfgs_alpha(scal, sres, gm, sa, n);
                else lbfgs_beta(scal, sres, gm, sa, sc, n);
            }
            if (post == 3) {
                __syncthreads();
                if (threadIdx.x == 0) {   // g . r without r: base . g + sum_j coef_j (g . s_j), in order
                    double dir = sres[n];
                    for (int j = 0; j < n; ++j) dir += sc[j] * sg[j];
                    scal[S_DIR] = dir;
                    __hip_atomic_store(tail.hout - S_RED + S_DIR, dir, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                }
            }
        }
        tail_signal(tail);
    }
}


// q = g - sum_j alpha_j y_j
__global__ void __launch_bounds__(kVecThreads)
k_axpy_q(long long n3, const double* __restrict__ g, VecPtrs Y, int n, const double* __restrict__ scal,
         double* __restrict__ q) {
    const long long i = static_cast<long long>(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    double v = g[i];
    for (int j = n - 1; j >= 0; --j) v -= scal[S_ALPHA + j] * Y.v[j][i];
    q[i] = v;
}

// r += sum_j coef_j s_j;  partial dir = g . r
__global__ void __launch_bounds__(kVecThreads)
k_axpy_r(long long n3, double* __restrict__ r, VecPtrs S, int n, const double* __restrict__ scal,
         const double* __restrict__ g, double* __restrict__ partials, RedTail tail) {
    __shared__ double sh[kVecThreads / 64];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long long i = static_cast<long long>(lb) * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (i < n3) {
        double v = r[i];
        for (int j = 0; j < n; ++j) v += S.v[j][i] * scal[S_COEF + j];
        r[i] = v;
        d = g[i] * v;
    }
    const double t = block_sum<kVecThreads>(d, sh);
    if (threadIdx.x == 0 && tail.nslots > 0) put_partial(partials, lb, 0, t);
    reduce_tail<kVecThreads, 1>(tail, partials);
}

__global__ void __launch_bounds__(kVecThreads)
k_dot(long long n3, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ partials,
      RedTail tail) {
    __shared__ double sh[kVecThreads / 64];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long long i = static_cast<long long>(lb) * blockDim.x + threadIdx.x;
    double d = 0.0, m = 0.0;
    if (i < n3) { d = a[i] * b[i]; m = fabs(a[i]); }
    const double t = block_sum<kVecThreads>(d, sh);
    const double tm = block_max<kVecThreads>(m, sh);
    if (threadIdx.x == 0 && tail.nslots >= 2) {
        put_partial(partials, lb, 0, t);
        put_partial(partials, lb, 1, tm);
    }
    reduce_tail<kVecThreads, 2>(tail, partials);
}

// x <- x - rate * dir;  s = x_new - x_old
__global__ void __launch_bounds__(kVecThreads)
k_xstep(long long n3, double* __restrict__ x, double rate, const double* __restrict__ dir, double* __restrict__ s) {
    const long long i = static_cast<long long>(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const double xo = x[i];
    const double xn = xo - rate * dir[i];
    x[i] = xn;
    if (s) s[i] = xn - xo;
}

// x -= rate r, s = x_new - x_old, with r = rv + sum_j coef_j s_j (k_axpy_r's expression)
__global__ void __launch_bounds__(kVecThreads)
k_xstep_r(long long n3, double* __restrict__ x, double rate, const double* __restrict__ rv, VecPtrs S, int n,
          const double* __restrict__ scal, double* __restrict__ s) {
    const long long i = static_cast<long long>(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    double v = rv[i];
    for (int j = 0; j < n; ++j) v += S.v[j][i] * scal[S_COEF + j];   // s may be one of S: read before it is written
    const double xo = x[i];
    const double xn = xo - rate * v;
    x[i] = xn;
    s[i] = xn - xo;
}

__global__ void __launch_bounds__(kVecThreads)
k_axpby(long long n3, double* __restrict__ out, const double* __restrict__ a, double alpha, const double* __restrict__ p) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    out[i] = a[i] + alpha * p[i];
}

__global__ void __launch_bounds__(kVecThreads)
k_copy(long long n, double* __restrict__ dst, const double* __restrict__ src) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void __launch_bounds__(kVecThreads)
k_diffnorm(long long n3, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ partials,
           RedTail tail) {
    __shared__ double sh[kVecThreads / 64];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long long i = static_cast<long long>(lb) * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (i < n3) { const double t = a[i] - b[i]; d = t * t; }
    const double t = block_sum<kVecThreads>(d, sh);
    if (threadIdx.x == 0 && tail.nslots > 0) put_partial(partials, lb, 0, t);
    reduce_tail<kVecThreads, 1>(tail, partials);
}

__global__ void __launch_bounds__(kVecThreads)
k_fill(long long n, double* __restrict__ dst, double v) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

// ---------------------------------------------------------------------------------
static SlotList mk_list(const int* slots, int n) {
    SlotList L;
    L.n = n;
    for (int i = 0; i < kMaxHist; ++i) L.slot[i] = i < n ? slots[i] : 0;
    return L;
}

void arm_tail(wrapd_ctx* c, int nslots, unsigned maxmask, int dst, bool host) {
    RedTail t;
    t.armed = true;
    t.abortf = c->d_red_abort;
    t.out = c->scal + dst;
    t.hout = host ? c->scal_host + dst : nullptr;
    t.hseq = host ? c->host_seq : nullptr;
    t.nslots = nslots;
    t.nstore = nslots;
    t.maxmask = maxmask;
    c->pending_tail = t;
}

RedTail take_tail(wrapd_ctx* c, int nblk, bool may_split) {
    RedTail t = c->pending_tail;
    c->pending_tail = RedTail{};
    if (!t.armed) return t;
    if (++c->red_epoch == 0) c->red_epoch = 1;   // the host's sequence word starts at 0
    t.epoch = c->red_epoch;
    if (t.hseq) c->host_epoch = t.epoch;
    if (may_split && nblk > c->tail_max_blocks) {   // one block would read too many partials: finalize apart
        c->pending_fin = wrapd_ctx::PendingFinalize{t.nslots, t.maxmask, static_cast<int>(t.out - c->scal), nblk,
                                                    t.hseq != nullptr};
        RedTail split;   // the blocks still store their partials, for the finalize launch
        split.nslots = t.nslots;
        return split;
    }
    return t;
}

void finish_reduction(wrapd_ctx* c) {
    const wrapd_ctx::PendingFinalize f = c->pending_fin;
    if (f.nblk == 0) return;
    c->pending_fin = wrapd_ctx::PendingFinalize{};
    hipLaunchKernelGGL(k_finalize, dim3(f.nslots), dim3(256), 0, c->stream, c->partials, f.nblk, f.nslots, f.maxmask,
                       c->scal + f.dst, f.host ? c->scal_host + f.dst : nullptr, f.host ? c->host_seq : nullptr,
                       c->red_epoch, c->d_fin_ticket);
}

void launch_gather_grad(wrapd_ctx* c, double* out, int mode, int new_pos, const int* list_slots, int nlist,
                        const int* alpha_slots, int nalpha) {
    ProfScope ps(c, "gather_grad");
    const int nb = blocks_for(c->n_f, kVecThreads);
    VecPtrs S{}, Y{}, SA{};
    SlotList L = mk_list(list_slots, nlist);
    for (int j = 0; j < nlist; ++j) { S.v[j] = c->hS[list_slots[j]]; Y.v[j] = c->hY[list_slots[j]]; }
    double* ynew = nullptr;
    const double* snew = nullptr;
    if (mode == 2) { ynew = c->hY[list_slots[new_pos]]; snew = c->hS[list_slots[new_pos]]; }
    // the next two-loop's alphas (mode 1 or 2; in mode 2 only for the list itself)
    if (mode == 0 || nalpha <= 0 || (mode == 2 && (nalpha != nlist || !std::equal(alpha_slots, alpha_slots + nalpha,
                                                                                    list_slots))))
        nalpha = 0;
    SlotList LA = mk_list(alpha_slots, nalpha);
    for (int j = 0; j < nalpha; ++j) SA.v[j] = c->hS[alpha_slots[j]];
    c->alpha_n = nalpha > 0 ? nalpha : -1;
    for (int j = 0; j < nalpha; ++j) c->alpha_slots[j] = alpha_slots[j];
    // mode 1 reduces into whatever the caller armed; mode 2 into S_RED, then stores the Gram
    // entries; the alpha dots are reduced after the stored slots
    if (mode == 2) arm_tail(c, 2 * nlist + 1, 0u, S_RED);
    if (mode != 0 && nalpha > 0) c->pending_tail.nslots = (mode == 1 ? 1 : 2 * nlist + 1) + nalpha;
    const RedTail t = mode == 0 ? RedTail{} : take_tail(c, nb, false);
    constexpr int kSmall = 16;   // mode 2 slots up to a memory of 5: 3 x 5 + 1
    const int need = mode == 2 ? 2 * nlist + 1 + nalpha : 0;
    auto k = mode == 0 ? k_gather<0, 1> : mode == 1 ? k_gather<1, 1 + kMaxHist>
                                                    : (need <= kSmall ? k_gather<2, kSmall> : k_gather<2, 3 * kMaxHist + 1>);
    hipLaunchKernelGGL(k, dim3(nb), dim3(kVecThreads), 0, c->stream, c->n_f, c->v2e_ptr, c->gstage, out,
                       c->grad_old, ynew, snew, S, Y, L, new_pos, LA, SA, c->partials, t, c->scal);
}

void launch_multidot(wrapd_ctx* c, const double* base, double* const* vecs, int nvec, int post, const int* slots,
                     const double* extra) {
    ProfScope ps(c, "multidot");
    VecPtrs V{};
    for (int j = 0; j < nvec; ++j) V.v[j] = vecs[j];
    const long long n3 = 3ll * c->n_f;
    const int nb = blocks_for(n3, kVecThreads * 4);
    if (post != 3) extra = nullptr;
    // read by the next launches only, except post 3's S_DIR (stored by the launch itself)
    arm_tail(c, nvec + (extra ? 1 : 0), 0u, S_RED, post == 3);
    if (post == 3) c->pending_tail.nstore = 0;
    const SlotList L = post ? mk_list(slots, nvec) : SlotList{};
    hipLaunchKernelGGL(k_multidot, dim3(nb), dim3(kVecThreads), 0, c->stream, n3, base, V, nvec, extra, c->partials,
                       take_tail(c, nb, false), c->scal, L, post);
}

void launch_axpy_q(wrapd_ctx* c, const double* g, const int* slots, int n) {
    ProfScope ps(c, "axpy_q");
    VecPtrs Y{};
    for (int j = 0; j < n; ++j) Y.v[j] = c->hY[slots[j]];
    const long long n3 = 3ll * c->n_f;
    hipLaunchKernelGGL(k_axpy_q, dim3(blocks_for(n3, kVecThreads)), dim3(kVecThreads), 0, c->stream, n3, g, Y,
                       n, c->scal, c->qv);
}

void launch_axpy_r(wrapd_ctx* c, const double* g, const int* slots, int n) {
    ProfScope ps(c, "axpy_r");
    VecPtrs S{};
    for (int j = 0; j < n; ++j) S.v[j] = c->hS[slots[j]];
    const long long n3 = 3ll * c->n_f;
    const int nb = blocks_for(n3, kVecThreads);
    arm_tail(c, 1, 0u, S_DIR);
    hipLaunchKernelGGL(k_axpy_r, dim3(nb), dim3(kVecThreads), 0, c->stream, n3, c->rv, S, n, c->scal, g,
                       c->partials, take_tail(c, nb, false));
}

void launch_dot(wrapd_ctx* c, const double* a, const double* b) {
    ProfScope ps(c, "dot");
    const long long n3 = 3ll * c->n_f;
    const int nb = blocks_for(n3, kVecThreads);
    arm_tail(c, 2, 2u, S_RED);
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kVecThreads), 0, c->stream, n3, a, b, c->partials, take_tail(c, nb, false));
}

void launch_diffnorm(wrapd_ctx* c, const double* a, const double* b, int dst) {
    const long long n3 = 3ll * c->n_f;
    const int nb = blocks_for(n3, kVecThreads);
    arm_tail(c, 1, 0u, dst);
    hipLaunchKernelGGL(k_diffnorm, dim3(nb), dim3(kVecThreads), 0, c->stream, n3, a, b, c->partials,
                       take_tail(c, nb, false));
}

void launch_xstep(wrapd_ctx* c, double rate, const double* dirv, double* s_out) {
    ProfScope ps(c, "xstep");
    const long long n3 = 3ll * c->n_f;
    hipLaunchKernelGGL(k_xstep, dim3(blocks_for(n3, kVecThreads)), dim3(kVecThreads), 0, c->stream, n3, c->xall,
                       rate, dirv, s_out);
}

void launch_xstep_r(wrapd_ctx* c, double rate, const int* slots, int n, double* s_out) {
    ProfScope ps(c, "xstep");
    VecPtrs S{};
    for (int j = 0; j < n; ++j) S.v[j] = c->hS[slots[j]];
    const long long n3 = 3ll * c->n_f;
    hipLaunchKernelGGL(k_xstep_r, dim3(blocks_for(n3, kVecThreads)), dim3(kVecThreads), 0, c->stream, n3, c->xall,
                       rate, c->rv, S, n, c->scal, s_out);
}

void launch_armijo_x(wrapd_ctx* c, double alpha, const double* p, double* xbuf) {
    const long long n3 = 3ll * c->n_f;
    hipLaunchKernelGGL(k_axpby, dim3(blocks_for(n3, kVecThreads)), dim3(kVecThreads), 0, c->stream, n3, c->xall,
                       xbuf, alpha, p);
}

void launch_copy(wrapd_ctx* c, double* dst, const double* src, long long n) {
    hipLaunchKernelGGL(k_copy, dim3(blocks_for(n, kVecThreads)), dim3(kVecThreads), 0, c->stream, n, dst, src);
}

void launch_fill(wrapd_ctx* c, double* dst, double v, long long n) {
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, kVecThreads)), dim3(kVecThreads), 0, c->stream, n, dst, v);
}

}  // namespace wd
