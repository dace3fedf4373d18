// Vertex-space kernels: the atomic-free D^T z gather, the SplitLBFGS vector work and
// the deterministic reductions.
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

struct SlotList {
    int n;
    int slot[kMaxHist];
};

struct VecPtrs {
    const double* v[kMaxHist];
};

__global__ void __launch_bounds__(256)
k_finalize(const double* __restrict__ partials, int nblk, int nslots, unsigned maxmask, double* __restrict__ out) {
    __shared__ double sh[4];
    for (int k = 0; k < nslots; ++k) {
        const bool mx = (maxmask >> k) & 1u;
        double acc = mx ? -1e308 : 0.0;
        for (int b = threadIdx.x; b < nblk; b += 256) {
            const double v = partials[b * kRedStride + k];
            acc = mx ? fmax(acc, v) : acc + v;
        }
        const double r = mx ? block_max<256>(acc, sh) : block_sum<256>(acc, sh);
        if (threadIdx.x == 0) out[k] = r;
    }
}

// grad[v] = sum of the staged element-corner gradients of vertex v, in element order.
// mode 1: also |grad - grad_old|^2 (SplitLBFGS.hpp:138).
// mode 2: y_new = grad - grad_old into slot `ns`, and the Gram entries of the new pair.
template <int MODE>
__global__ void __launch_bounds__(kVecThreads)
k_gather(int n_f, const int* __restrict__ ptr, const int* __restrict__ idx, const double* __restrict__ gstage,
         double* __restrict__ out, const double* __restrict__ gold, double* __restrict__ ynew,
         const double* __restrict__ snew, VecPtrs S, VecPtrs Y, SlotList L, int ns_pos,
         double* __restrict__ partials) {
    __shared__ double sh[kVecThreads / 64];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int v = lb * blockDim.x + threadIdx.x;
    double g[3] = {0.0, 0.0, 0.0};
    if (v < n_f) {
        for (int q = ptr[v]; q < ptr[v + 1]; ++q) {
            const double* p = gstage + 3ll * idx[q];
            g[0] += p[0];
            g[1] += p[1];
            g[2] += p[2];
        }
        out[3 * v + 0] = g[0];
        out[3 * v + 1] = g[1];
        out[3 * v + 2] = g[2];
    }
    if (MODE == 1) {
        double a = 0.0;
        if (v < n_f) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double d = g[c] - gold[3 * v + c];
                a += d * d;
            }
        }
        const double t = block_sum<kVecThreads>(a, sh);
        if (threadIdx.x == 0) partials[lb * kRedStride + 0] = t;
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
        for (int j = 0; j < L.n; ++j) {
            double a = 0.0, b = 0.0;
            if (v < n_f) {
                const double* yj = (j == ns_pos) ? nullptr : Y.v[j] + 3 * v;
                const double* sj = (j == ns_pos) ? nullptr : S.v[j] + 3 * v;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double yy = yj ? yj[c] : y[c];
                    const double ss = sj ? sj[c] : s[c];
                    a += s[c] * yy;   // s_new . y_j
                    b += ss * y[c];   // s_j . y_new
                }
            }
            const double ta = block_sum<kVecThreads>(a, sh);
            const double tb = block_sum<kVecThreads>(b, sh);
            if (threadIdx.x == 0) {
                partials[lb * kRedStride + 2 * j] = ta;
                partials[lb * kRedStride + 2 * j + 1] = tb;
            }
        }
        double d = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
        const double td = block_sum<kVecThreads>(d, sh);
        if (threadIdx.x == 0) partials[lb * kRedStride + 2 * L.n] = td;
    }
}

// Store the reduced Gram entries of the new pair and the y.y denominator.
__global__ void k_gram_store(double* __restrict__ scal, SlotList L, int ns_pos) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int sn = L.slot[ns_pos];
    for (int j = 0; j < L.n; ++j) {
        const int sj = L.slot[j];
        scal[S_GRAM + sn * kMaxHist + sj] = scal[S_RED + 2 * j];       // s_new . y_j
        scal[S_GRAM + sj * kMaxHist + sn] = scal[S_RED + 2 * j + 1];   // s_j . y_new
    }
    scal[S_DENOM] = scal[S_RED + 2 * L.n];
}

// partial dots of `base` with up to 8 vectors
__global__ void __launch_bounds__(kVecThreads)
k_multidot(long long n3, const double* __restrict__ base, VecPtrs V, int nvec, double* __restrict__ partials) {
    __shared__ double sh[kVecThreads / 64];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long long i0 = static_cast<long long>(lb) * blockDim.x * 4 + threadIdx.x;
    double acc[kMaxHist];
#pragma unroll
    for (int j = 0; j < kMaxHist; ++j) acc[j] = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long long i = i0 + static_cast<long long>(u) * blockDim.x;
        if (i < n3) {
            const double b = base[i];
#pragma unroll
            for (int j = 0; j < kMaxHist; ++j)
                if (j < nvec) acc[j] += b * V.v[j][i];
        }
    }
    for (int j = 0; j < nvec; ++j) {
        const double t = block_sum<kVecThreads>(acc[j], sh);
        if (threadIdx.x == 0) partials[lb * kRedStride + j] = t;
    }
}

// alpha recurrence (first loop), logical order 0 = oldest
__global__ void k_lbfgs_alpha(double* __restrict__ scal, SlotList L) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = L.n - 1; i >= 0; --i) {
        const int si = L.slot[i];
        double a = scal[S_RED + i];
        for (int j = i + 1; j < L.n; ++j) a -= scal[S_ALPHA + j] * scal[S_GRAM + si * kMaxHist + L.slot[j]];
        const double rho = 1.0 / scal[S_GRAM + si * kMaxHist + si];
        scal[S_ALPHA + i] = rho * a;
    }
}

// beta recurrence (second loop); stores alpha_i - beta_i
__global__ void k_lbfgs_beta(double* __restrict__ scal, SlotList L) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = 0; i < L.n; ++i) {
        const int si = L.slot[i];
        double b = scal[S_RED + i];
        for (int j = 0; j < i; ++j) b += scal[S_COEF + j] * scal[S_GRAM + L.slot[j] * kMaxHist + si];
        const double rho = 1.0 / scal[S_GRAM + si * kMaxHist + si];
        const double beta = rho * b;
        scal[S_COEF + i] = scal[S_ALPHA + i] - beta;
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
         const double* __restrict__ g, double* __restrict__ partials) {
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
    if (threadIdx.x == 0) partials[lb * kRedStride + 0] = t;
}

__global__ void __launch_bounds__(kVecThreads)
k_dot(long long n3, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ partials) {
    __shared__ double sh[kVecThreads / 64];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long long i = static_cast<long long>(lb) * blockDim.x + threadIdx.x;
    double d = 0.0, m = 0.0;
    if (i < n3) { d = a[i] * b[i]; m = fabs(a[i]); }
    const double t = block_sum<kVecThreads>(d, sh);
    const double tm = block_max<kVecThreads>(m, sh);
    if (threadIdx.x == 0) { partials[lb * kRedStride + 0] = t; partials[lb * kRedStride + 1] = tm; }
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
k_diffnorm(long long n3, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ partials) {
    __shared__ double sh[kVecThreads / 64];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long long i = static_cast<long long>(lb) * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (i < n3) { const double t = a[i] - b[i]; d = t * t; }
    const double t = block_sum<kVecThreads>(d, sh);
    if (threadIdx.x == 0) partials[lb * kRedStride + 0] = t;
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

void launch_finalize(wrapd_ctx* c, int nblk, int nslots, unsigned maxmask, int dst) {
    ProfScope ps(c, "finalize");
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, c->stream, c->partials, nblk, nslots, maxmask,
                       c->scal + dst);
}

void launch_gather_grad(wrapd_ctx* c, double* out, int mode, int new_pos, const int* list_slots, int nlist) {
    ProfScope ps(c, "gather_grad");
    const int nb = blocks_for(c->n_f, kVecThreads);
    VecPtrs S{}, Y{};
    SlotList L = mk_list(list_slots, nlist);
    for (int j = 0; j < nlist; ++j) { S.v[j] = c->hS[list_slots[j]]; Y.v[j] = c->hY[list_slots[j]]; }
    double* ynew = nullptr;
    const double* snew = nullptr;
    if (mode == 2) { ynew = c->hY[list_slots[new_pos]]; snew = c->hS[list_slots[new_pos]]; }
    if (mode == 0)
        hipLaunchKernelGGL(k_gather<0>, dim3(nb), dim3(kVecThreads), 0, c->stream, c->n_f, c->v2e_ptr, c->v2e_idx,
                           c->gstage, out, c->grad_old, ynew, snew, S, Y, L, new_pos, c->partials);
    else if (mode == 1)
        hipLaunchKernelGGL(k_gather<1>, dim3(nb), dim3(kVecThreads), 0, c->stream, c->n_f, c->v2e_ptr, c->v2e_idx,
                           c->gstage, out, c->grad_old, ynew, snew, S, Y, L, new_pos, c->partials);
    else
        hipLaunchKernelGGL(k_gather<2>, dim3(nb), dim3(kVecThreads), 0, c->stream, c->n_f, c->v2e_ptr, c->v2e_idx,
                           c->gstage, out, c->grad_old, ynew, snew, S, Y, L, new_pos, c->partials);
    if (mode == 2) {
        launch_finalize(c, nb, 2 * nlist + 1, 0u, S_RED);
        hipLaunchKernelGGL(k_gram_store, dim3(1), dim3(64), 0, c->stream, c->scal, L, new_pos);
    }
}

int finalize_blocks_n(wrapd_ctx* c) { return blocks_for(3ll * c->n_f, kVecThreads * 4); }

void launch_multidot(wrapd_ctx* c, const double* base, double* const* vecs, int nvec) {
    ProfScope ps(c, "multidot");
    VecPtrs V{};
    for (int j = 0; j < nvec; ++j) V.v[j] = vecs[j];
    const long long n3 = 3ll * c->n_f;
    const int nb = blocks_for(n3, kVecThreads * 4);
    hipLaunchKernelGGL(k_multidot, dim3(nb), dim3(kVecThreads), 0, c->stream, n3, base, V, nvec, c->partials);
    launch_finalize(c, nb, nvec, 0u, S_RED);
}

void launch_lbfgs_alpha(wrapd_ctx* c, const int* slots, int n) {
    hipLaunchKernelGGL(k_lbfgs_alpha, dim3(1), dim3(64), 0, c->stream, c->scal, mk_list(slots, n));
}

void launch_lbfgs_beta(wrapd_ctx* c, const int* slots, int n) {
    hipLaunchKernelGGL(k_lbfgs_beta, dim3(1), dim3(64), 0, c->stream, c->scal, mk_list(slots, n));
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
    hipLaunchKernelGGL(k_axpy_r, dim3(nb), dim3(kVecThreads), 0, c->stream, n3, c->rv, S, n, c->scal, g,
                       c->partials);
    launch_finalize(c, nb, 1, 0u, S_DIR);
}

void launch_dot(wrapd_ctx* c, const double* a, const double* b) {
    ProfScope ps(c, "dot");
    const long long n3 = 3ll * c->n_f;
    const int nb = blocks_for(n3, kVecThreads);
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kVecThreads), 0, c->stream, n3, a, b, c->partials);
    launch_finalize(c, nb, 2, 2u, S_RED);
}

void launch_diffnorm(wrapd_ctx* c, const double* a, const double* b) {
    const long long n3 = 3ll * c->n_f;
    const int nb = blocks_for(n3, kVecThreads);
    hipLaunchKernelGGL(k_diffnorm, dim3(nb), dim3(kVecThreads), 0, c->stream, n3, a, b, c->partials);
    launch_finalize(c, nb, 1, 0u, S_RED);
}

void launch_xstep(wrapd_ctx* c, double rate, const double* dirv, double* s_out) {
    ProfScope ps(c, "xstep");
    const long long n3 = 3ll * c->n_f;
    hipLaunchKernelGGL(k_xstep, dim3(blocks_for(n3, kVecThreads)), dim3(kVecThreads), 0, c->stream, n3, c->xall,
                       rate, dirv, s_out);
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
