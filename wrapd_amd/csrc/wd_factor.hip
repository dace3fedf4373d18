// Numeric supernodal Cholesky on the GPU into the solve layout
// G_s = [L11^{-1}; -L21 L11^{-1}] (the numeric half of LinearSolver::factorize,
// deps/mcloptlib/include/MCL/LinearSolver.hpp:279-298, redone on every reweight,
// src/GlobalStepBehavior.hpp:254-258).
//
// The schedule (levels, 64x64 tiles, task lists, pool offsets) comes from chol_plan on
// the host.  Four kernels, one launch per (level, step):
//   k_fassemble  one wave per front column: zero, store A entries, add children's
//                Schur-complement columns in child order (deterministic, no atomics)
//   k_fpotrf     one workgroup per supernode: Cholesky of the panel's diagonal tile in LDS
//   k_ftrsm      TRSM tiles below the diagonal tile, or GDIAG tiles of X = L^{-1}
//   k_fupdate    SYRK/GEMM trailing updates of F, or GEMM updates of X
// All arithmetic is FP64.  The two LDS tiles of k_ftrsm/k_fupdate take 64 KB.
#include "wd_ctx.hpp"
#include "chol_plan.hpp"

namespace wd {

constexpr int kT = wrapd::kFactorTile;   // 64
constexpr int kFactThreads = 256;

struct FactArgs {
    const int* col0;
    const int* fsize;
    const int64_t* foff;
    const int64_t* goff;
    double* pool;
    double* G;
};

__device__ __forceinline__ void front_tile(int t, int ns, int f, int& start, int& size) {
    const int P = (ns + kT - 1) / kT;
    if (t < P) {
        start = kT * t;
        size = min(kT, ns - start);
    } else {
        start = ns + kT * (t - P);
        size = min(kT, f - start);
    }
}

__global__ void k_fill_int(int* __restrict__ p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(64)
k_fassemble(const int2* __restrict__ tasks, FactArgs A, const double* __restrict__ Ax,
            const int64_t* __restrict__ cbase, const int64_t* __restrict__ acol_ptr, const int2* __restrict__ acol,
            const int64_t* __restrict__ ccol_ptr, const int2* __restrict__ ccol, const int64_t* __restrict__ relptr,
            const int* __restrict__ rel) {
    const int2 task = tasks[blockIdx.x];
    const int s = task.x, j = task.y;
    const int lane = threadIdx.x;
    const int f = A.fsize[s];
    const int ns = A.col0[s + 1] - A.col0[s];
    double* F = A.pool + A.foff[s] + static_cast<int64_t>(j) * f;
    for (int r = j + lane; r < f; r += 64) F[r] = 0.0;
    if (j < ns) {
        double* Gc = A.G + A.goff[s] + static_cast<int64_t>(j) * f;
        for (int r = lane; r < f; r += 64) Gc[r] = 0.0;
    }
    __syncthreads();
    const int64_t gc = cbase[s] + j;
    for (int64_t q = acol_ptr[gc] + lane; q < acol_ptr[gc + 1]; q += 64) {
        const int2 e = acol[q];
        F[e.x] = Ax[e.y];
    }
    __syncthreads();
    for (int64_t q = ccol_ptr[gc]; q < ccol_ptr[gc + 1]; ++q) {
        const int2 e = ccol[q];
        const int ch = e.x, jj = e.y;
        const int fc = A.fsize[ch];
        const int nsc = A.col0[ch + 1] - A.col0[ch];
        const int mc = fc - nsc;
        const double* src = A.pool + A.foff[ch] + static_cast<int64_t>(nsc + jj) * fc + nsc;
        const int* rl = rel + relptr[ch];
        for (int ii = jj + lane; ii < mc; ii += 64) F[rl[ii]] += src[ii];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// Cholesky of a diagonal tile held column-major in LDS (D[c * 64 + r], lower part)
__device__ __forceinline__ void lds_potrf(double* D, int nb, int* fail) {
    const int r = threadIdx.x & 63, cg = threadIdx.x >> 6;
    for (int p = 0; p < nb; ++p) {
        if (threadIdx.x == 0) {
            double d = D[p * 65];
            if (!(d > 0.0)) {
                *fail = 1;
                d = 1.0;
            }
            D[p * 65] = sqrt(d);
        }
        __syncthreads();
        if (threadIdx.x < nb - p - 1) D[p * 64 + p + 1 + threadIdx.x] /= D[p * 65];
        __syncthreads();
        if (r > p && r < nb) {
            const double lr = D[p * 64 + r];
            for (int c = p + 1 + cg; c <= r; c += 4) D[c * 64 + r] -= lr * D[p * 64 + c];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kFactThreads)
k_fpotrf(const int2* __restrict__ tasks, FactArgs A, int k, int* __restrict__ fail) {
    __shared__ double D[kT * kT];
    const int s = tasks[blockIdx.x].x;
    const int f = A.fsize[s];
    const int ns = A.col0[s + 1] - A.col0[s];
    const int c0 = kT * k;
    const int nb = min(kT, ns - c0);
    double* F = A.pool + A.foff[s];
    for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
        const int c = idx >> 6, r = idx & 63;
        D[idx] = (c < nb && r < nb && r >= c) ? F[static_cast<int64_t>(c0 + c) * f + c0 + r] : 0.0;
    }
    __syncthreads();
    lds_potrf(D, nb, fail);
    for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
        const int c = idx >> 6, r = idx & 63;
        if (c < nb && r < nb && r >= c) F[static_cast<int64_t>(c0 + c) * f + c0 + r] = D[idx];
    }
}

// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(kFactThreads)
k_ftrsm(const int2* __restrict__ tasks, FactArgs A, int k) {
    __shared__ double D[kT * kT];
    __shared__ double B[kT * kT];
    const int2 task = tasks[blockIdx.x];
    const int s = task.x;
    const int kind = task.y >> wrapd::kTaskKindShift;
    const int t = task.y & ((1 << wrapd::kTaskKindShift) - 1);
    const int f = A.fsize[s];
    const int ns = A.col0[s + 1] - A.col0[s];
    const int c0 = kT * k;
    const int nb = min(kT, ns - c0);
    const double* F = A.pool + A.foff[s];
    for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
        const int c = idx >> 6, r = idx & 63;
        D[idx] = (c < nb && r < nb && r >= c) ? F[static_cast<int64_t>(c0 + c) * f + c0 + r] : 0.0;
    }
    const int r = threadIdx.x & 63, cg = threadIdx.x >> 6;
    int t0, nt;
    front_tile(t, ns, f, t0, nt);
    if (kind == 0) {
        // rows [t0, t0+nt) of panel k: X L_kk^T = B
        double* Fw = A.pool + A.foff[s];
        for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
            const int c = idx >> 6, rr = idx & 63;
            B[idx] = (c < nb && rr < nt) ? Fw[static_cast<int64_t>(c0 + c) * f + t0 + rr] : 0.0;
        }
        __syncthreads();
        for (int p = 0; p < nb; ++p) {
            if (threadIdx.x < nt) B[p * 64 + threadIdx.x] /= D[p * 65];
            __syncthreads();
            if (r < nt) {
                const double xp = B[p * 64 + r];
                for (int c = p + 1 + cg; c < nb; c += 4) B[c * 64 + r] -= xp * D[p * 64 + c];
            }
            __syncthreads();
        }
        for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
            const int c = idx >> 6, rr = idx & 63;
            if (c < nb && rr < nt) Fw[static_cast<int64_t>(c0 + c) * f + t0 + rr] = B[idx];
        }
    } else {
        // X[k, t] = L_kk^{-1} X[k, t]: rows [c0, c0+nb), columns [t0, t0+nt) of G_s
        double* Gs = A.G + A.goff[s];
        for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
            const int j = idx >> 6, rr = idx & 63;
            double v = 0.0;
            if (j < nt && rr < nb) v = (t == k) ? (rr == j ? 1.0 : 0.0) : Gs[static_cast<int64_t>(t0 + j) * f + c0 + rr];
            B[idx] = v;
        }
        __syncthreads();
        for (int p = 0; p < nb; ++p) {
            if (threadIdx.x < nt) B[threadIdx.x * 64 + p] /= D[p * 65];
            __syncthreads();
            if (r > p && r < nb) {
                const double l = D[p * 64 + r];
                for (int j = cg; j < nt; j += 4) B[j * 64 + r] -= l * B[j * 64 + p];
            }
            __syncthreads();
        }
        for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
            const int j = idx >> 6, rr = idx & 63;
            if (j < nt && rr < nb) Gs[static_cast<int64_t>(t0 + j) * f + c0 + rr] = B[idx];
        }
    }
}

// ---------------------------------------------------------------------------------
// C[a, b] -= P_a P_b^T over the panel's nb columns; 16x16 threads, 4x4 outputs each
__global__ void __launch_bounds__(kFactThreads)
k_fupdate(const int2* __restrict__ tasks, FactArgs A, int k) {
    __shared__ double As[kT * kT];   // As[kk * 64 + i]
    __shared__ double Bs[kT * kT];   // Bs[kk * 64 + j]
    const int2 task = tasks[blockIdx.x];
    const int s = task.x;
    const int kind = task.y >> wrapd::kTaskKindShift;
    const int a = (task.y >> wrapd::kTaskTileShift) & wrapd::kTaskTileMask;
    const int b = task.y & wrapd::kTaskTileMask;
    const int f = A.fsize[s];
    const int ns = A.col0[s + 1] - A.col0[s];
    const int c0 = kT * k;
    const int nb = min(kT, ns - c0);
    int a0, na, b0, nbt;
    front_tile(a, ns, f, a0, na);
    front_tile(b, ns, f, b0, nbt);
    double* F = A.pool + A.foff[s];
    double* Gs = A.G + A.goff[s];
    for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
        const int kk = idx >> 6, i = idx & 63;
        As[idx] = (kk < nb && i < na) ? F[static_cast<int64_t>(c0 + kk) * f + a0 + i] : 0.0;
    }
    if (kind == 0) {
        for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
            const int kk = idx >> 6, j = idx & 63;
            Bs[idx] = (kk < nb && j < nbt) ? F[static_cast<int64_t>(c0 + kk) * f + b0 + j] : 0.0;
        }
    } else {
        for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
            const int j = idx >> 6, kk = idx & 63;
            Bs[kk * 64 + j] = (kk < nb && j < nbt) ? Gs[static_cast<int64_t>(b0 + j) * f + c0 + kk] : 0.0;
        }
    }
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) acc[ii][jj] = 0.0;
    for (int kk = 0; kk < nb; ++kk) {
        double av[4], bv[4];
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) av[ii] = As[kk * 64 + tx + 16 * ii];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) bv[jj] = Bs[kk * 64 + ty + 16 * jj];
#pragma unroll
        for (int ii = 0; ii < 4; ++ii)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) acc[ii][jj] = fma(av[ii], bv[jj], acc[ii][jj]);
    }
    double* C = kind == 0 ? F : Gs;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        const int j = ty + 16 * jj;
        if (j >= nbt) continue;
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) {
            const int i = tx + 16 * ii;
            if (i >= na) continue;
            if (kind == 0 && a == b && i < j) continue;   // diagonal tiles: lower part only
            C[static_cast<int64_t>(b0 + j) * f + a0 + i] -= acc[ii][jj];
        }
    }
}

// ---------------------------------------------------------------------------------
template <typename T>
static int up_vec(wrapd_ctx* c, T*& dptr, const std::vector<T>& h) {
    if (dptr) { (void)hipFree(dptr); dptr = nullptr; }
    const size_t bytes = std::max<size_t>(sizeof(T), h.size() * sizeof(T));
    if (check_hip(c, hipMalloc(&dptr, bytes), "hipMalloc(factor plan)")) return 1;
    if (!h.empty() && check_hip(c, hipMemcpy(dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice),
                                "hipMemcpy(factor plan)"))
        return 1;
    return 0;
}

int factor_prepare_device(wrapd_ctx* c) {
    const wrapd::CholSymbolic& S = c->sym;
    try {
        wrapd::chol_plan(S, c->fplan);
    } catch (const std::exception& ex) {
        c->last_error = ex.what();
        return WRAPD_ERR_STATE;
    }
    const wrapd::FactorPlan& P = c->fplan;
    int rc = 0;
    rc |= up_vec(c, c->d_foff, P.foff);
    rc |= up_vec(c, c->d_cbase, P.cbase);
    rc |= up_vec(c, c->d_acol_ptr, P.acol_ptr);
    rc |= up_vec(c, c->d_acol, P.acol);
    rc |= up_vec(c, c->d_ccol_ptr, P.ccol_ptr);
    rc |= up_vec(c, c->d_ccol, P.ccol);
    rc |= up_vec(c, c->d_relptr, S.relptr);
    rc |= up_vec(c, c->d_rel, S.rel);
    rc |= up_vec(c, c->d_ftasks, P.tasks);
    std::vector<int> zero1(1, 0);
    rc |= up_vec(c, c->d_ffail, zero1);
    if (c->pool) { (void)hipFree(c->pool); c->pool = nullptr; }
    rc |= check_hip(c, hipMalloc(&c->pool, std::max<size_t>(8, P.pool_size * sizeof(double))), "hipMalloc(front pool)");
    return rc ? WRAPD_ERR_HIP : WRAPD_OK;
}

// queues the whole factorization on c->stream; the caller reads c->d_ffail afterwards
void launch_factor(wrapd_ctx* c) {
    const wrapd::FactorPlan& P = c->fplan;
    FactArgs A{c->d_col0, c->d_fsize, c->d_foff, c->d_goff, c->pool, c->G};
    const int2* tasks = reinterpret_cast<const int2*>(c->d_ftasks);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->d_ffail, 1, 0);
    for (const wrapd::FactorLaunch& L : P.launches) {
        const int2* t = tasks + L.offset;
        switch (L.kind) {
        case wrapd::FK_ASSEMBLE:
            hipLaunchKernelGGL(k_fassemble, dim3(L.count), dim3(64), 0, c->stream, t, A, c->Ax, c->d_cbase,
                               c->d_acol_ptr, reinterpret_cast<const int2*>(c->d_acol), c->d_ccol_ptr,
                               reinterpret_cast<const int2*>(c->d_ccol), c->d_relptr, c->d_rel);
            break;
        case wrapd::FK_POTRF:
            hipLaunchKernelGGL(k_fpotrf, dim3(L.count), dim3(kFactThreads), 0, c->stream, t, A, L.panel, c->d_ffail);
            break;
        case wrapd::FK_TRSM:
            hipLaunchKernelGGL(k_ftrsm, dim3(L.count), dim3(kFactThreads), 0, c->stream, t, A, L.panel);
            break;
        default:
            hipLaunchKernelGGL(k_fupdate, dim3(L.count), dim3(kFactThreads), 0, c->stream, t, A, L.panel);
            break;
        }
    }
}

}  // namespace wd
