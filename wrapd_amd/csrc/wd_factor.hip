// Numeric supernodal Cholesky on the GPU into the solve layout
// G_s = [L11^{-1}; -L21 L11^{-1}] (the numeric half of LinearSolver::factorize,
// deps/mcloptlib/include/MCL/LinearSolver.hpp:279-298, redone on every reweight,
// src/GlobalStepBehavior.hpp:254-258).
//
// The schedule (levels, 64x64 tiles, task lists, pool offsets) comes from chol_plan on
// the host.  Four kernels, one launch per (level, step):
//   k_fassemble  one wave per front column: zero, store A entries, add children's
//                Schur-complement columns in child order (deterministic, no atomics)
//   k_fpotrf     one workgroup per supernode: Cholesky L_kk of the panel's diagonal tile
//                and its inverse in LDS; L_kk goes to the tile's lower half, the strict
//                lower part of L_kk^{-1} (transposed) to its otherwise unused upper half
//   k_ftrsm      TRSM tiles F[t, k] <- F[t, k] L_kk^{-T}, or GDIAG tiles
//                X[k, c] <- L_kk^{-1} X[k, c], both as products with the stored inverse
//   k_fupdate    SYRK/GEMM updates of the supernode's own later columns, or GEMM
//                updates of X = L11^{-1}
//   k_ffinal     the Schur complement F[a, b] -= L21 L21^T and G's bottom -L21 L11^{-1},
//                each output tile accumulated over all ns columns in registers
// Every product runs on the matrix cores in 64x64 output tiles (v_mfma_f64_16x16x4:
// each of the 4 waves owns a 32x32 quarter as 2x2 16x16 accumulators) over 64-deep
// operand chunks staged in two 32 KB LDS tiles; the accumulator layout puts
// consecutive lanes on consecutive rows, so tile stores are coalesced.  All FP64.
#include "wd_ctx.hpp"
#include "chol_plan.hpp"
#include "device/wd_reduce.hpp"

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
    if (j < ns) {   // L11^{-1} rows only: G's bottom rows are assigned whole later
        double* Gc = A.G + A.goff[s] + static_cast<int64_t>(j) * wrapd::g_ld(f);
        for (int r = lane; r < ns; r += 64) Gc[r] = 0.0;
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
typedef double dbl4 __attribute__((ext_vector_type(4)));

// acc += sum_kk As[kk][i] Bs[kk][j], kk < nk (a multiple of 4, operands zero-padded).
// v_mfma_f64_16x16x4 lane layout: A[row l%16][k l/16], B[k l/16][col l%16], D register q
// of lane l = D[4q + l/16][l%16].  The product is formed as D = Bs^T-side x As-side, so
// that D's column index (consecutive lanes) is the output row i.
__device__ __forceinline__ void tile_zero(dbl4 acc[2][2]) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
}

__device__ __forceinline__ void tile_mma(const double* __restrict__ As, const double* __restrict__ Bs, int nk,
                                         dbl4 acc[2][2]) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int j0 = 32 * (w >> 1), i0 = 32 * (w & 1);
    const int li = l & 15, lk = l >> 4;
    for (int kk = 0; kk < nk; kk += 4) {
        const int k = (kk + lk) * kT;
        const double b0 = Bs[k + j0 + li], b1 = Bs[k + j0 + 16 + li];
        const double a0 = As[k + i0 + li], a1 = As[k + i0 + 16 + li];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a1, acc[1][1], 0, 0, 0);
    }
}

// visits every accumulator entry: fn(i, j, value) with (i, j) in the 64x64 tile; within a
// wave, lanes l and l+1 hold rows i and i+1 of the same column j
template <typename Fn>
__device__ __forceinline__ void tile_for_each(const dbl4 acc[2][2], Fn fn) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int j0 = 32 * (w >> 1), i0 = 32 * (w & 1);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) fn(i0 + 16 * b + (l & 15), j0 + 16 * a + 4 * q + (l >> 4), acc[a][b][q]);
}

// Linv[i][q] of panel (c0, nb) from the diagonal tile: 1/L_ii on the diagonal, the stored
// transpose above it, zero below the diagonal of the inverse's upper part
__device__ __forceinline__ double linv_at(const double* __restrict__ F, int f, int c0, int i, int q) {
    if (q > i) return 0.0;
    if (q == i) return 1.0 / F[static_cast<int64_t>(c0 + i) * f + c0 + i];
    return F[static_cast<int64_t>(c0 + i) * f + c0 + q];
}

// ---------------------------------------------------------------------------------
// Cholesky of the diagonal tile and the inverse of its factor, blocked by 16.
//
// The 64x64 tile D (LDS, column-major D[c * 64 + r]) is factored in four block steps:
// wave 0 factors the 16x16 diagonal block in registers (lane r holds row r; columns are
// broadcast with v_readlane, no barrier inside the block) and inverts it; the panel below
// becomes L_IJ = A_IJ L_JJ^{-T} and the trailing blocks get A_IK -= L_IJ L_KJ^T, both as
// 16x16x16 products on the matrix cores, one block per wave.  The inverse of the whole
// factor is then assembled block row by block row, X_IJ = -X_II sum_K L_IK X_KJ.  Three
// barriers per block step instead of three per column.
__device__ __forceinline__ double lane_bcast(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// the 16x16x4 product of a 16-row and a 16-column operand, 4 k-steps (k = 16): lane l
// supplies a[kc] = Aop[l % 16][4 kc + l / 16] and b[kc] = Bop[4 kc + l / 16][l % 16]
__device__ __forceinline__ dbl4 mma16(const double a[4], const double b[4], dbl4 acc) {
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kc], b[kc], acc, 0, 0, 0);
    return acc;
}

__global__ void __launch_bounds__(kFactThreads)
k_fpotrf(const int2* __restrict__ tasks, FactArgs A, int k, int* __restrict__ fail) {
    __shared__ double D[kT * kT];
    __shared__ double X[kT * kT];   // L^{-1}, column-major, lower blocks only
    const int s = tasks[blockIdx.x].x;
    const int f = A.fsize[s];
    const int ns = A.col0[s + 1] - A.col0[s];
    const int c0 = kT * k;
    const int nb = min(kT, ns - c0);
    const int nblk = (nb + 15) >> 4;   // 16-blocks holding tile columns
    double* F = A.pool + A.foff[s];
    // pad rows/columns beyond nb with the identity so every 16-block factors cleanly
    for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
        const int c = idx >> 6, r = idx & 63;
        D[idx] = (c < nb && r < nb) ? (r >= c ? F[static_cast<int64_t>(c0 + c) * f + c0 + r] : 0.0)
                                    : (r == c ? 1.0 : 0.0);
        X[idx] = 0.0;
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int li = l & 15, lk = l >> 4;
    int bad = 0;
    for (int J = 0; J < nblk; ++J) {
        const int j0 = 16 * J;
        if (w == 0) {
            // Cholesky of D_JJ: lane li holds row li (lanes 16..63 repeat rows, write nothing)
            double a[16], x[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) a[c] = D[(j0 + c) * 64 + j0 + li];
#pragma unroll
            for (int p = 0; p < 16; ++p) {
                double d = lane_bcast(a[p], p);
                if (!(d > 0.0)) {
                    bad = 1;
                    d = 1.0;
                }
                const double sq = sqrt(d);
                const double inv = 1.0 / sq;
                a[p] = li == p ? sq : (li > p ? a[p] * inv : a[p]);
#pragma unroll
                for (int c = p + 1; c < 16; ++c) {
                    const double lc = lane_bcast(a[p], c);   // L[c][p]
                    if (li >= c) a[c] -= a[p] * lc;
                }
            }
            // X_JJ = L_JJ^{-1} by substitution on the identity (row li in x)
#pragma unroll
            for (int c = 0; c < 16; ++c) x[c] = c == li ? 1.0 : 0.0;
#pragma unroll
            for (int p = 0; p < 16; ++p) {
                const double rp = 1.0 / lane_bcast(a[p], p);
#pragma unroll
                for (int c = 0; c <= p; ++c) x[c] = li == p ? x[c] * rp : x[c];
#pragma unroll
                for (int c = 0; c <= p; ++c) {
                    const double xpc = lane_bcast(x[c], p);   // X[p][c]
                    if (li > p) x[c] -= a[p] * xpc;
                }
            }
            if (l < 16) {
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    if (li >= c) {
                        D[(j0 + c) * 64 + j0 + li] = a[c];
                        X[(j0 + c) * 64 + j0 + li] = x[c];
                    }
                }
            }
        }
        __syncthreads();
        // panel: L_IJ = A_IJ X_JJ^T for I > J, one block per wave.  Computed transposed,
        // C^T = X_JJ A_IJ^T, so consecutive lanes hold consecutive rows of L_IJ.
        {
            const int I = J + 1 + w;
            if (w < 3 && I < 4) {
                double a[4], b[4];
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) {
                    const int kk = 4 * kc + lk;
                    a[kc] = X[(j0 + kk) * 64 + j0 + li];       // X_JJ[li][kk]
                    b[kc] = D[(j0 + kk) * 64 + 16 * I + li];   // A_IJ[li][kk]
                }
                dbl4 acc = dbl4{0.0, 0.0, 0.0, 0.0};
                acc = mma16(a, b, acc);   // acc[q] = C^T[4q + lk][li] = L_IJ[li][4q + lk]
#pragma unroll
                for (int q = 0; q < 4; ++q) D[(j0 + 4 * q + lk) * 64 + 16 * I + li] = acc[q];
            }
        }
        __syncthreads();
        // trailing blocks: A_IK -= L_IJ L_KJ^T for J < K <= I < 4
        {
            int t = 0;
            for (int K = J + 1; K < 4; ++K)
                for (int I = K; I < 4; ++I, ++t) {
                    if ((t & 3) != w) continue;
                    double a[4], b[4];
#pragma unroll
                    for (int kc = 0; kc < 4; ++kc) {
                        const int kk = 4 * kc + lk;
                        a[kc] = D[(j0 + kk) * 64 + 16 * K + li];   // L_KJ[li][kk]
                        b[kc] = D[(j0 + kk) * 64 + 16 * I + li];   // L_IJ[li][kk]
                    }
                    dbl4 acc = dbl4{0.0, 0.0, 0.0, 0.0};
                    acc = mma16(a, b, acc);   // acc[q] = (L_IJ L_KJ^T)[li][4q + lk]
#pragma unroll
                    for (int q = 0; q < 4; ++q) D[(16 * K + 4 * q + lk) * 64 + 16 * I + li] -= acc[q];
                }
        }
        __syncthreads();
    }
    if (w == 0 && l == 0 && bad) *fail = 1;
    // off-diagonal blocks of the inverse: X_IJ = -X_II (sum_{K=J}^{I-1} L_IK X_KJ), by block rows
    for (int I = 1; I < nblk; ++I) {
        const int J = w;
        if (J < I) {
            dbl4 acc = dbl4{0.0, 0.0, 0.0, 0.0};   // T = sum_K L_IK X_KJ, acc[q] = T[4q + lk][li]
            for (int K = J; K < I; ++K) {
                double a[4], b[4];
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) {
                    const int kk = 4 * kc + lk;
                    a[kc] = D[(16 * K + kk) * 64 + 16 * I + li];   // L_IK[li][kk]
                    b[kc] = X[(16 * J + li) * 64 + 16 * K + kk];   // X_KJ[kk][li]
                }
                acc = mma16(a, b, acc);
            }
            // acc[q] = T[4q + lk][li] is the B operand layout for k-step q
            double a[4], b[4];
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) {
                a[kc] = X[(16 * I + 4 * kc + lk) * 64 + 16 * I + li];   // X_II[li][4kc + lk]
                b[kc] = acc[kc];
            }
            dbl4 r = dbl4{0.0, 0.0, 0.0, 0.0};
            r = mma16(a, b, r);   // r[q] = (X_II T)[4q + lk][li]
#pragma unroll
            for (int q = 0; q < 4; ++q) X[(16 * J + li) * 64 + 16 * I + 4 * q + lk] = -r[q];
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
        const int c = idx >> 6, rr = idx & 63;
        if (c >= nb || rr >= nb) continue;
        double* dst = F + static_cast<int64_t>(c0 + c) * f + c0 + rr;
        if (rr >= c) *dst = D[idx];                 // L_kk (lower, with the diagonal)
        else *dst = X[rr * 64 + c];                 // (L_kk^{-1})[c][rr] above the diagonal
    }
}

// ---------------------------------------------------------------------------------
// TRSM (kind 0): F[t, k] <- F[t, k] L_kk^{-T};  GDIAG (kind 1): X[k, c] <- L_kk^{-1} X[k, c]
__global__ void __launch_bounds__(kFactThreads)
k_ftrsm(const int2* __restrict__ tasks, FactArgs A, int k) {
    __shared__ double As[kT * kT];   // As[q * 64 + i]
    __shared__ double Bs[kT * kT];   // Bs[q * 64 + j]
    const int2 task = tasks[blockIdx.x];
    const int s = task.x;
    const int kind = task.y >> wrapd::kTaskKindShift;
    const int t = task.y & ((1 << wrapd::kTaskKindShift) - 1);
    const int f = A.fsize[s];
    const int ns = A.col0[s + 1] - A.col0[s];
    const int c0 = kT * k;
    const int nb = min(kT, ns - c0);
    double* F = A.pool + A.foff[s];
    double* Gs = A.G + A.goff[s];
    const int lg = wrapd::g_ld(f);
    int t0, nt;
    front_tile(t, ns, f, t0, nt);
    if (kind == 1 && t == k) {   // X[k, k] = L_kk^{-1}
        for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
            const int j = idx >> 6, i = idx & 63;
            if (j < nb && i < nb) Gs[static_cast<int64_t>(c0 + j) * lg + c0 + i] = linv_at(F, f, c0, i, j);
        }
        return;
    }
    if (kind == 0) {
        // C[i][j] = sum_q B[i][q] Linv[j][q]
        for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
            const int q = idx >> 6, i = idx & 63;
            As[idx] = (q < nb && i < nt) ? F[static_cast<int64_t>(c0 + q) * f + t0 + i] : 0.0;
            Bs[idx] = (q < nb && i < nb) ? linv_at(F, f, c0, i, q) : 0.0;   // here i plays j
        }
    } else {
        // C[i][j] = sum_q Linv[i][q] X[q][j], X = G rows [c0, c0+nb), columns [t0, t0+nt)
        for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
            const int q = idx >> 6, i = idx & 63;
            As[idx] = (q < nb && i < nb) ? linv_at(F, f, c0, i, q) : 0.0;
        }
        for (int idx = threadIdx.x; idx < kT * kT; idx += kFactThreads) {
            const int j = idx >> 6, q = idx & 63;
            Bs[q * 64 + j] = (q < nb && j < nt) ? Gs[static_cast<int64_t>(t0 + j) * lg + c0 + q] : 0.0;
        }
    }
    __syncthreads();
    dbl4 acc[2][2];
    tile_zero(acc);
    tile_mma(As, Bs, (nb + 3) & ~3, acc);
    if (kind == 0) {
        tile_for_each(acc, [&](int i, int j, double v) {
            if (i < nt && j < nb) F[static_cast<int64_t>(c0 + j) * f + t0 + i] = v;
        });
    } else {
        tile_for_each(acc, [&](int i, int j, double v) {
            if (i < nb && j < nt) Gs[static_cast<int64_t>(t0 + j) * lg + c0 + i] = v;
        });
    }
}

// ---------------------------------------------------------------------------------
// SYRK/GEMM (kind 0): F[a, b] -= L[a, k] L[b, k]^T;  GUPD (kind 1): X[i, c] -= L[i, k] X[k, c]
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
    const int lg = wrapd::g_ld(f);
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
            Bs[kk * 64 + j] = (kk < nb && j < nbt) ? Gs[static_cast<int64_t>(b0 + j) * lg + c0 + kk] : 0.0;
        }
    }
    __syncthreads();
    dbl4 acc[2][2];
    tile_zero(acc);
    tile_mma(As, Bs, (nb + 3) & ~3, acc);
    double* C = kind == 0 ? F : Gs;
    const int ldc = kind == 0 ? f : lg;
    const bool diag = kind == 0 && a == b;
    tile_for_each(acc, [&](int i, int j, double v) {
        if (i < na && j < nbt && !(diag && i < j)) C[static_cast<int64_t>(b0 + j) * ldc + a0 + i] -= v;
    });
}


// ---------------------------------------------------------------------------------
// SCHUR (kind 0): F[a, b] -= sum_q L[a, q] L[b, q]^T over all panels q < P, tiles a >= b >= P;
// GBOT (kind 1): G[a, c] = -sum_{q >= c} L[a, q] X[q, c], X = L11^{-1}, rows a >= P
__global__ void __launch_bounds__(kFactThreads)
k_ffinal(const int2* __restrict__ tasks, FactArgs A) {
    __shared__ double As[kT * kT];   // As[kk * 64 + i]
    __shared__ double Bs[kT * kT];   // Bs[kk * 64 + j]
    const int2 task = tasks[blockIdx.x];
    const int s = task.x;
    const int kind = task.y >> wrapd::kTaskKindShift;
    const int a = (task.y >> wrapd::kTaskTileShift) & wrapd::kTaskTileMask;
    const int b = task.y & wrapd::kTaskTileMask;
    const int f = A.fsize[s];
    const int ns = A.col0[s + 1] - A.col0[s];
    const int P = (ns + kT - 1) / kT;
    int a0, na, b0, nbt;
    front_tile(a, ns, f, a0, na);
    front_tile(b, ns, f, b0, nbt);
    const double* F = A.pool + A.foff[s];
    double* Fw = A.pool + A.foff[s];
    double* Gs = A.G + A.goff[s];
    const int lg = wrapd::g_ld(f);
    // chunk q+1 is fetched into registers while chunk q runs on the matrix cores
    constexpr int kPer = kT * kT / kFactThreads;   // 16 entries per thread and operand
    double ra[kPer], rb[kPer];
    auto fetch = [&](int q) {
        const int q0 = kT * q;
        const int nq = min(kT, ns - q0);
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const int idx = threadIdx.x + kFactThreads * r;
            const int kk = idx >> 6, i = idx & 63;
            ra[r] = (kk < nq && i < na) ? F[static_cast<int64_t>(q0 + kk) * f + a0 + i] : 0.0;
            if (kind == 0) {
                rb[r] = (kk < nq && i < nbt) ? F[static_cast<int64_t>(q0 + kk) * f + b0 + i] : 0.0;
            } else {   // transposed: consecutive threads read consecutive rows kk of G
                const int j = idx >> 6, k2 = idx & 63;
                rb[r] = (k2 < nq && j < nbt) ? Gs[static_cast<int64_t>(b0 + j) * lg + q0 + k2] : 0.0;
            }
        }
    };
    dbl4 acc[2][2];
    tile_zero(acc);
    const int qfirst = kind == 0 ? 0 : b;
    fetch(qfirst);
    for (int q = qfirst; q < P; ++q) {
        const int nq = min(kT, ns - kT * q);
        __syncthreads();   // the previous chunk has been consumed
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const int idx = threadIdx.x + kFactThreads * r;
            As[idx] = ra[r];
            if (kind == 0) Bs[idx] = rb[r];
            else Bs[(idx & 63) * 64 + (idx >> 6)] = rb[r];
        }
        __syncthreads();
        if (q + 1 < P) fetch(q + 1);
        tile_mma(As, Bs, (nq + 3) & ~3, acc);
    }
    if (kind == 0) {
        const bool diag = a == b;
        tile_for_each(acc, [&](int i, int j, double v) {
            if (i < na && j < nbt && !(diag && i < j)) Fw[static_cast<int64_t>(b0 + j) * f + a0 + i] -= v;
        });
    } else {
        tile_for_each(acc, [&](int i, int j, double v) {
            if (i < na && j < nbt) Gs[static_cast<int64_t>(b0 + j) * lg + a0 + i] = -v;
        });
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
        case wrapd::FK_UPDATE:
            hipLaunchKernelGGL(k_fupdate, dim3(L.count), dim3(kFactThreads), 0, c->stream, t, A, L.panel);
            break;
        default:
            hipLaunchKernelGGL(k_ffinal, dim3(L.count), dim3(kFactThreads), 0, c->stream, t, A);
            break;
        }
    }
}

}  // namespace wd
