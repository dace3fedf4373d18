// Numeric supernodal Cholesky on the GPU into the solve layout
// G_s = [L11^{-1}; -L21 L11^{-1}] (the numeric half of LinearSolver::factorize,
// deps/mcloptlib/include/MCL/LinearSolver.hpp:279-298, redone on every reweight,
// src/GlobalStepBehavior.hpp:254-258).
//
// The schedule (levels, 64x64 tiles, task lists, pool offsets) comes from chol_plan on
// the host.  One launch per (level, step) and kernel:
//   k_fassemble  one wave per chunk of a front column, built in LDS: A entries, then the
//                children's Schur-complement columns in child order (deterministic, no
//                atomics), stored once; a level's fronts in groups, each after the previous
//                level's FINAL launch that its children are in
//   k_fpotrf     one workgroup per supernode: Cholesky L_kk of panel 0's diagonal tile and
//                its inverse in LDS; L_kk goes to the tile's lower half, L_kk^{-1} to G's
//                diagonal tile (k, k), which is also the diagonal tile of L11^{-1}
//   k_fpotrf16   the same for supernodes of at most 16 columns, one wave in registers
//   k_ftrsm      TRSM tiles F[t, k] <- F[t, k] L_kk^{-T} of L11, or GDIAG tiles
//                X[k, c] <- L_kk^{-1} X[k, c], both as products with the stored inverse
//   k_fupdate    SYRK/GEMM updates of L11's later columns, or GEMM updates of X = L11^{-1};
//                the task updating the next panel's diagonal tile then factors it
//                (k_fpotrf's work), so only panel 0 has a POTRF launch
//   k_fl21       L21 by panel, left-looking, on levels where the panel steps cover L11 only
//   k_ffinal     once X is complete: the Schur complement F[a, b] -= L21 L21^T and G's
//                bottom -L21 X, each 64x64 output tile accumulated over its whole k range
// The panel steps touch only the supernode's diagonal block (ns x ns), so the m update rows
// are read and written once per pass instead of once per 64-column panel.
// Every panel-step product runs on the matrix cores in 64x64 output tiles
// (v_mfma_f64_16x16x4: with 4 waves each owns a 32x32 quarter as 2x2 16x16 accumulators,
// with 8 waves (the update, TRSM and L21 kernels) a 32x16 part as 2x1) over 64-deep operand
// chunks staged in LDS; the accumulator layout puts consecutive lanes on consecutive rows, so
// tile stores are coalesced.  Tiles are loaded with every thread's loads (16 with 4 waves, 8
// with 8) in flight at once (tile_load).  All FP64.
#include "wd_ctx.hpp"
#include "chol_plan.hpp"
#include "device/wd_reduce.hpp"

namespace wd {

constexpr int kT = wrapd::kFactorTile;   // 64
constexpr int kFactThreads = 256;
// the L21 kernel's workgroup: eight waves, each a 32x16 part of the 64x64 tile, so each
// thread holds half the operand and accumulator registers of a four-wave one
constexpr int kL21Threads = 512;
constexpr int kTrsmThreads = 512;   // the TRSM kernel's workgroup, split the same way
// the update kernel's: the same split; its diagonal task's factorization runs on the first
// four waves (the others only take part in its barriers)
constexpr int kUpdThreads = 512;

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
// One wave per chunk q of front column j: rows [max(j, q C), min(f, (q + 1) C)), C the plan's
// chunk, are built in LDS (A's entries, then each child's Schur-complement column added in
// child order: deterministic, no atomics) and written to the front once, so the front is
// neither zeroed in memory nor read back.  The plan's per-child chunk table gives the child
// rows that land in the chunk (rel is increasing), so each child value is read once.  The
// column's first chunk also zeroes rows < ns of G's column j (L11^{-1}'s) for the TRSM and
// update steps.  A task needs at most C doubles of LDS (the launch sizes it to its largest
// chunk), so many waves share a compute unit and their global loads overlap.
struct AsmArgs {
    const double* Ax;
    const int64_t* cbase;
    const int64_t* acol_ptr;
    const int2* acol;
    const int64_t* ccol_ptr;
    const int2* ccol;
    const int64_t* relptr;
    const int* rel;
    const int64_t* cchunk_ptr;
    const int* cchunk;
    int chunk;
};

// lane k's 64-bit value, for every lane
__device__ __forceinline__ int64_t bcast64(int64_t v, int k) {
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(v), k);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(v >> 32), k);
    return (static_cast<int64_t>(hi) << 32) | static_cast<uint32_t>(lo);
}

__global__ void __launch_bounds__(64)
k_fassemble(const int2* __restrict__ tasks, FactArgs A, AsmArgs M) {
    extern __shared__ double col[];   // rows r0 .. r1 - 1 of the column
    const int2 task = tasks[blockIdx.x];
    const int s = task.x, j = task.y & 0xFFFF, q = task.y >> 16;
    const int lane = threadIdx.x;
    const int f = A.fsize[s];
    const int ns = A.col0[s + 1] - A.col0[s];
    const int r0 = max(j, q * M.chunk), r1 = min(f, (q + 1) * M.chunk);
    const int len = r1 - r0;
    for (int r = lane; r < len; r += 64) col[r] = 0.0;
    if (j < ns && q == j / M.chunk) {   // L11^{-1}'s rows only: G's bottom rows are assigned whole later
        double* Gc = A.G + A.goff[s] + static_cast<int64_t>(j) * wrapd::g_ld(f);
        for (int r = lane; r < ns; r += 64) Gc[r] = 0.0;
    }
    __syncthreads();
    const int64_t gc = M.cbase[s] + j;
    const int64_t k0 = M.ccol_ptr[gc], nkid = M.ccol_ptr[gc + 1] - k0;
    for (int64_t e = M.acol_ptr[gc] + lane; e < M.acol_ptr[gc + 1]; e += 64) {
        const int2 a = M.acol[e];
        if (a.x >= r0 && a.x < r1) col[a.x - r0] = M.Ax[a.y];
    }
    // The children's contributions, 64 children at a time: lane i loads child i's sizes,
    // offsets and chunk bounds (all children's loads in flight together), and the loop over
    // the children, in order, takes them from that lane.
    for (int64_t kb = 0; kb < nkid; kb += 64) {
        const int nb = static_cast<int>(min<int64_t>(64, nkid - kb));
        int64_t msrc = 0, mrel = 0;
        int mlo = 0, mhi = 0;
        if (lane < nb) {
            const int2 e = M.ccol[k0 + kb + lane];
            const int ch = e.x, jj = e.y;
            const int fc = A.fsize[ch];
            const int nsc = A.col0[ch + 1] - A.col0[ch];
            msrc = A.foff[ch] + static_cast<int64_t>(nsc + jj) * fc + nsc;
            mrel = M.relptr[ch];
            const int* cs = M.cchunk + M.cchunk_ptr[ch];
            mlo = max(jj, cs[q]);
            mhi = cs[q + 1];
        }
        for (int k = 0; k < nb; ++k) {
            __syncthreads();   // the previous contribution (or A's entries) is in
            const double* src = A.pool + bcast64(msrc, k);
            const int* rl = M.rel + bcast64(mrel, k);
            const int lo = __builtin_amdgcn_readlane(mlo, k), hi = __builtin_amdgcn_readlane(mhi, k);
            // four rows per lane per round, their loads issued together (clamped in range)
            for (int i0 = lo + lane; i0 < hi; i0 += 4 * 64) {
                int rr[4];
                double vv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int ii = min(i0 + 64 * u, hi - 1);
                    rr[u] = rl[ii];
                    vv[u] = src[ii];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (i0 + 64 * u < hi) col[rr[u] - r0] += vv[u];
            }
        }
    }
    __syncthreads();
    double* F = A.pool + A.foff[s] + static_cast<int64_t>(j) * f + r0;
    for (int r = lane; r < len; r += 64) F[r] = col[r];
}

// ---------------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

// acc += sum_kk As[kk][i] Bs[kk][j], kk < nk (a multiple of 4, operands zero-padded).
// v_mfma_f64_16x16x4 lane layout: A[row l%16][k l/16], B[k l/16][col l%16], D register q
// of lane l = D[4q + l/16][l%16].  The product is formed as D = Bs^T-side x As-side, so
// that D's column index (consecutive lanes) is the output row i.
// How NT threads split a 64x64 tile: NT / 64 waves, two along the output rows i (32 rows
// each) and the rest along the columns j (SJ columns, NBJ 16-column blocks each); tile_load
// gives each thread R entries.
template <int NT>
struct TileWaves {
    static constexpr int NW = NT / 64;
    static constexpr int SJ = 64 / (NW / 2);
    static constexpr int NBJ = SJ / 16;
    static constexpr int R = 64 / NW;
};

template <int NBJ>
__device__ __forceinline__ void tile_zero(dbl4 (&acc)[NBJ][2]) {
#pragma unroll
    for (int a = 0; a < NBJ; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
}

template <int NT = kFactThreads>
__device__ __forceinline__ void tile_mma(const double* __restrict__ As, const double* __restrict__ Bs, int nk,
                                         dbl4 (&acc)[TileWaves<NT>::NBJ][2]) {
    constexpr int NBJ = TileWaves<NT>::NBJ;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int j0 = TileWaves<NT>::SJ * (w >> 1), i0 = 32 * (w & 1);
    const int li = l & 15, lk = l >> 4;
    for (int kk = 0; kk < nk; kk += 4) {
        const int k = (kk + lk) * kT;
        double b[NBJ];
#pragma unroll
        for (int x = 0; x < NBJ; ++x) b[x] = Bs[k + j0 + 16 * x + li];
        const double a0 = As[k + i0 + li], a1 = As[k + i0 + 16 + li];
#pragma unroll
        for (int x = 0; x < NBJ; ++x) {
            acc[x][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[x], a0, acc[x][0], 0, 0, 0);
            acc[x][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[x], a1, acc[x][1], 0, 0, 0);
        }
    }
}

// Operand loads go through buffer resources (a 32-bit byte offset per lane, a uniform base
// in scalar registers): addresses cost no 64-bit registers, and an offset past the block's
// end reads as zero, so no load needs a clamped address.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t blk_rsrc(const double* base, int64_t n_doubles) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), static_cast<short>(0),
                                             static_cast<int>(n_doubles * 8), 0x00020000);
}
__device__ __forceinline__ double blk_ld(__amdgpu_buffer_rsrc_t r, uint32_t elem) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, elem * 8u, 0, 0));
}
// Entries that must read as zero are loaded from past the buffer's end (the load returns
// zero) rather than masked after the load, so a prefetch does not wait for its data where it
// is issued: element indexes are below 2^28 (every buffer here is under 2 GB), and setting
// bit 28 puts the byte offset past 2^31.  Arithmetic, not a select, so the compiler does
// not split the load into two branches (whose common destination costs a full wait).
__device__ __forceinline__ uint32_t zero_unless(bool ok, uint32_t elem) {
    return elem | (static_cast<uint32_t>(!ok) << 28);
}
// a buffer window for a tile given by its first entry: 2^28 - 1 doubles (the tile's entries
// lie within 64 leading dimensions of it)
constexpr int64_t kTileWindow = (int64_t{1} << 28) - 1;

// 64x64 tile of a column-major block (leading dimension ld) into registers: thread t's
// entry r (0..15) is (row t % 64, column t / 64 + 4 r), that is LDS index t + 256 r of a
// column-major tile.  Rows >= nr and columns >= nc read as zero (from past the buffer's
// end), so the 16 loads carry no branch or select and stay in flight together.
template <int NT = kFactThreads>
__device__ __forceinline__ void tile_load(const double* __restrict__ src, int64_t ld, int nr, int nc,
                                          double (&v)[TileWaves<NT>::R]) {
    constexpr int NW = TileWaves<NT>::NW;
    const __amdgpu_buffer_rsrc_t rs = blk_rsrc(src, kTileWindow);
    const int row = threadIdx.x & 63, col0 = threadIdx.x >> 6;
    const uint32_t o = row + col0 * static_cast<uint32_t>(ld);
#pragma unroll
    for (int r = 0; r < TileWaves<NT>::R; ++r)
        v[r] = blk_ld(rs, zero_unless(row < nr && col0 + NW * r < nc, o + static_cast<uint32_t>(NW * r * ld)));
}

// tile_load_h through a buffer resource (rows and columns past nr, nc read as zero): the
// entries of half h (columns 32 h .. 32 h + 31) of tile_load's
template <int NT = kFactThreads>
__device__ __forceinline__ void tile_ldb_h(__amdgpu_buffer_rsrc_t rs, uint32_t off0, uint32_t ld, int nr, int nc,
                                           double (&v)[TileWaves<NT>::R], int h) {
    constexpr int NW = TileWaves<NT>::NW, H = TileWaves<NT>::R / 2;
    const int row = threadIdx.x & 63, col0 = threadIdx.x >> 6;
    const uint32_t o = off0 + row + col0 * ld;
#pragma unroll
    for (int r = H * h; r < H * h + H; ++r)
        v[r] = blk_ld(rs, zero_unless(row < nr && col0 + NW * r < nc, o + static_cast<uint32_t>(NW * r) * ld));
}
// the registers of tile_load into an LDS tile, as loaded (T[col * 64 + row]) or transposed
template <int NT = kFactThreads>
__device__ __forceinline__ void tile_to_lds(double* T, const double (&v)[TileWaves<NT>::R]) {
#pragma unroll
    for (int r = 0; r < TileWaves<NT>::R; ++r) T[threadIdx.x + NT * r] = v[r];
}
template <int NT = kFactThreads>
__device__ __forceinline__ void tile_to_lds_t(double* T, const double (&v)[TileWaves<NT>::R]) {
    constexpr int NW = TileWaves<NT>::NW;
    const int row = threadIdx.x & 63, col0 = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < TileWaves<NT>::R; ++r) T[row * 64 + col0 + NW * r] = v[r];
}

// Half-depth staging: the k range 32 h .. 32 h + 31 of a 64-deep product into a 32 x 64 LDS
// operand (T[(k - 32 h) * 64 + i]), so a product takes 32 KB of LDS instead of 64 KB and is
// run as tile_mma over each half in turn: the same products in the same k order.
// tile_to_lds_h: the k index is the tile's column (entries r = 8 h .. 8 h + 7 of tile_load);
// tile_to_lds_t_h: the k index is the tile's row (threads of rows 32 h .. 32 h + 31).
constexpr int kHalf = kT / 2;
template <int NT = kFactThreads>
__device__ __forceinline__ void tile_to_lds_h(double* T, const double (&v)[TileWaves<NT>::R], int h) {
    constexpr int H = TileWaves<NT>::R / 2;
#pragma unroll
    for (int r = 0; r < H; ++r) T[threadIdx.x + NT * r] = v[H * h + r];
}
template <int NT = kFactThreads>
__device__ __forceinline__ void tile_to_lds_t_h(double* T, const double (&v)[TileWaves<NT>::R], int h) {
    constexpr int NW = TileWaves<NT>::NW;
    const int row = threadIdx.x & 63, col0 = threadIdx.x >> 6;
    if ((row >> 5) == h) {
#pragma unroll
        for (int r = 0; r < TileWaves<NT>::R; ++r) T[(row - kHalf * h) * 64 + col0 + NW * r] = v[r];
    }
}
__device__ __forceinline__ int half_depth(int nk, int h) {   // the half's k count, a multiple of 4
    const int n = min(kHalf, nk - kHalf * h);
    return n > 0 ? (n + 3) & ~3 : 0;
}

// accumulator tile into a column-major block C (leading dimension ldc): C[i][j] -= acc
// (kSub) or C[i][j] = scale * acc, for i < nr, j < nc, and i >= j when lower_only.  With
// kSub the 16 old values are loaded first, all at once (clamped addresses).
template <bool kSub, int NT = kFactThreads>
__device__ __forceinline__ void tile_write(double* __restrict__ C, int64_t ldc, int nr, int nc, bool lower_only,
                                           const dbl4 (&acc)[TileWaves<NT>::NBJ][2], double scale) {
    constexpr int NBJ = TileWaves<NT>::NBJ;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int j0 = TileWaves<NT>::SJ * (w >> 1), i0 = 32 * (w & 1);
    double old[NBJ * 8];
    if (kSub) {
#pragma unroll
        for (int a = 0; a < NBJ; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = i0 + 16 * b + (l & 15), j = j0 + 16 * a + 4 * q + (l >> 4);
                    old[(a * 2 + b) * 4 + q] =
                        C[static_cast<int64_t>(j < nc ? j : 0) * ldc + (i < nr ? i : 0)];
                }
    }
#pragma unroll
    for (int a = 0; a < NBJ; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = i0 + 16 * b + (l & 15), j = j0 + 16 * a + 4 * q + (l >> 4);
                if (i < nr && j < nc && !(lower_only && i < j)) {
                    double* d = C + static_cast<int64_t>(j) * ldc + i;
                    *d = kSub ? old[(a * 2 + b) * 4 + q] - acc[a][b][q] : scale * acc[a][b][q];
                }
            }
}

// ---------------------------------------------------------------------------------
// Cholesky of the diagonal tile and the inverse of its factor, blocked by 16.
//
// The 64x64 tile D (LDS, column-major D[c * 64 + r]) is factored in four block steps:
// wave 0 factors the 16x16 diagonal block in registers and inverts it (lane (r, g) holds row
// r of columns 4 m + g; columns move between lanes by ds_bpermute, no barrier inside the
// block); the panel below becomes L_IJ = A_IJ L_JJ^{-T} and the trailing blocks get
// A_IK -= L_IJ L_KJ^T, both as 16x16x16 products on the matrix cores, one block per wave.
// Block row J of the inverse, X_JI = -X_JJ sum_K L_JK X_KI, is formed by the waves the
// panel step leaves idle.  Three barriers per block step instead of three per column.
__device__ __forceinline__ double lane_bcast(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// lane src's v, for each lane (src may differ between lanes; every lane of the wave active)
__device__ __forceinline__ double lane_fetch(double v, int src) {
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

// the 16x16x4 product of a 16-row and a 16-column operand, 4 k-steps (k = 16): lane l
// supplies a[kc] = Aop[l % 16][4 kc + l / 16] and b[kc] = Bop[4 kc + l / 16][l % 16]
__device__ __forceinline__ dbl4 mma16(const double a[4], const double b[4], dbl4 acc) {
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kc], b[kc], acc, 0, 0, 0);
    return acc;
}

// The Cholesky of a 16x16 block and its inverse by one wave, in registers: lane (li, g = lane
// / 16) holds row li of columns 4 m + g (m < 4) in a[m] (the block, then L) and x[m] (the
// identity, then L^{-1}); columns move between lanes by ds_bpermute.  L^{-1} is formed by
// substitution on the identity, its step p right after the Cholesky's column p, which is
// final then.  Every entry gets the operations of the row-per-lane loop, in the same order.
// A pivot that is not positive is taken as 1 and sets bad.  All 64 lanes take part.
__device__ __forceinline__ void block16_factor(double a[4], double x[4], int li, int g, int& bad) {
    double d = lane_bcast(a[0], 0);   // A[0][0]
#pragma unroll
    for (int p = 0; p < 16; ++p) {
        const int gp = p & 3, mp = p >> 2;   // column p: a[mp] of the lanes of group gp
        if (!(d > 0.0)) {
            bad = 1;
            d = 1.0;
        }
        // A[p+1][p+1] with the updates of the columns before p; column p's update of it
        // is the operation its own lane applies below, done here as well so the next
        // pivot does not wait for the column exchange
        double dn = p < 15 ? lane_bcast(a[(p + 1) >> 2], (p + 1) + 16 * ((p + 1) & 3)) : 0.0;
        // correctly rounded: a hardware estimate with Newton steps saved 2.4 us per
        // tile but changed L in the last bits, enough to move the sample runs past
        // the reference's own round-off spread
        const double sq = sqrt(d);
        const double inv = 1.0 / sq;
        if (g == gp) a[mp] = li == p ? sq : a[mp] * inv;   // rows above p: never read
        if (p < 15) {
            const double l1 = lane_bcast(a[mp], (p + 1) + 16 * gp);   // L[p+1][p]
            dn -= l1 * l1;
        }
        const double lp = lane_fetch(a[mp], li + 16 * gp);   // L[li][p]
#pragma unroll
        for (int m = mp; m < 4; ++m) {
            const double lc = lane_fetch(a[mp], 4 * m + g + 16 * gp);   // L[4m+g][p]
            if (m > mp || g > gp) a[m] -= lp * lc;   // columns after p (rows above: never read)
        }
        const double xs = li == p ? inv : 1.0;   // row p of X is final after this
#pragma unroll
        for (int m = 0; m <= mp; ++m) x[m] *= xs;
#pragma unroll
        for (int m = 0; m <= mp; ++m) {
            const double xpc = lane_fetch(x[m], p + 16 * g);   // X[p][4m+g]
            if (li > p && (m < mp || g <= gp)) x[m] -= lp * xpc;
        }
        d = dn;
    }
}

// The inverse X in LDS: entry (r, c) at c * 64 + (r ^ (c % 16)), so that the 16 lanes
// reading one row across 16 columns hit 16 different banks.
__device__ __forceinline__ int xs_at(int r, int c) { return c * 64 + (r ^ (c & 15)); }

// X_IJ = -X_II (sum_{K=J}^{I-1} L_IK X_KJ), by one wave: L from D, X_II and X_KJ (K < I) final
__device__ __forceinline__ void inverse_block(const double* D, double* X, int I, int J) {
    const int l = threadIdx.x & 63;
    const int li = l & 15, lk = l >> 4;
    dbl4 acc = dbl4{0.0, 0.0, 0.0, 0.0};   // T = sum_K L_IK X_KJ, acc[q] = T[4q + lk][li]
    for (int K = J; K < I; ++K) {
        double a[4], b[4];
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) {
            const int kk = 4 * kc + lk;
            a[kc] = D[(16 * K + kk) * 64 + 16 * I + li];   // L_IK[li][kk]
            b[kc] = X[xs_at(16 * K + kk, 16 * J + li)];     // X_KJ[kk][li]
        }
        acc = mma16(a, b, acc);
    }
    // acc[q] = T[4q + lk][li] is the B operand layout for k-step q
    double a[4], b[4];
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
        a[kc] = X[xs_at(16 * I + li, 16 * I + 4 * kc + lk)];   // X_II[li][4kc + lk]
        b[kc] = acc[kc];
    }
    dbl4 r = dbl4{0.0, 0.0, 0.0, 0.0};
    r = mma16(a, b, r);   // r[q] = (X_II T)[4q + lk][li]
#pragma unroll
    for (int q = 0; q < 4; ++q) X[xs_at(16 * I + 4 * q + lk, 16 * J + li)] = -r[q];
}

// The 64x64 tile D (LDS, column-major D[c * 64 + r], lower triangle; rows and columns from
// nb on hold the identity) is factored in place, L^{-1} built in X (LDS, zeroed by the
// caller, entry (r, c) at xs_at(r, c)); then L goes to the front's lower triangle at Fkk
// (leading dimension f) and L^{-1} (zero above the diagonal) to Gkk (leading dimension lg).
// The first 256 threads do the work; a caller with more waves brings them to every barrier
// (they do nothing else).  The caller has synchronized after filling D and X.
__device__ __forceinline__ void potrf_lds(double* D, double* X, int nb, double* __restrict__ Fkk, int64_t f,
                          double* __restrict__ Gkk, int64_t lg, int* __restrict__ fail) {
    const int nblk = (nb + 15) >> 4;   // 16-blocks holding tile columns
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int li = l & 15, lk = l >> 4;
    int bad = 0;
    for (int J = 0; J < nblk; ++J) {
        const int j0 = 16 * J;
        if (w == 0) {
            // Cholesky of D_JJ and L_JJ^{-1} (block16_factor's lane layout)
            const int g = lk;
            double a[4], x[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                a[m] = D[(j0 + 4 * m + g) * 64 + j0 + li];
                x[m] = 4 * m + g == li ? 1.0 : 0.0;
            }
            block16_factor(a, x, li, g, bad);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int c = 4 * m + g;
                if (li >= c) {
                    D[(j0 + c) * 64 + j0 + li] = a[m];
                    X[xs_at(j0 + li, j0 + c)] = x[m];
                }
            }
        }
        __syncthreads();
        // block row J of the inverse, block J2 < J on wave 3 - J2: it reads L_JK (K < J), X_JJ
        // and the earlier block rows, none of which the panel below writes
        if (w < 4 && w > 2 - J && 3 - w < J) inverse_block(D, X, J, 3 - w);
        // panel: L_IJ = A_IJ X_JJ^T for I > J, one block per wave.  Computed transposed,
        // C^T = X_JJ A_IJ^T, so consecutive lanes hold consecutive rows of L_IJ.
        {
            const int I = J + 1 + w;
            if (w < 3 && I < 4) {
                double a[4], b[4];
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) {
                    const int kk = 4 * kc + lk;
                    a[kc] = X[xs_at(j0 + li, j0 + kk)];       // X_JJ[li][kk]
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
    // L_kk to the front's lower triangle, L_kk^{-1} (zero above the diagonal) to G's
    // diagonal tile (k, k), where the TRSM tasks read it and the solves use it
    {
        const int r = threadIdx.x & 63, cb = threadIdx.x >> 6;
        if (r < nb && threadIdx.x < kFactThreads) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int c = cb + 4 * q;
                if (c < nb) {
                    if (r >= c) Fkk[static_cast<int64_t>(c) * f + r] = D[threadIdx.x + kFactThreads * q];
                    Gkk[c * lg + r] = X[xs_at(r, c)];
                }
            }
        }
    }
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
    double* F = A.pool + A.foff[s];
    double* Fkk = F + static_cast<int64_t>(c0) * f + c0;
    {
        // the lower triangle; rows/columns beyond nb padded with the identity so every
        // 16-block factors cleanly
        double v[16];
        tile_load(Fkk, f, nb, nb, v);
        const int r = threadIdx.x & 63, cb = threadIdx.x >> 6;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = cb + 4 * q;
            D[threadIdx.x + kFactThreads * q] = r < c ? 0.0 : (c < nb ? v[q] : (r == c ? 1.0 : 0.0));
            X[threadIdx.x + kFactThreads * q] = 0.0;
        }
    }
    __syncthreads();
    const int64_t lg = wrapd::g_ld(f);
    potrf_lds(D, X, nb, Fkk, f, A.G + A.goff[s] + static_cast<int64_t>(c0) * lg + c0, lg, fail);
}

// Panel 0's diagonal tile of a supernode of at most 16 columns (FK_POTRF16): potrf_lds's
// only block step for such a tile (rows and columns from nb on are the identity; nothing of
// them is stored) by one wave, in registers, one supernode per 64-thread workgroup.
__global__ void __launch_bounds__(64)
k_fpotrf16(const int2* __restrict__ tasks, FactArgs A, int* __restrict__ fail) {
    const int s = tasks[blockIdx.x].x;
    const int f = A.fsize[s];
    const int nb = A.col0[s + 1] - A.col0[s];
    double* Fkk = A.pool + A.foff[s];
    double* Gkk = A.G + A.goff[s];
    const int64_t lg = wrapd::g_ld(f);
    const int li = threadIdx.x & 15, g = threadIdx.x >> 4;
    double a[4], x[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int c = 4 * m + g;
        a[m] = li < c ? 0.0 : (c < nb ? (li < nb ? Fkk[static_cast<int64_t>(c) * f + li] : 0.0) : (li == c ? 1.0 : 0.0));
        x[m] = c == li ? 1.0 : 0.0;
    }
    int bad = 0;
    block16_factor(a, x, li, g, bad);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int c = 4 * m + g;
        if (li < nb && c < nb) {
            if (li >= c) Fkk[static_cast<int64_t>(c) * f + li] = a[m];
            Gkk[c * lg + li] = li >= c ? x[m] : 0.0;
        }
    }
    if (threadIdx.x == 0 && bad) *fail = 1;
}

// ---------------------------------------------------------------------------------
// TRSM (kind 0): F[t, k] <- F[t, k] L_kk^{-T};  GDIAG (kind 1): X[k, c] <- L_kk^{-1} X[k, c],
// both as products with L_kk^{-1}, which k_fpotrf left in G's diagonal tile (k, k)
template <int NT>
__global__ void __launch_bounds__(NT)
k_ftrsm(const int2* __restrict__ tasks, FactArgs A, int k) {
    constexpr int NBJ = TileWaves<NT>::NBJ, R = TileWaves<NT>::R;
    __shared__ double As[kHalf * kT];   // As[(q - 32 h) * 64 + i], half h of the k range
    __shared__ double Bs[kHalf * kT];   // Bs[(q - 32 h) * 64 + j]
    const int2 task = tasks[blockIdx.x];
    const int s = task.x;
    const int kind = wrapd::task_kind(task.y);
    const int t = task.y & ((1 << wrapd::kTaskKindShift) - 1);
    const int f = A.fsize[s];
    const int ns = A.col0[s + 1] - A.col0[s];
    const int c0 = kT * k;
    const int nb = min(kT, ns - c0);
    double* F = A.pool + A.foff[s];
    double* Gs = A.G + A.goff[s];
    const int lg = wrapd::g_ld(f);
    if (kind == 1 && t == k) return;   // X[k, k] = L_kk^{-1}: stored by k_fpotrf
    int t0, nt;
    front_tile(t, ns, f, t0, nt);
    const double* Linv = Gs + static_cast<int64_t>(c0) * lg + c0;   // column-major, leading dimension lg
    double va[R], vb[R];
    if (kind == 0) {
        // C[i][j] = sum_q B[i][q] Linv[j][q]: As[q][i] = B[i][q], Bs[q][j] = Linv[j][q]
        tile_load<NT>(F + static_cast<int64_t>(c0) * f + t0, f, nt, nb, va);
        tile_load<NT>(Linv, lg, nb, nb, vb);
    } else {
        // C[i][j] = sum_q Linv[i][q] X[q][j], X = G rows [c0, c0+nb), columns [t0, t0+nt)
        tile_load<NT>(Linv, lg, nb, nb, va);
        tile_load<NT>(Gs + static_cast<int64_t>(t0) * lg + c0, lg, nb, nt, vb);
    }
    dbl4 acc[NBJ][2];
    tile_zero(acc);
    for (int h = 0; h < 2; ++h) {
        const int nh = half_depth(nb, h);
        if (nh == 0) break;
        if (h) __syncthreads();   // the first half has been read
        tile_to_lds_h<NT>(As, va, h);
        if (kind == 0) tile_to_lds_h<NT>(Bs, vb, h);
        else tile_to_lds_t_h<NT>(Bs, vb, h);
        __syncthreads();
        tile_mma<NT>(As, Bs, nh, acc);
    }
    if (kind == 0) tile_write<false, NT>(F + static_cast<int64_t>(c0) * f + t0, f, nt, nb, false, acc, 1.0);
    else tile_write<false, NT>(Gs + static_cast<int64_t>(t0) * lg + c0, lg, nb, nt, false, acc, 1.0);
}

// ---------------------------------------------------------------------------------
// SYRK/GEMM (kind 0): F[a, b] -= L[a, k] L[b, k]^T;  GUPD (kind 1): X[i, c] -= L[i, k] X[k, c]
template <int NT>
__global__ void __launch_bounds__(NT)
k_fupdate(const int2* __restrict__ tasks, FactArgs A, int k, int* __restrict__ fail) {
    constexpr int NBJ = TileWaves<NT>::NBJ, R = TileWaves<NT>::R;
    __shared__ double As[kT * kT];   // As[kk * 64 + i]
    __shared__ double Bs[kT * kT];   // Bs[kk * 64 + j]
    const int2 task = tasks[blockIdx.x];
    const int s = task.x;
    const int kind = wrapd::task_kind(task.y);
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
    double va[R], vb[R];
    tile_load<NT>(F + static_cast<int64_t>(c0) * f + a0, f, na, nb, va);
    if (kind == 0) {
        tile_load<NT>(F + static_cast<int64_t>(c0) * f + b0, f, nbt, nb, vb);
        tile_to_lds<NT>(Bs, vb);
    } else {
        tile_load<NT>(Gs + static_cast<int64_t>(b0) * lg + c0, lg, nb, nbt, vb);
        tile_to_lds_t<NT>(Bs, vb);
    }
    tile_to_lds<NT>(As, va);
    __syncthreads();
    dbl4 acc[NBJ][2];
    tile_zero(acc);
    tile_mma<NT>(As, Bs, (nb + 3) & ~3, acc);
    if (kind == 0 && a == b && a == k + 1) {
        // The next panel's diagonal tile: its update stays in registers and the tile is
        // factored here (the POTRF of panel k + 1), overlapping the launch's other tasks.
        // As and Bs become that factorization's D and X.
        double* C = F + static_cast<int64_t>(b0) * f + a0;
        const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
        const int j0 = TileWaves<NT>::SJ * (w >> 1), i0 = 32 * (w & 1);
        double nv[NBJ * 8];
#pragma unroll
        for (int x = 0; x < NBJ; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = i0 + 16 * y + (l & 15), j = j0 + 16 * x + 4 * q + (l >> 4);
                    nv[(x * 2 + y) * 4 + q] =
                        C[static_cast<int64_t>(j < nbt ? j : 0) * f + (i < na ? i : 0)] - acc[x][y][q];
                }
        __syncthreads();   // every wave is done reading As and Bs
#pragma unroll
        for (int x = 0; x < NBJ; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = i0 + 16 * y + (l & 15), j = j0 + 16 * x + 4 * q + (l >> 4);
                    As[j * 64 + i] = i < j ? 0.0 : ((i < nbt && j < nbt) ? nv[(x * 2 + y) * 4 + q] : (i == j ? 1.0 : 0.0));
                }
#pragma unroll
        for (int r = 0; r < kT * kT / NT; ++r) Bs[threadIdx.x + NT * r] = 0.0;
        __syncthreads();
        potrf_lds(As, Bs, nbt, C, f, Gs + static_cast<int64_t>(b0) * lg + b0, lg, fail);
        return;
    }
    if (kind == 0) tile_write<true, NT>(F + static_cast<int64_t>(b0) * f + a0, f, na, nbt, a == b, acc, 1.0);
    else tile_write<true, NT>(Gs + static_cast<int64_t>(b0) * lg + a0, lg, na, nbt, false, acc, 1.0);
}

// ---------------------------------------------------------------------------------
// L21 at panel k (FK_L21): one workgroup per 64-row tile t of the update rows, in place,
//   F[t, k] -= L[t, p] L[k, p]^T   for p = 0, 1, .., k - 1, each product 64 deep from zero
//   L[t, k]  = F[t, k] L_kk^{-T}   (the product with the diagonal tile's stored inverse)
// When the panel steps cover the update rows, k_fupdate and k_ftrsm apply exactly these
// tile products (tile_mma) in this order, so L21 is the same to the last bit either way; here
// F[t, k] stays in registers between the products, so each update-row tile is written once
// instead of once per panel.  The next product's tiles are loaded while the current one runs
// on the matrix cores.
template <int NT>
__global__ void __launch_bounds__(NT)
k_fl21(const int2* __restrict__ tasks, FactArgs A, int k) {
    constexpr int NBJ = TileWaves<NT>::NBJ, R = TileWaves<NT>::R;
    __shared__ double As[kHalf * kT];   // As[(q - 32 h) * 64 + i], half h of the k range
    __shared__ double Bs[kHalf * kT];   // Bs[(q - 32 h) * 64 + j]
    const int2 task = tasks[blockIdx.x];
    const int s = task.x;
    const int f = A.fsize[s];
    const int ns = A.col0[s + 1] - A.col0[s];
    int a0, na;
    front_tile(task.y, ns, f, a0, na);
    double* F = A.pool + A.foff[s];
    const double* Gs = A.G + A.goff[s];
    const int lg = wrapd::g_ld(f);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int j0 = TileWaves<NT>::SJ * (w >> 1), i0 = 32 * (w & 1);
    const int c0 = kT * k;
    const int nbk = min(kT, ns - c0);
    // the operand tiles of product p: L[t, p] and L[k, p], from the front as a buffer
    const __amdgpu_buffer_rsrc_t rF = blk_rsrc(F, static_cast<int64_t>(f) * f);
    double va[R], vb[R];
    auto load_upd_h = [&](int p, int h) {
        tile_ldb_h<NT>(rF, kT * p * f + a0, f, na, kT, va, h);
        tile_ldb_h<NT>(rF, kT * p * f + c0, f, nbk, kT, vb, h);
    };
    if (k > 0) {
        load_upd_h(0, 0);
        load_upd_h(0, 1);
    }
    // run = F[t, k] in the accumulator layout (output (i, j) = (i0 + 16 b + l % 16,
    // j0 + 16 a + 4 q + l / 16)); positions outside the tile hold zero
    dbl4 run[NBJ][2];
#pragma unroll
    for (int x = 0; x < NBJ; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = i0 + 16 * y + (l & 15), j = j0 + 16 * x + 4 * q + (l >> 4);
                const double v = blk_ld(rF, (c0 + j) * f + a0 + i);
                run[x][y][q] = (i < na && j < nbk) ? v : 0.0;
            }
    for (int p = 0; p < k; ++p) {   // panels p < k < P are a full 64 columns
        dbl4 acc[NBJ][2];
        tile_zero(acc);
        for (int h = 0; h < 2; ++h) {
            if (p > 0 || h) __syncthreads();   // the previous half has been read
            tile_to_lds_h<NT>(As, va, h);
            tile_to_lds_h<NT>(Bs, vb, h);
            __syncthreads();
            if (p + 1 < k) load_upd_h(p + 1, h);   // this half is staged: load the next product's
            tile_mma<NT>(As, Bs, kHalf, acc);
        }
#pragma unroll
        for (int x = 0; x < NBJ; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int q = 0; q < 4; ++q) run[x][y][q] = run[x][y][q] - acc[x][y][q];
    }
    // L[t, k] = run L_kk^{-T}: run becomes the A operand (As[j * 64 + i] = run(i, j), j the
    // k index of this product), staged half by half
    tile_load<NT>(Gs + static_cast<int64_t>(c0) * lg + c0, lg, nbk, nbk, vb);
    dbl4 acc[NBJ][2];
    tile_zero(acc);
    for (int h = 0; h < 2; ++h) {
        const int nh = half_depth(nbk, h);
        if (nh == 0) break;
        if (k > 0 || h) __syncthreads();
#pragma unroll
        for (int x = 0; x < NBJ; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = i0 + 16 * y + (l & 15), j = j0 + 16 * x + 4 * q + (l >> 4);
                    if ((j >> 5) == h) As[(j - kHalf * h) * kT + i] = run[x][y][q];
                }
        tile_to_lds_h<NT>(Bs, vb, h);
        __syncthreads();
        tile_mma<NT>(As, Bs, nh, acc);
    }
    tile_write<false, NT>(F + static_cast<int64_t>(c0) * f + a0, f, na, nbk, false, acc, 1.0);
}

// ---------------------------------------------------------------------------------
// SCHUR (kind 0): F[a, b] -= sum_q L[a, q] L[b, q]^T over all ns columns, update tiles a >= b;
// GBOT (kind 1): G[a, c] = -sum_{q >= c} L[a, q] X[q, c], X = L11^{-1}.
// TILE x TILE output tiles over the whole k range: the NT / 64 waves split the tile into
// parts of 16x16 accumulators.  The k range is walked in KC-deep chunks, chunk q+1 fetched
// into registers while chunk q runs on the matrix cores; each output's products are
// accumulated in k order whatever the tiling, so the tile shape does not change the bits.
// Both operands are staged index-major (As[i * (KC + 1) + k], Bs[j * (KC + 1) + k]): the
// padding keeps the stores of either load orientation and the matrix-core reads free of bank
// conflicts.  At 64x64 with 4 waves a kernel needs 82 registers, so five waves share a SIMD
// and one's loads and barriers hide behind the others' products (128x128 tiles at two or four
// waves per SIMD were 2.5-3.4 ms slower per C5 refactorization; held to 80 registers, six waves
// per SIMD measured the same as five).
constexpr int kFK = wrapd::kFinalChunk;    // k-chunk

// R x C block (R * C = NT * NV, NT threads) of a column-major source (element offset off0,
// leading dimension ld) into registers: thread t's entry r is (row t % R, column t / R +
// (NT / R) r).
// The k dimension (rows for R = kFK, columns otherwise) is zero from nk on; entries past the
// other dimension's end hold whatever the block holds there (they only reach outputs that are
// not stored) or zero past the block's end.
template <int R, int KC, int NV, int NT>
__device__ __forceinline__ void blk_load(__amdgpu_buffer_rsrc_t rs, uint32_t off0, uint32_t ld, int nk, double v[NV]) {
    constexpr int CS = NT / R;
    const int row = threadIdx.x % R, col0 = threadIdx.x / R;
    const uint32_t o = off0 + row + col0 * ld;
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        const bool ok = R == KC ? row < nk : col0 + CS * r < nk;
        v[r] = blk_ld(rs, zero_unless(ok, o + static_cast<uint32_t>(CS * r) * ld));
    }
}

// the registers of blk_load<R> into an index-major LDS operand T[idx * (KC + 1) + k]: with
// R = TILE the block's rows are the index (entry r: idx t % TILE, k t / TILE + (NT / TILE) r),
// with R = KC its rows are k (entry r: k t % KC, idx t / KC + (NT / KC) r)
template <int R, int TILE, int KC, int NV, int NT>
__device__ __forceinline__ void blk_to_lds(double* T, const double v[NV]) {
    const int t = threadIdx.x;
    constexpr int CS = NT / R;
    constexpr int LD = KC + 1;
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        if (R == TILE) T[(t % TILE) * LD + t / TILE + CS * r] = v[r];
        else T[(t / KC + CS * r) * LD + t % KC] = v[r];
    }
}

template <int TILE, int KC, int NT>
__global__ void __launch_bounds__(NT, 2)
k_ffinal(const int2* __restrict__ tasks, FactArgs A) {
    constexpr int kLd = KC + 1;          // LDS operand row length (padded)
    constexpr int NV = TILE * KC / NT;   // operand entries per thread
    constexpr int WI = 2, WJ = NT / 128; // waves along the output rows i and columns j
    constexpr int SI = TILE / WI, SJ = TILE / WJ;   // a wave's part of the tile
    constexpr int NBI = SI / 16, NBJ = SJ / 16;     // its 16x16 blocks
    __shared__ double As[TILE * kLd];    // As[i * kLd + k]
    __shared__ double Bs[TILE * kLd];    // Bs[j * kLd + k]
    const int2 task = tasks[blockIdx.x];
    const int s = task.x;
    const int kind = wrapd::task_kind(task.y);
    const int u = (task.y >> wrapd::kTaskTileShift) & wrapd::kTaskTileMask;
    const int v = task.y & wrapd::kTaskTileMask;
    const int f = A.fsize[s];
    const int ns = A.col0[s + 1] - A.col0[s];
    const int a0 = ns + TILE * u, na = min(TILE, f - a0);
    const int b0 = kind == 0 ? ns + TILE * v : TILE * v;
    const int nbt = min(TILE, (kind == 0 ? f : ns) - b0);
    double* F = A.pool + A.foff[s];
    double* Gs = A.G + A.goff[s];
    const int lg = wrapd::g_ld(f);
    // the front (f x f) and G_s (lg x ns) as buffers
    const __amdgpu_buffer_rsrc_t rF = blk_rsrc(F, static_cast<int64_t>(f) * f);
    const __amdgpu_buffer_rsrc_t rG = blk_rsrc(Gs, static_cast<int64_t>(lg) * ns);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int li = l & 15, lk = l >> 4;
    const int i0 = SI * (w % WI), j0 = SJ * (w / WI);
    // a diagonal SCHUR super-tile keeps its lower half: a wave whose part lies wholly above
    // the diagonal is idle
    const bool idle = kind == 0 && u == v && i0 + SI <= j0;
    const int qfirst = kind == 0 ? 0 : b0;   // X[q, c] = 0 for q < c
    double ra[NV], rb[NV];
    auto fetch = [&](int q0) {
        const int nq = min(KC, ns - q0);
        blk_load<TILE, KC, NV, NT>(rF, q0 * f + a0, f, nq, ra);                  // L[a0 + i, q0 + k]
        if (kind == 0) blk_load<TILE, KC, NV, NT>(rF, q0 * f + b0, f, nq, rb);   // L[b0 + j, q0 + k]
        else blk_load<KC, KC, NV, NT>(rG, b0 * lg + q0, lg, nq, rb);             // X[q0 + k, b0 + j]
    };
    dbl4 acc[NBJ][NBI];
#pragma unroll
    for (int x = 0; x < NBJ; ++x)
#pragma unroll
        for (int y = 0; y < NBI; ++y) acc[x][y] = dbl4{0.0, 0.0, 0.0, 0.0};
    fetch(qfirst);
    for (int q0 = qfirst; q0 < ns; q0 += KC) {
        const int nq = min(KC, ns - q0);
        __syncthreads();   // the previous chunk has been consumed
        blk_to_lds<TILE, TILE, KC, NV, NT>(As, ra);
        if (kind == 0) blk_to_lds<TILE, TILE, KC, NV, NT>(Bs, rb);
        else blk_to_lds<KC, TILE, KC, NV, NT>(Bs, rb);
        __syncthreads();
        if (q0 + KC < ns) fetch(q0 + KC);
        if (!idle) {
            const int nk = (nq + 3) & ~3;
            for (int kk = 0; kk < nk; kk += 4) {
                const int k = kk + lk;
                double bj[NBJ], ai[NBI];
#pragma unroll
                for (int x = 0; x < NBJ; ++x) bj[x] = Bs[(j0 + 16 * x + li) * kLd + k];
#pragma unroll
                for (int y = 0; y < NBI; ++y) ai[y] = As[(i0 + 16 * y + li) * kLd + k];
#pragma unroll
                for (int x = 0; x < NBJ; ++x)
#pragma unroll
                    for (int y = 0; y < NBI; ++y)
                        acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(bj[x], ai[y], acc[x][y], 0, 0, 0);
            }
        }
    }
    if (idle) return;
    // acc[x][y][q] is output (i0 + 16 y + li, j0 + 16 x + 4 q + lk)
    double* C = kind == 0 ? F + static_cast<int64_t>(b0) * f + a0 : Gs + static_cast<int64_t>(b0) * lg + a0;
    const int64_t ldc = kind == 0 ? f : lg;
    const bool lower = kind == 0 && u == v;
#pragma unroll
    for (int x = 0; x < NBJ; ++x) {
        double old[NBI * 4];
        if (kind == 0) {
#pragma unroll
            for (int y = 0; y < NBI; ++y)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = i0 + 16 * y + li, j = j0 + 16 * x + 4 * q + lk;
                    old[y * 4 + q] = C[static_cast<int64_t>(j < nbt ? j : 0) * ldc + (i < na ? i : 0)];
                }
        }
#pragma unroll
        for (int y = 0; y < NBI; ++y)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = i0 + 16 * y + li, j = j0 + 16 * x + 4 * q + lk;
                if (i < na && j < nbt && !(lower && i < j))
                    C[static_cast<int64_t>(j) * ldc + i] = kind == 0 ? old[y * 4 + q] - acc[x][y][q] : -acc[x][y][q];
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
        wrapd::chol_plan(S, c->fplan, c->plan_opt);
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
    rc |= up_vec(c, c->d_cchunk_ptr, P.cchunk_ptr);
    rc |= up_vec(c, c->d_cchunk, P.cchunk);
    rc |= up_vec(c, c->d_ftasks, P.tasks);
    std::vector<int> zero1(1, 0);
    rc |= up_vec(c, c->d_ffail, zero1);
    if (!c->fstream2) {
        rc |= check_hip(c, hipStreamCreateWithFlags(&c->fstream2, hipStreamNonBlocking), "hipStreamCreate");
        rc |= check_hip(c, hipEventCreateWithFlags(&c->fev_fork, hipEventDisableTiming), "hipEventCreate");
        rc |= check_hip(c, hipEventCreateWithFlags(&c->fev_join, hipEventDisableTiming), "hipEventCreate");
    }
    size_t nfinal = 0;   // one event per FINAL launch of a level
    for (const wrapd::FactorLaunch& L : P.launches) nfinal = std::max(nfinal, static_cast<size_t>(L.ord + 1));
    while (c->fev_final.size() < nfinal) {
        hipEvent_t e = nullptr;
        rc |= check_hip(c, hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        if (!e) break;
        c->fev_final.push_back(e);
    }
    if (c->pool) { (void)hipFree(c->pool); c->pool = nullptr; }
    rc |= check_hip(c, hipMalloc(&c->pool, std::max<size_t>(8, P.pool_size * sizeof(double))), "hipMalloc(front pool)");
    return rc ? WRAPD_ERR_HIP : WRAPD_OK;
}

static void launch_final(const wrapd::FactorLaunch& L, const int2* t, const FactArgs& A, hipStream_t st) {
    hipLaunchKernelGGL((k_ffinal<wrapd::kFinalTile, kFK, kFactThreads>), dim3(L.count), dim3(kFactThreads), 0, st, t,
                       A);
}

// queues the whole factorization on c->stream (with the side launches on c->fstream2, forked
// after the main-stream launch before them; an assembly launch waits for the second stream's
// launches up to the previous level's FINAL launch its fronts need (FactorLaunch::wait), the
// level's last one for all of them, and so does the end); the caller reads c->d_ffail afterwards
void launch_factor(wrapd_ctx* c) {
    const wrapd::FactorPlan& P = c->fplan;
    FactArgs A{c->d_col0, c->d_fsize, c->d_foff, c->d_goff, c->pool, c->G};
    AsmArgs M{c->Ax, c->d_cbase, c->d_acol_ptr, reinterpret_cast<const int2*>(c->d_acol), c->d_ccol_ptr,
              reinterpret_cast<const int2*>(c->d_ccol), c->d_relptr, c->d_rel, c->d_cchunk_ptr, c->d_cchunk,
              P.asm_chunk};
    const int2* tasks = reinterpret_cast<const int2*>(c->d_ftasks);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->d_ffail, 1, 0);
    bool side_busy = false;
    auto join = [&] {
        if (!side_busy) return;
        (void)hipEventRecord(c->fev_join, c->fstream2);
        (void)hipStreamWaitEvent(c->stream, c->fev_join, 0);
        side_busy = false;
    };
    for (const wrapd::FactorLaunch& L : P.launches) {
        const int2* t = tasks + L.offset;
        if (L.kind == wrapd::FK_ASSEMBLE) {
            if (L.wait < 0 || L.wait >= static_cast<int>(c->fev_final.size()))
                join();
            else if (side_busy)
                (void)hipStreamWaitEvent(c->stream, c->fev_final[L.wait], 0);
        }
        if (L.side) {
            (void)hipEventRecord(c->fev_fork, c->stream);
            (void)hipStreamWaitEvent(c->fstream2, c->fev_fork, 0);
            if (L.kind == wrapd::FK_L21) {
                hipLaunchKernelGGL(k_fl21<kL21Threads>, dim3(L.count), dim3(kL21Threads), 0, c->fstream2, t, A, L.panel);
            } else {
                launch_final(L, t, A, c->fstream2);
                if (L.ord >= 0 && L.ord < static_cast<int>(c->fev_final.size()))
                    (void)hipEventRecord(c->fev_final[L.ord], c->fstream2);
            }
            side_busy = true;
            continue;
        }
        switch (L.kind) {
        case wrapd::FK_ASSEMBLE:
            hipLaunchKernelGGL(k_fassemble, dim3(L.count), dim3(64), static_cast<size_t>(L.rows) * sizeof(double),
                               c->stream, t, A, M);
            break;
        case wrapd::FK_POTRF:
            hipLaunchKernelGGL(k_fpotrf, dim3(L.count), dim3(kFactThreads), 0, c->stream, t, A, L.panel, c->d_ffail);
            break;
        case wrapd::FK_POTRF16:
            hipLaunchKernelGGL(k_fpotrf16, dim3(L.count), dim3(64), 0, c->stream, t, A, c->d_ffail);
            break;
        case wrapd::FK_TRSM:
            hipLaunchKernelGGL(k_ftrsm<kTrsmThreads>, dim3(L.count), dim3(kTrsmThreads), 0, c->stream, t, A, L.panel);
            break;
        case wrapd::FK_UPDATE:
            hipLaunchKernelGGL(k_fupdate<kUpdThreads>, dim3(L.count), dim3(kUpdThreads), 0, c->stream, t, A, L.panel, c->d_ffail);
            break;
        case wrapd::FK_L21:
            hipLaunchKernelGGL(k_fl21<kL21Threads>, dim3(L.count), dim3(kL21Threads), 0, c->stream, t, A, L.panel);
            break;
        default:
            launch_final(L, t, A, c->stream);
            break;
        }
    }
    join();
}

}  // namespace wd
