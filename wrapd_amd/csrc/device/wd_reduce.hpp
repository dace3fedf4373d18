// Deterministic block reductions and the XCD-aware block mapping.
#pragma once

#include <hip/hip_runtime.h>

namespace wd {

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs.  Remap so that
// each XCD works on one contiguous range of the logical grid (neighbouring elements
// share vertices, so their gathers then hit the same L2).
__device__ __forceinline__ int xcd_block(int bid, int nblk) {
    const int full = (nblk / 8) * 8;
    if (bid >= full) return bid;
    const int per = nblk / 8;
    return (bid % 8) * per + bid / 8;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Sum over the block, fixed order; the result is valid in thread 0.  sh: NT/64 doubles.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) r += sh[i];
    }
    __syncthreads();
    return r;
}

// Block sums of n <= KS values per thread, each in the order of block_sum, with one
// barrier for all of them; the result of value k is returned in thread k (k < n).
// sh: KS x NT/64 doubles.
template <int NT, int KS>
__device__ __forceinline__ double block_sums(const double (&v)[KS], int n, double (*sh)[NT / 64]) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < KS; ++k)
        if (k < n) {
            const double t = wave_sum(v[k]);
            if (l == 0) sh[k][w] = t;
        }
    __syncthreads();
    double r = 0.0;
    if (static_cast<int>(threadIdx.x) < n) {
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) r += sh[threadIdx.x][i];
    }
    return r;
}

// The same with some values reduced by max (bit k of maxmask), in the order of block_max.
template <int NT, int KS>
__device__ __forceinline__ double block_sums_max(const double (&v)[KS], unsigned maxmask, double (*sh)[NT / 64]) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        const double t = ((maxmask >> k) & 1u) ? wave_max(v[k]) : wave_sum(v[k]);
        if (l == 0) sh[k][w] = t;
    }
    __syncthreads();
    double r = 0.0;
    const int k = threadIdx.x;
    if (k < KS) {
        if ((maxmask >> k) & 1u) {
            r = sh[k][0];
#pragma unroll
            for (int i = 1; i < NT / 64; ++i) r = fmax(r, sh[k][i]);
        } else {
#pragma unroll
            for (int i = 0; i < NT / 64; ++i) r += sh[k][i];
        }
    }
    return r;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* sh) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 1; i < NT / 64; ++i) r = fmax(r, sh[i]);
    }
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------------
// Values handed from one workgroup to another in the same launch are stored and loaded at
// agent scope (written through to memory, read past this XCD's L2).  Where the reader
// polls the values themselves, they hold kPendingBits until written: a signalling NaN with
// a payload, which arithmetic never yields (it only yields quiet NaNs).  Each 8-byte word
// is written once per use, so a word that is not pending is final.
constexpr unsigned long long kPendingBits = 0xFFF40000FFF40000ull;
__device__ __forceinline__ bool pending(double v) {
    return static_cast<unsigned long long>(__double_as_longlong(v)) == kPendingBits;
}
__device__ __forceinline__ double pending_value() { return __longlong_as_double(static_cast<long long>(kPendingBits)); }
__device__ __forceinline__ double ld_agent(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------------
// Reductions finished in the producing launch.  Every block stores its partials (slot k
// of block lb at partials[k * kTailBlocks + lb]: slot-major, so the reading block's loads
// are contiguous) at agent scope, exactly the R.nslots slots the reduction reads.  The
// block dispatched last (blockIdx.x == gridDim.x - 1: every other block is already
// running, and none waits for anything) polls the partials until none is pending, sets
// them back to pending for the next reduction, and sums them in the order of a 256-thread
// finalize: lane v takes blocks v, v + 256, ...; lanes are summed in waves of 64 and the 4
// waves in order.  So the result is the one a separate one-block finalize launch would
// produce, without the launch (and without a done-flag per block, which would cost its
// writer a store completion and the reader a round trip).
constexpr int kTailSlots = 32;        // reduction slots per block
constexpr int kTailBlocks = 1 << 16;  // blocks per reducing launch (the partials' row length)
constexpr int kTailSpinLimit = 1 << 20;

struct RedTail {
    bool armed = false;          // finished in this launch; otherwise (with nslots) by a finalize launch
    int* abortf = nullptr;       // sticky: a wait ran out (the results are then NaN)
    double* out = nullptr;       // nslots results
    double* hout = nullptr;      // their copy in host memory (pinned), or nullptr when the host does not read them
    unsigned* hseq = nullptr;    // host memory: set to epoch once the copies (and any work after) are stored, or nullptr
    unsigned epoch = 0;
    int nslots = 0;              // the slots every block stores (0: the launch reduces nothing)
    int nstore = 0;              // the leading slots stored to out (and hout); the rest only to shres
    unsigned maxmask = 0;        // slots reduced with max instead of sum
};

__device__ __forceinline__ void put_partial(double* partials, int lb, int k, double v) {
    st_agent(partials + static_cast<long long>(k) * kTailBlocks + lb, v);
}

// Tells the host that the host copies of this launch are stored: thread 0 of the block that
// finished the reduction, after every store the host will read (each thread waited for its
// own before the block's last barrier).
__device__ __forceinline__ void tail_signal(const RedTail R) {
    if (threadIdx.x != 0 || !R.hseq) return;
    __builtin_amdgcn_s_waitcnt(0);
    __hip_atomic_store(R.hseq, R.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Returns true in the block that finished the reduction (the results are in R.out, and
// in shres when given: LDS of at least KS doubles).  KS: the most slots this call site
// reduces (R.nslots <= KS).  SIGNAL: call tail_signal here (false: the caller does, after
// work of its own that the host reads).
template <int NT, int KS, bool SIGNAL = true>
__device__ __forceinline__ bool reduce_tail(const RedTail R, double* partials, double* shres = nullptr) {
    static_assert(NT == 128 || NT == 256, "reduce_tail: 128 or 256 threads");
    static_assert(KS >= 1 && KS <= kTailSlots, "reduce_tail: slot count");
    if (!R.armed) return false;
    const int nblk = gridDim.x;
    if (blockIdx.x != nblk - 1) return false;
    __shared__ int sbad;
    if (threadIdx.x == 0) sbad = 0;
    // The partials, at agent scope (past this XCD's L2), BB blocks' slots per batch.
    constexpr int BB = KS > 8 ? 2 : 4;
    constexpr int V = 256 / NT;   // lanes of the 256-lane order per thread
    __shared__ double wsum[KS][4];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int ns = R.nslots;
    const double pv = pending_value();
    bool bad = false;
    __syncthreads();
#pragma unroll
    for (int h = 0; h < V; ++h) {
        double acc[KS];
#pragma unroll
        for (int k = 0; k < KS; ++k) acc[k] = ((R.maxmask >> k) & 1u) ? -1e308 : 0.0;
        for (int b0 = threadIdx.x + h * NT; b0 < nblk; b0 += 256 * BB) {
            double v[BB][KS];
            auto load = [&](bool only_pending) {
#pragma unroll
                for (int u = 0; u < BB; ++u) {
                    const int b = b0 + u * 256;
#pragma unroll
                    for (int k = 0; k < KS; ++k)
                        if (b < nblk && k < ns && (!only_pending || pending(v[u][k])))
                            v[u][k] = ld_agent(partials + static_cast<long long>(k) * kTailBlocks + b);
                }
            };
#pragma unroll
            for (int u = 0; u < BB; ++u)
#pragma unroll
                for (int k = 0; k < KS; ++k) v[u][k] = 0.0;
            load(false);
            for (int polls = 1;; ++polls) {   // until no partial of this batch is pending
                bool any = false;
#pragma unroll
                for (int u = 0; u < BB; ++u)
#pragma unroll
                    for (int k = 0; k < KS; ++k) any = any || pending(v[u][k]);
                if (!any || bad) break;
                if ((polls & 63) == 0 &&
                    (polls >= kTailSpinLimit ||
                     __hip_atomic_load(R.abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
                    __hip_atomic_store(R.abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    bad = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
                load(true);
            }
#pragma unroll
            for (int u = 0; u < BB; ++u) {
                const int b = b0 + u * 256;
#pragma unroll
                for (int k = 0; k < KS; ++k)
                    if (b < nblk && k < ns) {   // pending again for the next reduction
                        st_agent(partials + static_cast<long long>(k) * kTailBlocks + b, pv);
                        acc[k] = ((R.maxmask >> k) & 1u) ? fmax(acc[k], v[u][k]) : acc[k] + v[u][k];
                    }
            }
        }
#pragma unroll
        for (int k = 0; k < KS; ++k)
            if (k < ns) {   // uniform across the block
                const double t = ((R.maxmask >> k) & 1u) ? wave_max(acc[k]) : wave_sum(acc[k]);
                if (l == 0) wsum[k][w + h * (NT / 64)] = t;
            }
    }
    if (bad) sbad = 1;
    __syncthreads();
    bad = sbad != 0;
    if (static_cast<int>(threadIdx.x) < ns) {
        const int k = threadIdx.x;
        double r;
        if ((R.maxmask >> k) & 1u) {
            r = wsum[k][0];
            for (int i = 1; i < 4; ++i) r = fmax(r, wsum[k][i]);
        } else {
            r = 0.0;
            for (int i = 0; i < 4; ++i) r += wsum[k][i];
        }
        if (bad) r = __builtin_nan("");
        if (shres) shres[k] = r;
        if (k < R.nstore) R.out[k] = r;
        if (R.hout && k < R.nstore) {
            __hip_atomic_store(R.hout + k, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __builtin_amdgcn_s_waitcnt(0);
        }
    }
    __syncthreads();
    if (SIGNAL) tail_signal(R);
    return true;
}

}  // namespace wd
