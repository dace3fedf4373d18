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

}  // namespace wd
