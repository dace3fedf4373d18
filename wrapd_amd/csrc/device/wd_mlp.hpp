// The learned prox warm start: an MLP 6 -> 64 -> 64 -> 3 with GELU (tanh form), BF16
// operands on the matrix cores, FP32 accumulation.
//
// Input  (6): sorted target singular values s0 >= s1 >= s2, log(w^2/vol / mu),
//             log(lambda / mu), log(mu), each standardised with the training set's
//             mean and standard deviation.
// Output (3): the prox solution minus s_hat, in units of the training targets'
//             standard deviation (plus their mean).
//
// FP32 master parameters, row-major (out x in):
//   W1[64][6] b1[64] W2[64][64] b2[64] W3[3][64] b3[3]         (4803 floats)
// Normalisation: in_mean[6] in_std[6] out_mean[3] out_std[3]   (18 floats)
//
// Matrix-core layout (v_mfma_f32_16x16x16bf16_1k, one wave, 16 samples per product):
// each layer is computed transposed, Z^T = W X^T, with A = a 16x16 weight tile
// (lane l: row l%16, k = 4(l/16) + 0..3) and B = 16 features x 16 samples (lane l:
// k = 4(l/16) + 0..3, sample l%16).  The accumulator holds Z^T[4(l/16) + i][l%16], which
// is exactly the B operand of the next layer's k-block: activations never leave the lane.
// The BF16 weight copies are stored tile-ready: W1p[64][16] (inputs zero-padded),
// W2[64][64], W3p[16][64] (outputs zero-padded), and for the backward pass the
// transposes W2t[64][64], W3t[64][16].
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace wd {

constexpr int kMlpIn = 6, kMlpHid = 64, kMlpOut = 3;
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpHid * kMlpIn;      // 384
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHid;               // 448
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHid * kMlpHid;     // 4544
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHid;               // 4608
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpOut * kMlpHid;     // 4800
constexpr int kMlpParams = kMlpOffB3 + kMlpOut;              // 4803
constexpr int kMlpNorm = 18;
// BF16 tile copies (uint16 elements)
constexpr int kMlpBfW1p = 0;                     // 64 x 16
constexpr int kMlpBfW2 = kMlpBfW1p + 64 * 16;     // 64 x 64
constexpr int kMlpBfW3p = kMlpBfW2 + 64 * 64;     // 16 x 64
constexpr int kMlpBfW2t = kMlpBfW3p + 16 * 64;    // 64 x 64
constexpr int kMlpBfW3t = kMlpBfW2t + 64 * 64;    // 64 x 16
constexpr int kMlpBfSize = kMlpBfW3t + 64 * 16;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__host__ __device__ inline uint16_t f2bf(float x) {
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16); }
__device__ inline float bfround(float x) { return bf2f(f2bf(x)); }

// GELU, tanh form: 0.5 x (1 + tanh(c (x + a x^3))), and its derivative
__device__ inline float gelu(float x) {
    const float c = 0.7978845608028654f, a = 0.044715f;
    return 0.5f * x * (1.0f + tanhf(c * (x + a * x * x * x)));
}
__device__ inline float gelu_d(float x) {
    const float c = 0.7978845608028654f, a = 0.044715f;
    const float t = tanhf(c * (x + a * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * c * (1.0f + 3.0f * a * x * x);
}

__device__ inline f32x4 mfma_bf16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A operand of a weight tile: rows [16 tr, +16), k-block t, row-major matrix of width ld
// (a multiple of 4, so a lane's 4 values are one aligned 8-byte load)
__device__ inline bf16x4 ld_tile_a(const uint16_t* __restrict__ W, int ld, int tr, int t, int lane) {
    return *reinterpret_cast<const bf16x4*>(W + (16 * tr + (lane & 15)) * ld + 16 * t + 4 * (lane >> 4));
}

__device__ inline f32x4 ld_bias(const float* __restrict__ b, int rsub) {
    f32x4 v;
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = b[rsub + i];
    return v;
}

// B operand from an accumulator tile (4 rows x this lane's sample), rounded to BF16
__device__ inline bf16x4 acc_to_b(f32x4 v) {
    bf16x4 r;
    r[0] = static_cast<short>(f2bf(v[0]));
    r[1] = static_cast<short>(f2bf(v[1]));
    r[2] = static_cast<short>(f2bf(v[2]));
    r[3] = static_cast<short>(f2bf(v[3]));
    return r;
}

struct MlpDevice {
    const uint16_t* bf;    // kMlpBfSize BF16 tiles
    const float* master;   // kMlpParams (biases are read from here)
    const float* norm;     // kMlpNorm
};

// raw prox inputs -> standardised features (FP32)
__device__ inline void mlp_features(const float* __restrict__ norm, float s0, float s1, float s2, float wov, float mu,
                                    float lam, float f[kMlpIn]) {
    const float raw[kMlpIn] = {s0, s1, s2, logf(wov / mu), logf(lam / mu), logf(mu)};
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) f[i] = (raw[i] - norm[i]) / norm[6 + i];
}

// Forward pass of one 16-sample block inside a wave.  xb: this lane's B operand of the
// standardised input block (features 4(l/16) + 0..3 of sample l%16, zero beyond 6).
// Returns the output accumulator (lane l < 16 holds outputs 0..2 of sample l in [0..2]).
// If z1/z2 are given, the pre-activations are kept (training).
__device__ inline f32x4 mlp_block_forward(const MlpDevice& M, bf16x4 xb, int lane, f32x4* z1 = nullptr,
                                          f32x4* z2 = nullptr, bf16x4* h1o = nullptr, bf16x4* h2o = nullptr) {
    const int rsub = 4 * (lane >> 4);
    bf16x4 h1[4], h2[4];
#pragma unroll
    for (int to = 0; to < 4; ++to) {
        f32x4 acc;
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = M.master[kMlpOffB1 + 16 * to + rsub + i];
        acc = mfma_bf16(ld_tile_a(M.bf + kMlpBfW1p, 16, to, 0, lane), xb, acc);
        if (z1) z1[to] = acc;
        f32x4 g;
#pragma unroll
        for (int i = 0; i < 4; ++i) g[i] = gelu(acc[i]);
        h1[to] = acc_to_b(g);
        if (h1o) h1o[to] = h1[to];
    }
#pragma unroll
    for (int to = 0; to < 4; ++to) {
        f32x4 acc;
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = M.master[kMlpOffB2 + 16 * to + rsub + i];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc = mfma_bf16(ld_tile_a(M.bf + kMlpBfW2, 64, to, t, lane), h1[t], acc);
        if (z2) z2[to] = acc;
        f32x4 g;
#pragma unroll
        for (int i = 0; i < 4; ++i) g[i] = gelu(acc[i]);
        h2[to] = acc_to_b(g);
        if (h2o) h2o[to] = h2[to];
    }
    f32x4 y;
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] = (rsub + i < kMlpOut) ? M.master[kMlpOffB3 + rsub + i] : 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) y = mfma_bf16(ld_tile_a(M.bf + kMlpBfW3p, 64, 0, t, lane), h2[t], y);
    return y;
}

// B operand of block b (samples 16b .. 16b+15 of the wave) from per-lane features
__device__ inline bf16x4 mlp_gather_input(const float f[kMlpIn], int b, int lane) {
    const int src = 16 * b + (lane & 15);
    const int k0 = 4 * (lane >> 4);
    float v[kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) v[i] = __shfl(f[i], src, 64);
    bf16x4 r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k = k0 + j;
        float x = 0.0f;
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) x = (k == i) ? v[i] : x;
        r[j] = static_cast<short>(f2bf(x));
    }
    return r;
}

// Whole-wave inference: every lane passes its own features and receives its own three
// outputs (de-standardised: prox solution minus s_hat).  All 64 lanes must call it.
// The wave's 64 samples are 4 blocks of 16; each weight tile is loaded once and applied
// to all 4 (the same products, in the same order, as mlp_block_forward per block).
// Blocks with no lane in `want` (a wave-uniform lane mask) are skipped; their outputs are 0.
__device__ inline void mlp_wave_predict(const MlpDevice& M, const float f[kMlpIn], int lane, float out[kMlpOut],
                                        unsigned long long want = ~0ull) {
    const int rsub = 4 * (lane >> 4);
    bf16x4 xb[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) xb[b] = mlp_gather_input(f, b, lane);
    bf16x4 h1[4][4], h2[4][4];   // [block][k-block]
#pragma unroll
    for (int to = 0; to < 4; ++to) {
        const bf16x4 w = ld_tile_a(M.bf + kMlpBfW1p, 16, to, 0, lane);
        const f32x4 bias = ld_bias(M.master + kMlpOffB1 + 16 * to, rsub);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (!((want >> (16 * b)) & 0xFFFFull)) continue;
            const f32x4 acc = mfma_bf16(w, xb[b], bias);
            f32x4 g;
#pragma unroll
            for (int i = 0; i < 4; ++i) g[i] = gelu(acc[i]);
            h1[b][to] = acc_to_b(g);
        }
    }
#pragma unroll
    for (int to = 0; to < 4; ++to) {
        bf16x4 w[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) w[t] = ld_tile_a(M.bf + kMlpBfW2, 64, to, t, lane);
        const f32x4 bias = ld_bias(M.master + kMlpOffB2 + 16 * to, rsub);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (!((want >> (16 * b)) & 0xFFFFull)) continue;
            f32x4 acc = bias;
#pragma unroll
            for (int t = 0; t < 4; ++t) acc = mfma_bf16(w[t], h1[b][t], acc);
            f32x4 g;
#pragma unroll
            for (int i = 0; i < 4; ++i) g[i] = gelu(acc[i]);
            h2[b][to] = acc_to_b(g);
        }
    }
    bf16x4 w3[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) w3[t] = ld_tile_a(M.bf + kMlpBfW3p, 64, 0, t, lane);
    f32x4 b3;
#pragma unroll
    for (int i = 0; i < 4; ++i) b3[i] = (rsub + i < kMlpOut) ? M.master[kMlpOffB3 + rsub + i] : 0.0f;
    out[0] = out[1] = out[2] = 0.0f;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        if (!((want >> (16 * b)) & 0xFFFFull)) continue;
        f32x4 y = b3;
#pragma unroll
        for (int t = 0; t < 4; ++t) y = mfma_bf16(w3[t], h2[b][t], y);
        // lane m < 16 holds sample 16b + m in y[0..2]
        const int src = lane & 15;
        const float y0 = __shfl(y[0], src, 64), y1 = __shfl(y[1], src, 64), y2 = __shfl(y[2], src, 64);
        if ((lane >> 4) == b) {
            out[0] = y0 * M.norm[15] + M.norm[12];
            out[1] = y1 * M.norm[16] + M.norm[13];
            out[2] = y2 * M.norm[17] + M.norm[14];
        }
    }
}

}  // namespace wd
