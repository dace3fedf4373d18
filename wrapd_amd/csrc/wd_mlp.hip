// Learned prox warm start: training and batch inference of the MLP of device/wd_mlp.hpp.
//
// Kernels:
//   k_mlp_pack      FP32 master parameters -> BF16 matrix-core tiles (and transposes)
//   k_mlp_infer     batch inference, one sample per lane, 16-sample products per wave
//   k_mlp_train_fb  forward + backward of a minibatch: each wave takes 64 samples per
//                   round (4 blocks of 16), keeps the layer gradients in registers and
//                   accumulates the weight gradients on the matrix cores from LDS
//                   copies of the activations; writes one partial gradient per wave
//   k_mlp_adam      sums the partial gradients in wave order (deterministic) and takes
//                   an Adam step on the FP32 master parameters
//   k_mlp_augment   the tuples' w^2/vol redrawn log-uniformly over a range and their
//                   targets re-solved by the FP64 Newton prox (device/wd_nh.hpp)
// Training tuples are 9 floats: s0 s1 s2 w^2/vol mu lambda | prox solution x0 x1 x2.
#include <algorithm>
#include <cmath>
#include <vector>

#include "wd_ctx.hpp"
#include "device/wd_mlp.hpp"
#include "device/wd_nh.hpp"

namespace wd {

__global__ void k_mlp_pack(const float* __restrict__ P, uint16_t* __restrict__ bf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpBfSize) return;
    float v = 0.0f;
    if (i < kMlpBfW2) {                       // W1p 64 x 16
        const int r = (i - kMlpBfW1p) / 16, k = (i - kMlpBfW1p) % 16;
        v = k < kMlpIn ? P[kMlpOffW1 + r * kMlpIn + k] : 0.0f;
    } else if (i < kMlpBfW3p) {               // W2 64 x 64
        v = P[kMlpOffW2 + (i - kMlpBfW2)];
    } else if (i < kMlpBfW2t) {               // W3p 16 x 64
        const int r = (i - kMlpBfW3p) / 64, k = (i - kMlpBfW3p) % 64;
        v = r < kMlpOut ? P[kMlpOffW3 + r * 64 + k] : 0.0f;
    } else if (i < kMlpBfW3t) {               // W2t 64 x 64: W2t[r][k] = W2[k][r]
        const int r = (i - kMlpBfW2t) / 64, k = (i - kMlpBfW2t) % 64;
        v = P[kMlpOffW2 + k * 64 + r];
    } else {                                  // W3t 64 x 16: W3t[r][k] = W3[k][r]
        const int r = (i - kMlpBfW3t) / 16, k = (i - kMlpBfW3t) % 16;
        v = k < kMlpOut ? P[kMlpOffW3 + k * 64 + r] : 0.0f;
    }
    bf[i] = f2bf(v);
}

// in: n x 6 raw prox inputs (s0 s1 s2 w^2/vol mu lambda); out: n x 3 predicted solutions
__global__ void __launch_bounds__(256)
k_mlp_infer(int n, const float* __restrict__ in, float* __restrict__ out, MlpDevice M) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    float f[kMlpIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float s[3] = {0.f, 0.f, 0.f};
    if (i < n) {
        const float* p = in + 6ll * i;
        s[0] = p[0]; s[1] = p[1]; s[2] = p[2];
        mlp_features(M.norm, p[0], p[1], p[2], p[3], p[4], p[5], f);
    }
    float y[kMlpOut];
    mlp_wave_predict(M, f, lane, y);
    if (i < n) {
        out[3ll * i + 0] = s[0] + y[0];
        out[3ll * i + 1] = s[1] + y[1];
        out[3ll * i + 2] = s[2] + y[2];
    }
}

__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// sample source: either the list idx (n entries, used in order) or, with idx == nullptr,
// indices drawn from (seed, position) over n_data tuples
struct TrainBatch {
    const float* tuples;
    long long n_data;
    const int* idx;
    long long n;          // samples in the minibatch
    uint64_t seed;
};

__device__ inline long long sample_index(const TrainBatch& B, long long pos) {
    if (B.idx) return B.idx[pos];
    return static_cast<long long>(mix64(B.seed ^ (0x632BE59BD9B4E019ull * static_cast<uint64_t>(pos + 1))) %
                                  static_cast<uint64_t>(B.n_data));
}

// one wave per block; rounds of 64 samples; partial[wave][kMlpParams], lpart[wave]
__global__ void __launch_bounds__(64)
k_mlp_train_fb(MlpDevice M, TrainBatch B, int rounds, float* __restrict__ partial, float* __restrict__ lpart) {
    __shared__ uint16_t s_h1[64 * 16], s_h2[64 * 16], s_dz1[64 * 16], s_dz2[64 * 16], s_x[16 * 16], s_dy[16 * 16];
    const int lane = threadIdx.x;
    const int rsub = 4 * (lane >> 4), col = lane & 15;
    const float inv_n = 1.0f / static_cast<float>(B.n);
    f32x4 gW1[4], gW2[4][4], gW3[4], gb1[4], gb2[4], gb3;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gW1[a] = gW3[a] = gb1[a] = gb2[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < 4; ++b) gW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    gb3 = f32x4{0.f, 0.f, 0.f, 0.f};
    float loss = 0.0f;
    for (int rnd = 0; rnd < rounds; ++rnd) {
        const long long pos = (static_cast<long long>(blockIdx.x) * rounds + rnd) * 64 + lane;
        const bool valid = pos < B.n;
        float f[kMlpIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, tn[3] = {0.f, 0.f, 0.f};
        if (valid) {
            const float* t = B.tuples + 9ll * sample_index(B, pos);
            mlp_features(M.norm, t[0], t[1], t[2], t[3], t[4], t[5], f);
#pragma unroll
            for (int k = 0; k < 3; ++k) tn[k] = ((t[6 + k] - t[k]) - M.norm[12 + k]) / M.norm[15 + k];
        }
        const float w = valid ? 1.0f : 0.0f;
#pragma unroll 1
        for (int b = 0; b < 4; ++b) {
            const bf16x4 xb = mlp_gather_input(f, b, lane);
            f32x4 z1[4], z2[4];
            bf16x4 h1[4], h2[4];
            const f32x4 y = mlp_block_forward(M, xb, lane, z1, z2, h1, h2);
            // targets and weights of this block's samples, in the output layout
            const int src = 16 * b + col;
            const float t0 = __shfl(tn[0], src, 64), t1 = __shfl(tn[1], src, 64), t2 = __shfl(tn[2], src, 64);
            const float wb = __shfl(w, src, 64);
            f32x4 dy = {0.f, 0.f, 0.f, 0.f};
            if (rsub == 0) {
                const float e0 = y[0] - t0, e1 = y[1] - t1, e2 = y[2] - t2;
                loss += wb * (e0 * e0 + e1 * e1 + e2 * e2);
                dy[0] = 2.0f * wb * e0 * inv_n;
                dy[1] = 2.0f * wb * e1 * inv_n;
                dy[2] = 2.0f * wb * e2 * inv_n;
            }
            const bf16x4 dyb = acc_to_b(dy);
            // dH2 = W3^T dY, dZ2 = dH2 * gelu'(Z2)
            f32x4 dz2[4];
#pragma unroll
            for (int to = 0; to < 4; ++to) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma_bf16(ld_tile_a(M.bf + kMlpBfW3t, 16, to, 0, lane), dyb, acc);
#pragma unroll
                for (int i = 0; i < 4; ++i) dz2[to][i] = acc[i] * gelu_d(z2[to][i]);
            }
            // dH1 = W2^T dZ2, dZ1 = dH1 * gelu'(Z1)
            bf16x4 dz2b[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) dz2b[t] = acc_to_b(dz2[t]);
            f32x4 dz1[4];
#pragma unroll
            for (int to = 0; to < 4; ++to) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int t = 0; t < 4; ++t) acc = mfma_bf16(ld_tile_a(M.bf + kMlpBfW2t, 64, to, t, lane), dz2b[t], acc);
#pragma unroll
                for (int i = 0; i < 4; ++i) dz1[to][i] = acc[i] * gelu_d(z1[to][i]);
            }
            // bias gradients (per lane, summed over lanes at the end)
#pragma unroll
            for (int to = 0; to < 4; ++to) {
                gb1[to] += dz1[to];
                gb2[to] += dz2[to];
            }
            gb3 += dy;
            // feature-major LDS copies [feature][sample] of the operands of the weight
            // gradients (rows rsub + i of each tile, sample col)
            __syncthreads();
#pragma unroll
            for (int to = 0; to < 4; ++to)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = 16 * to + rsub + i;
                    s_h1[r * 16 + col] = static_cast<uint16_t>(h1[to][i]);
                    s_h2[r * 16 + col] = static_cast<uint16_t>(h2[to][i]);
                    s_dz1[r * 16 + col] = f2bf(dz1[to][i]);
                    s_dz2[r * 16 + col] = static_cast<uint16_t>(dz2b[to][i]);
                }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                s_x[(rsub + i) * 16 + col] = static_cast<uint16_t>(xb[i]);
                s_dy[(rsub + i) * 16 + col] = static_cast<uint16_t>(dyb[i]);
            }
            __syncthreads();
            // A operand: rows of a [feature][sample] array (lane: row l%16, samples 4(l/16)..)
            auto lda = [&](const uint16_t* S, int tr) {
                const uint16_t* p = S + (16 * tr + col) * 16 + rsub;
                bf16x4 v;
                v[0] = static_cast<short>(p[0]); v[1] = static_cast<short>(p[1]);
                v[2] = static_cast<short>(p[2]); v[3] = static_cast<short>(p[3]);
                return v;
            };
            // B operand: samples x features (lane: samples 4(l/16).., feature column l%16)
            auto ldb = [&](const uint16_t* S, int tc) { return lda(S, tc); };
#pragma unroll
            for (int ti = 0; ti < 4; ++ti) {
                const bf16x4 bh1 = ldb(s_h1, ti), bh2 = ldb(s_h2, ti);
#pragma unroll
                for (int to = 0; to < 4; ++to) gW2[to][ti] = mfma_bf16(lda(s_dz2, to), bh1, gW2[to][ti]);
                gW3[ti] = mfma_bf16(lda(s_dy, 0), bh2, gW3[ti]);
            }
            const bf16x4 bx = ldb(s_x, 0);
#pragma unroll
            for (int to = 0; to < 4; ++to) gW1[to] = mfma_bf16(lda(s_dz1, to), bx, gW1[to]);
        }
    }
    // bias gradients: sum over the 16 lanes (samples) that share rows
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int to = 0; to < 4; ++to) {
                gb1[to][i] += __shfl_xor(gb1[to][i], m, 64);
                gb2[to][i] += __shfl_xor(gb2[to][i], m, 64);
            }
            gb3[i] += __shfl_xor(gb3[i], m, 64);
        }
    }
    float* g = partial + static_cast<long long>(blockIdx.x) * kMlpParams;
#pragma unroll
    for (int to = 0; to < 4; ++to)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 16 * to + rsub + i;
            if (col < kMlpIn) g[kMlpOffW1 + r * kMlpIn + col] = gW1[to][i];
#pragma unroll
            for (int ti = 0; ti < 4; ++ti) g[kMlpOffW2 + r * 64 + 16 * ti + col] = gW2[to][ti][i];
            if (col == 0) {
                g[kMlpOffB1 + r] = gb1[to][i];
                g[kMlpOffB2 + r] = gb2[to][i];
            }
        }
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (rsub + i < kMlpOut) g[kMlpOffW3 + (rsub + i) * 64 + 16 * ti + col] = gW3[ti][i];
    if (col == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (rsub + i < kMlpOut) g[kMlpOffB3 + rsub + i] = gb3[i];
    }
    for (int m = 1; m < 64; m <<= 1) loss += __shfl_xor(loss, m, 64);
    if (lane == 0) lpart[blockIdx.x] = loss * inv_n;
}

// Adam (Kingma & Ba 2015) on the summed partial gradients; lr <= 0 only sums
__global__ void __launch_bounds__(256)
k_mlp_adam(const float* __restrict__ partial, const float* __restrict__ lpart, int nparts, float* __restrict__ P,
           float* __restrict__ m1, float* __restrict__ m2, float* __restrict__ gsum, float* __restrict__ loss_out,
           float lr, float beta1, float beta2, float eps, int t) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        float l = 0.0f;
        for (int w = 0; w < nparts; ++w) l += lpart[w];
        *loss_out = l;
    }
    if (p >= kMlpParams) return;
    float g = 0.0f;
    for (int w = 0; w < nparts; ++w) g += partial[static_cast<long long>(w) * kMlpParams + p];
    gsum[p] = g;
    if (lr <= 0.0f) return;
    const float a = beta1 * m1[p] + (1.0f - beta1) * g;
    const float b = beta2 * m2[p] + (1.0f - beta2) * g * g;
    m1[p] = a;
    m2[p] = b;
    const float ah = a / (1.0f - powf(beta1, static_cast<float>(t)));
    const float bh = b / (1.0f - powf(beta2, static_cast<float>(t)));
    P[p] -= lr * ah / (sqrtf(bh) + eps);
}

// w^2/vol redrawn log-uniformly in [lo, hi]; targets re-solved by the FP64 Newton prox
__global__ void __launch_bounds__(256)
k_mlp_augment(long long n, float* __restrict__ tuples, double lo, double hi, uint64_t seed) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* t = tuples + 9 * i;
    const double u = static_cast<double>(mix64(seed ^ (0xD1B54A32D192ED03ull * static_cast<uint64_t>(i + 1))) >> 11) *
                     (1.0 / 9007199254740992.0);
    const double wov = exp(log(lo) + u * (log(hi) - log(lo)));
    ProxProblem<double> P;
    P.p.mu = t[4];
    P.p.lambda = t[5];
    P.x0[0] = t[0]; P.x0[1] = t[1]; P.x0[2] = t[2];
    P.wov = wov;
    double x[3] = {t[6], t[7], t[8]};   // the logged solution is a good start
    double gn = 0.0;
    const int it = prox_newton(P, x, 60, 1e-12, gn);
    if (it < 0 || !(gn <= 1e-9 * (P.p.mu + wov) * 4.0) || !(x[0] > 0.0 && x[1] > 0.0 && x[2] > 0.0)) return;
    t[3] = static_cast<float>(wov);
    t[6] = static_cast<float>(x[0]);
    t[7] = static_cast<float>(x[1]);
    t[8] = static_cast<float>(x[2]);
}

// ---------------------------------------------------------------------------------
static int mlp_alloc(wrapd_ctx* c) {
    if (c->mlp_master) return 0;
    int rc = 0;
    rc |= check_hip(c, hipMalloc(&c->mlp_master, kMlpParams * sizeof(float)), "hipMalloc(mlp)");
    rc |= check_hip(c, hipMalloc(&c->mlp_m1, kMlpParams * sizeof(float)), "hipMalloc(mlp)");
    rc |= check_hip(c, hipMalloc(&c->mlp_m2, kMlpParams * sizeof(float)), "hipMalloc(mlp)");
    rc |= check_hip(c, hipMalloc(&c->mlp_gsum, kMlpParams * sizeof(float)), "hipMalloc(mlp)");
    rc |= check_hip(c, hipMalloc(&c->mlp_norm, kMlpNorm * sizeof(float)), "hipMalloc(mlp)");
    rc |= check_hip(c, hipMalloc(&c->mlp_bf, kMlpBfSize * sizeof(uint16_t)), "hipMalloc(mlp)");
    rc |= check_hip(c, hipMalloc(&c->mlp_loss, 16 * sizeof(float)), "hipMalloc(mlp)");
    return rc;
}

MlpDevice mlp_device(wrapd_ctx* c) { return MlpDevice{c->mlp_bf, c->mlp_master, c->mlp_norm}; }

int mlp_upload(wrapd_ctx* c, const float* params, const float* norm) {
    if (mlp_alloc(c)) return WRAPD_ERR_HIP;
    if (check_hip(c, hipMemcpyAsync(c->mlp_master, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, c->stream),
                  "mlp params") ||
        check_hip(c, hipMemcpyAsync(c->mlp_norm, norm, kMlpNorm * sizeof(float), hipMemcpyHostToDevice, c->stream),
                  "mlp norm"))
        return WRAPD_ERR_HIP;
    hipLaunchKernelGGL(k_mlp_pack, dim3(blocks_for(kMlpBfSize, 256)), dim3(256), 0, c->stream, c->mlp_master, c->mlp_bf);
    c->mlp_ready = true;
    return WRAPD_OK;
}

int mlp_infer(wrapd_ctx* c, const float* d_in, int n, float* d_out) {
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, d_in, d_out, mlp_device(c));
    return WRAPD_OK;
}

// the loss and its gradient summed over n given tuples (no Adam step), into caller-given
// buffers: partial (waves x kMlpParams), lpart (waves), gsum (kMlpParams), loss (1), with
// waves = mlp_grad_waves(n).  The same kernels and order as mlp_step, so the same bits.
void launch_mlp_grad(wrapd_ctx* c, const float* d_tuples, const int* d_idx, long long n, float* partial, float* lpart,
                     float* gsum, float* loss) {
    const int rounds = 4;
    const int waves = mlp_grad_waves(n);
    TrainBatch B{d_tuples, n, d_idx, n, 0};
    hipLaunchKernelGGL(k_mlp_train_fb, dim3(waves), dim3(64), 0, c->stream, mlp_device(c), B, rounds, partial, lpart);
    hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(kMlpParams, 256)), dim3(256), 0, c->stream, partial, lpart, waves,
                       c->mlp_master, c->mlp_m1, c->mlp_m2, gsum, loss, 0.0f, 0.9f, 0.999f, 1e-8f, 1);
}

// one minibatch forward/backward (+ Adam when lr > 0); waves of `rounds` x 64 samples
int mlp_step(wrapd_ctx* c, const float* d_tuples, long long n_data, const int* d_idx, long long n, uint64_t seed,
             float lr, int t) {
    const int rounds = 4;
    const int waves = static_cast<int>((n + 64ll * rounds - 1) / (64ll * rounds));
    if (c->mlp_partial_waves < waves) {
        if (c->mlp_partial) (void)hipFree(c->mlp_partial);
        if (c->mlp_lpart) (void)hipFree(c->mlp_lpart);
        c->mlp_partial = nullptr;
        c->mlp_lpart = nullptr;
        if (check_hip(c, hipMalloc(&c->mlp_partial, static_cast<size_t>(waves) * kMlpParams * sizeof(float)), "mlp partial") ||
            check_hip(c, hipMalloc(&c->mlp_lpart, static_cast<size_t>(waves) * sizeof(float)), "mlp partial"))
            return WRAPD_ERR_HIP;
        c->mlp_partial_waves = waves;
    }
    TrainBatch B{d_tuples, n_data, d_idx, n, seed};
    hipLaunchKernelGGL(k_mlp_train_fb, dim3(waves), dim3(64), 0, c->stream, mlp_device(c), B, rounds, c->mlp_partial,
                       c->mlp_lpart);
    hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(kMlpParams, 256)), dim3(256), 0, c->stream, c->mlp_partial,
                       c->mlp_lpart, waves, c->mlp_master, c->mlp_m1, c->mlp_m2, c->mlp_gsum, c->mlp_loss, lr, 0.9f,
                       0.999f, 1e-8f, t);
    if (lr > 0.0f)
        hipLaunchKernelGGL(k_mlp_pack, dim3(blocks_for(kMlpBfSize, 256)), dim3(256), 0, c->stream, c->mlp_master,
                           c->mlp_bf);
    return WRAPD_OK;
}

__global__ void __launch_bounds__(256)
k_mlp_gather(long long m, long long stride, const float* __restrict__ src, float* __restrict__ dst) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m * 9) return;
    const long long row = i / 9, k = i - row * 9;
    dst[i] = src[row * stride * 9 + k];
}

// every stride-th tuple of the log, contiguous on the host (m rows)
int mlp_sample(wrapd_ctx* c, long long stride, long long m, float* host) {
    float* d = nullptr;
    if (check_hip(c, hipMalloc(&d, std::max<long long>(1, m) * 9 * sizeof(float)), "mlp sample")) return WRAPD_ERR_HIP;
    hipLaunchKernelGGL(k_mlp_gather, dim3(blocks_for(m * 9, 256)), dim3(256), 0, c->stream, m, stride, c->mlp_log, d);
    int st = WRAPD_OK;
    if (check_hip(c, hipMemcpyAsync(host, d, m * 9 * sizeof(float), hipMemcpyDeviceToHost, c->stream), "mlp sample") ||
        check_hip(c, hipStreamSynchronize(c->stream), "mlp sample"))
        st = WRAPD_ERR_HIP;
    (void)hipFree(d);
    return st;
}

void launch_mlp_augment(wrapd_ctx* c, float* d_tuples, long long n, double lo, double hi, uint64_t seed) {
    hipLaunchKernelGGL(k_mlp_augment, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, d_tuples, lo, hi, seed);
}

}  // namespace wd
