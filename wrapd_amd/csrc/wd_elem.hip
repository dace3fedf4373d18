// Element kernels: one tetrahedron per lane over structure-of-arrays element state.
//
//   k_setup       TetElements ctor (src/TetElements.cpp:90-142)
//   k_init_duals  initialize_dual_vars + polar_potential_energy (src/TetElements.cpp:145-183, 638-675)
//   k_defo        update_defo_cache fused with global_obj_value/grad, potential_energy_x,
//                 inversion_count, candidate weights and add_to_b
//                 (src/TetElements.cpp:564-596, 678-787, 598-635, 444-468, 391-432, 470-481)
//   k_local       polar_update (src/TetElements.cpp:256-356) with the candidate weights fused
//   k_local_ua    update (src/TetElements.cpp:193-254), rotation-unaware local step
//   k_update_w    update_actual_weights (src/TetElements.cpp:435-442)
//   k_assemble_A  update_A_coeffs (src/TetElements.cpp:498-510) as a per-entry gather
#include <type_traits>

#include <hipcub/hipcub.hpp>

#include "wd_ctx.hpp"
#include "device/wd_math.hpp"
#include "device/wd_nh.hpp"
#include "device/wd_mlp.hpp"
#include "device/wd_reduce.hpp"

namespace wd {

__device__ __forceinline__ double ldD(const double* Di, int n_ep, int r, int k, int e) {
    return Di[(r * 4 + k) * n_ep + e];
}

struct MatArgs {
    const double* mu_e;   // nullptr: homogeneous
    const double* lam_e;
    double mu, lam;
    double beta_min, beta_max;
};

__device__ __forceinline__ NHParams<double> mat_of(const MatArgs& m, int e) {
    NHParams<double> p;
    p.mu = m.mu_e ? m.mu_e[e] : m.mu;
    p.lambda = m.lam_e ? m.lam_e[e] : m.lam;
    return p;
}

// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(kElemThreads)
k_setup(int n_e, int n_ep, const int4* __restrict__ tv, const double* __restrict__ rest,
        double* __restrict__ Di, double* __restrict__ vol, double* __restrict__ kst,
        double* __restrict__ wsq, double* __restrict__ sig, MatArgs m, int model, double beta_static,
        int* __restrict__ bad) {
    const int e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (e >= n_ep) return;
    if (e >= n_e) {   // padding lanes: harmless values
#pragma unroll
        for (int i = 0; i < 12; ++i) Di[i * n_ep + e] = 0.0;
        vol[e] = 1.0; kst[e] = 1.0; wsq[e] = 1.0;
        sig[e] = sig[n_ep + e] = sig[2 * n_ep + e] = 1.0;
        return;
    }
    const int4 t = tv[e];
    const int id[4] = {t.x, t.y, t.z, t.w};
    M3<double> E;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) E(r, c) = rest[3 * id[c + 1] + r] - rest[3 * id[0] + r];
    M3<double> C;
    C(0, 0) = E(1, 1) * E(2, 2) - E(1, 2) * E(2, 1);
    C(0, 1) = E(0, 2) * E(2, 1) - E(0, 1) * E(2, 2);
    C(0, 2) = E(0, 1) * E(1, 2) - E(0, 2) * E(1, 1);
    C(1, 0) = E(1, 2) * E(2, 0) - E(1, 0) * E(2, 2);
    C(1, 1) = E(0, 0) * E(2, 2) - E(0, 2) * E(2, 0);
    C(1, 2) = E(0, 2) * E(1, 0) - E(0, 0) * E(1, 2);
    C(2, 0) = E(1, 0) * E(2, 1) - E(1, 1) * E(2, 0);
    C(2, 1) = E(0, 1) * E(2, 0) - E(0, 0) * E(2, 1);
    C(2, 2) = E(0, 0) * E(1, 1) - E(0, 1) * E(1, 0);
    const double det = E(0, 0) * C(0, 0) + E(0, 1) * C(1, 0) + E(0, 2) * C(2, 0);
    const double id_ = 1.0 / det;
    M3<double> Ei;
#pragma unroll
    for (int i = 0; i < 9; ++i) Ei.m[i] = C.m[i] * id_;
    const double v = det3(E) / 6.0;
    if (v < 0.0) bad[0] = 1;
    vol[e] = v;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        Di[(r * 4 + 0) * n_ep + e] = -Ei(0, r) - Ei(1, r) - Ei(2, r);
#pragma unroll
        for (int i = 0; i < 3; ++i) Di[(r * 4 + i + 1) * n_ep + e] = Ei(i, r);
    }
    const NHParams<double> p = mat_of(m, e);
    const double k = model == 1 ? 2.0066 * p.mu + 1.0122 * p.lambda : 1.0;
    kst[e] = k;
    wsq[e] = beta_static * k * v;
    sig[e] = sig[n_ep + e] = sig[2 * n_ep + e] = 1.0;
}

// F from positions; `split` keeps free and fixed corner sums apart as the cached
// deformation of the reference does (free part + D_fix x_fix, src/TetElements.cpp:575-585).
// Ft(r,c) = sum_k D(r,k) x_k[c]; F = Ft^T.  Fx receives the fixed part of Ft.
template <bool SPLIT>
__device__ __forceinline__ void gather_F(const int4 vi, const double* __restrict__ Di, int n_ep, int e, int n_f,
                                         const double* __restrict__ xall, M3<double>& F, double d[3][4],
                                         M3<double>* Fx_t) {
    const int id[4] = {vi.x, vi.y, vi.z, vi.w};
    double x[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) x[k][c] = xall[3 * id[k] + c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 4; ++k) d[r][k] = ldD(Di, n_ep, r, k, e);
    M3<double> Ff, Fx;
#pragma unroll
    for (int i = 0; i < 9; ++i) { Ff.m[i] = 0.0; Fx.m[i] = 0.0; }
    bool anyfix = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool fr = !SPLIT || id[k] < n_f;
        anyfix |= !fr;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (fr) Ff(r, c) += d[r][k] * x[k][c];
                else Fx(r, c) += d[r][k] * x[k][c];
            }
    }
    if (anyfix) {
#pragma unroll
        for (int i = 0; i < 9; ++i) Ff.m[i] += Fx.m[i];
    }
    if (Fx_t) *Fx_t = Fx;
    F = transpose(Ff);
}

__device__ __forceinline__ void ld9(const double* a, int n_ep, int e, M3<double>& M) {
#pragma unroll
    for (int i = 0; i < 9; ++i) M.m[i] = a[i * n_ep + e];
}
__device__ __forceinline__ void st9(double* a, int n_ep, int e, const M3<double>& M) {
#pragma unroll
    for (int i = 0; i < 9; ++i) a[i * n_ep + e] = M.m[i];
}

// ---------------------------------------------------------------------------------
template <int RA>
__global__ void __launch_bounds__(kElemThreads)
k_init_duals(int n_e, int n_ep, int n_f, const int4* __restrict__ vidx, const double* __restrict__ Di,
             const double* __restrict__ xall, const double* __restrict__ vol, double* __restrict__ U,
             double* __restrict__ P, MatArgs m, int model, double* __restrict__ partials, RedTail tail) {
    __shared__ double sh[kElemThreads / 64];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int e = lb * blockDim.x + threadIdx.x;
    double pe = 0.0;
    if (e < n_e) {
        M3<double> F;
        double d[3][4];
        gather_F<false>(reinterpret_cast<const int4*>(vidx)[e], Di, n_ep, e, n_f, xall, F, d, nullptr);
        M3<double> Z;
        if (RA) {
            double s[3];
            M3<double> Uu, V;
            svd3(F, s, Uu, V);
            Z = transpose(udvt(V, s, V));
            if (model == 1) {
                double s2[3];
                M3<double> U2, V2;
                svd3(transpose(Z), s2, U2, V2);
                pe = vol[e] * nh_psi(s2, mat_of(m, e));
            } else {
                M3<double> D = Z;
                D(0, 0) -= 1.0; D(1, 1) -= 1.0; D(2, 2) -= 1.0;
                pe = 0.5 * 1.0 * vol[e] * sqnorm9(D);
            }
        } else {
            Z = transpose(F);
        }
        M3<double> zero;
#pragma unroll
        for (int i = 0; i < 9; ++i) zero.m[i] = 0.0;
        st9(U, n_ep, e, zero);
        st9(P, n_ep, e, Z);
    }
    const double t = block_sum<kElemThreads>(pe, sh);
    if (threadIdx.x == 0 && tail.nslots > 0) put_partial(partials, lb, 0, t);
    reduce_tail<kElemThreads, 1>(tail, partials);
}

// ---------------------------------------------------------------------------------
template <int MODE, int MODEL>
__global__ void __launch_bounds__(kElemThreads)
k_defo(int n_e, int n_ep, int n_f, const int4* __restrict__ vidx, const double* __restrict__ Di,
       const double* __restrict__ xall, const double* __restrict__ P, const double* __restrict__ wsq,
       const double* __restrict__ vol, const double* __restrict__ kst, double* __restrict__ symF,
       double* __restrict__ gstage, const int4* __restrict__ cpos, double* __restrict__ cand, MatArgs m,
       int use_hessian, double* __restrict__ partials, RedTail tail) {
    __shared__ double shk[3][kElemThreads / 64];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int e = lb * blockDim.x + threadIdx.x;
    double val = 0.0, en = 0.0, flip = 0.0;
    if (e < n_e) {
        M3<double> F, Fx;
        double d[3][4];
        const int4 vi = vidx[e];
        // where the corner gradients go: each free corner's row of the vertex-major staging
        const int4 cp = (MODE & (DM_GRAD | DM_BVEC)) ? cpos[e] : make_int4(-1, -1, -1, -1);
        const int cq[4] = {cp.x, cp.y, cp.z, cp.w};
        if (MODE & DM_DIRECT) gather_F<false>(vi, Di, n_ep, e, n_f, xall, F, d, &Fx);
        else gather_F<true>(vi, Di, n_ep, e, n_f, xall, F, d, &Fx);
        if (MODE & DM_FT_CACHE) st9(symF, n_ep, e, transpose(F));
        double s[3];
        M3<double> Uu, V;
        const bool need_svd = (MODE & (DM_WRITE_SYMF | DM_VALUE | DM_GRAD | DM_ENERGY | DM_CAND)) &&
                              !((MODE & DM_FT_CACHE) && !(MODE & (DM_ENERGY | DM_CAND)));
        if (need_svd) svd3(F, s, Uu, V);
        M3<double> S;
        if (MODE & (DM_WRITE_SYMF | DM_VALUE | DM_GRAD)) S = udvt(V, s, V);
        if ((MODE & DM_WRITE_SYMF) && !(MODE & DM_FT_CACHE)) st9(symF, n_ep, e, S);
        const double w2 = (MODE & (DM_VALUE | DM_GRAD | DM_BVEC)) ? wsq[e] : 0.0;
        M3<double> Pm;
        if (MODE & (DM_VALUE | DM_GRAD | DM_BVEC)) ld9(P, n_ep, e, Pm);
        if (MODE & (DM_VALUE | DM_GRAD)) {
            // P here is (Z - U)^T in the reference's notation: P_ref = (Z - U)^T
            const M3<double> Pt = transpose(Pm);
            if (MODE & DM_VALUE) {
                M3<double> D;
#pragma unroll
                for (int i = 0; i < 9; ++i) D.m[i] = S.m[i] - Pt.m[i];
                val = (0.5 * w2) * sqnorm9(D);
            }
            if (MODE & DM_GRAD) {
                const M3<double> Q = matmul_tn(V, matmul(Pt, V));
                M3<double> K;
                K(0, 0) = Q(0, 0) / s[0];
                K(1, 1) = Q(1, 1) / s[1];
                K(2, 2) = Q(2, 2) / s[2];
                K(0, 1) = K(1, 0) = (Q(0, 1) + Q(1, 0)) / (s[0] + s[1]);
                K(0, 2) = K(2, 0) = (Q(0, 2) + Q(2, 0)) / (s[0] + s[2]);
                K(1, 2) = K(2, 1) = (Q(1, 2) + Q(2, 1)) / (s[1] + s[2]);
                M3<double> US = Uu;
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) US(r, c) *= s[c];
                const M3<double> T = matmul(matmul(US, K), transpose(V));
                M3<double> Mm;
#pragma unroll
                for (int i = 0; i < 9; ++i) Mm.m[i] = F.m[i] - T.m[i];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (cq[k] >= 0) {
                        double* g = gstage + 3ll * cq[k];
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            g[c] = d[0][k] * (w2 * Mm(c, 0)) + d[1][k] * (w2 * Mm(c, 1)) + d[2][k] * (w2 * Mm(c, 2));
                    }
            }
        }
        if (MODE & DM_BVEC) {
            // add_to_b: w^2 D^T (Z - U - D_fix x_fix), the cache here holds Ft so P = Z - U
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (cq[k] >= 0) {
                    double* g = gstage + 3ll * cq[k];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        double acc = 0.0;
#pragma unroll
                        for (int r = 0; r < 3; ++r) acc += (w2 * d[r][k]) * (Pm(r, c) - Fx(r, c));
                        g[c] = acc;
                    }
                }
        }
        if (MODE & DM_ENERGY) {
            if (MODEL == 1) {
                en = vol[e] * nh_psi(s, mat_of(m, e));
            } else {
                const M3<double> R = matmul(Uu, transpose(V));
                M3<double> D;
#pragma unroll
                for (int i = 0; i < 9; ++i) D.m[i] = F.m[i] - R.m[i];
                en = 0.5 * 1.0 * vol[e] * sqnorm9(D);
            }
            flip = (s[0] < 0.0 || s[1] < 0.0 || s[2] < 0.0) ? 1.0 : 0.0;
        }
        if (MODE & DM_CAND) {
            const NHParams<double> p = mat_of(m, e);
            const double k = kst[e];
            double cs = k;
            if (use_hessian) cs = nh_candidate_stiffness(s, p, k * m.beta_min, k * m.beta_max);
            cand[e] = vol[e] * cs;
        }
    }
    if (MODE & (DM_VALUE | DM_ENERGY)) {   // slots value, energy, flips: the reduction's nslots of them
        const double vals[3] = {val, en, flip};
        const double t = block_sums<kElemThreads, 3>(vals, 3, shk);
        if (static_cast<int>(threadIdx.x) < tail.nslots) put_partial(partials, lb, threadIdx.x, t);
        reduce_tail<kElemThreads, 3>(tail, partials);
    }
}

// ---------------------------------------------------------------------------------
// Local step.  T is the precision of the SVD and the prox; the ADMM state stays FP64.
template <typename T>
__device__ __forceinline__ M3<T> cvt(const M3<double>& a) {
    M3<T> b;
#pragma unroll
    for (int i = 0; i < 9; ++i) b.m[i] = static_cast<T>(a.m[i]);
    return b;
}

// one training tuple (s_hat, w^2/vol, mu, lambda | prox solution) for the learned warm start
__device__ __forceinline__ void log_tuple(float* __restrict__ tlog, int e, const double shat[3], double wov,
                                          const NHParams<double>& pd, const double x[3]) {
    float* t = tlog + 9ll * e;
    t[0] = static_cast<float>(shat[0]);
    t[1] = static_cast<float>(shat[1]);
    t[2] = static_cast<float>(shat[2]);
    t[3] = static_cast<float>(wov);
    t[4] = static_cast<float>(pd.mu);
    t[5] = static_cast<float>(pd.lambda);
    t[6] = static_cast<float>(x[0]);
    t[7] = static_cast<float>(x[1]);
    t[8] = static_cast<float>(x[2]);
}

// Stretch prox of one element in precision T.  Returns the new Z (FP64) and the
// prox solution; iters gets the solver iterations.  ok = false when the result is unusable.
template <typename T, int PROX>
__device__ __forceinline__ void elem_prox(const M3<double>& Sin, const double sl[3], NHParams<double> pd, double wov,
                                          int prox_iters, double rtol, M3<double>& Znew, double x_out[3],
                                          double& iters, bool& ok, double& shat_min, double& evals_out,
                                          double sh_out[3]) {
    const M3<T> St = cvt<T>(Sin);
    T sh[3];
    M3<T> Ut, Vt;
    svd3(St, sh, Ut, Vt);
    shat_min = static_cast<double>(sh[2]);
    sh_out[0] = static_cast<double>(sh[0]);
    sh_out[1] = static_cast<double>(sh[1]);
    sh_out[2] = static_cast<double>(sh[2]);
    ProxProblem<T> P;
    P.p.mu = static_cast<T>(pd.mu);
    P.p.lambda = static_cast<T>(pd.lambda);
    P.x0[0] = sh[0]; P.x0[1] = sh[1]; P.x0[2] = sh[2];
    P.wov = static_cast<T>(wov);
    T x[3] = {static_cast<T>(sl[0]), static_cast<T>(sl[1]), static_cast<T>(sl[2])};
    int it;
    if (PROX == WRAPD_PROX_LBFGS) {
        int evals = 0;
        it = prox_lbfgs(P, x, prox_iters, evals);
        evals_out += static_cast<double>(evals);
    } else {
        T gn = T(0);
        it = prox_newton(P, x, prox_iters, static_cast<T>(rtol), gn);
        evals_out += static_cast<double>(it < 0 ? 0 : it);
    }
    iters = it < 0 ? 0.0 : static_cast<double>(it);
    const M3<T> Zt = udvt(Vt, x, Ut);
    ok = true;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        Znew.m[i] = static_cast<double>(Zt.m[i]);
        ok &= isfinite(Znew.m[i]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) x_out[i] = static_cast<double>(x[i]);
}

// The local step's 11 reduction slots: 0 bef, 1 aft, 2 pe, 3 max ratio, 4 prox iterations,
// 5 FP64 fallbacks, 6 inverted, 7 learned starts accepted, 8 evaluations, 9 max
// evaluations, 10 learned starts used.
// The local step's 11 block partials (slots 3 and 9 by max), reduced with one barrier and
// stored by threads 0..10.
constexpr unsigned kLocalMaxMask = (1u << 3) | (1u << 9);
__device__ __forceinline__ void store_local_partials(double* partials, int lb, double r0, double r1, double r2,
                                                     double r3, double r4, double r5, double r6, double r7,
                                                     double r8, double r9, double r10,
                                                     double (*shk)[kLocalThreads / 64], int nslots) {
    const double v[11] = {r0, r1, r2, r3, r4, r5, r6, r7, r8, r9, r10};
    const double t = block_sums_max<kLocalThreads, 11>(v, kLocalMaxMask, shk);
    if (static_cast<int>(threadIdx.x) < min(nslots, 11)) put_partial(partials, lb, threadIdx.x, t);
}

// The FP32 Newton local steps fit 168 registers (a few spilled to scratch), which gives three
// waves per SIMD instead of two; the FP64 and L-BFGS variants need more and keep one or two.
constexpr int local_min_blocks(int prec, int prox) { return prec == WRAPD_FP32 && prox != WRAPD_PROX_LBFGS ? 3 : 1; }

template <int MODEL, int PREC, int PROX>
__global__ void __launch_bounds__(kLocalThreads, local_min_blocks(PREC, PROX))
k_local(int n_e, int n_ep, const double* __restrict__ symF, double* __restrict__ U, double* __restrict__ P,
        const double* __restrict__ wsq, const double* __restrict__ vol, const double* __restrict__ kst,
        double* __restrict__ cand, double* __restrict__ sig, MatArgs m, int prox_iters, double rtol,
        double fallback_sigma, int update_cand, int use_hessian, double* __restrict__ partials,
        float* __restrict__ tlog, float* __restrict__ fbinfo, RedTail tail) {
    __shared__ double shk[11][kLocalThreads / 64];   // the block reductions' wave results
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int e = lb * blockDim.x + threadIdx.x;
    double bef = 0.0, aft = 0.0, pe = 0.0, ratio = 0.0, iters = 0.0, fallback = 0.0, neg = 0.0, evals = 0.0;
    bool defer = false;   // an FP32 result unusable here: redone in FP64 by k_local_fallback
    if (e < n_e) {
        M3<double> S, Uo, Po;
        ld9(symF, n_ep, e, S);
        ld9(U, n_ep, e, Uo);
        ld9(P, n_ep, e, Po);
        const double w2 = wsq[e], v = vol[e];
        {   // bef = 1/2 w^2 |symF - (Z - U)|^2 with P = Z - U stored transposed-free
            M3<double> D;
#pragma unroll
            for (int i = 0; i < 9; ++i) D.m[i] = S.m[i] - Po.m[i];
            bef = 0.5 * w2 * sqnorm9(D);
        }
        M3<double> Zn;
        double x[3] = {1.0, 1.0, 1.0};
        if (MODEL == 1) {
            M3<double> Sin;
#pragma unroll
            for (int i = 0; i < 9; ++i) Sin.m[i] = S.m[i] + Uo.m[i];
            Sin = transpose(Sin);
            const NHParams<double> pd = mat_of(m, e);
            const double sl[3] = {sig[e], sig[n_ep + e], sig[2 * n_ep + e]};
            bool ok;
            double smin, shat[3];
            elem_prox<typename std::conditional<PREC == WRAPD_FP32, float, double>::type, PROX>(
                Sin, sl, pd, w2 / v, prox_iters, rtol, Zn, x, iters, ok, smin, evals, shat);
            if (PREC == WRAPD_FP32)
                defer = !ok || smin < fallback_sigma || x[0] < 0.0 || x[1] < 0.0 || x[2] < 0.0;
            if (!defer) {
                if (tlog) log_tuple(tlog, e, shat, w2 / v, pd, x);
                pe = v * nh_psi(x, pd);
                neg = (x[0] < 0.0 || x[1] < 0.0 || x[2] < 0.0) ? 1.0 : 0.0;
                sig[e] = x[0];
                sig[n_ep + e] = x[1];
                sig[2 * n_ep + e] = x[2];
            }
        } else {
#pragma unroll
            for (int i = 0; i < 9; ++i) Zn.m[i] = S.m[i] + Uo.m[i];
            Zn.m[0] += 1.0; Zn.m[4] += 1.0; Zn.m[8] += 1.0;
#pragma unroll
            for (int i = 0; i < 9; ++i) Zn.m[i] *= 0.5;
            M3<double> D = S;
            D.m[0] -= 1.0; D.m[4] -= 1.0; D.m[8] -= 1.0;
            pe = 0.5 * 1.0 * v * sqnorm9(D);
        }
        if (PREC == WRAPD_FP32) fbinfo[e] = defer ? static_cast<float>(evals + 1.0) : 0.0f;
        if (defer) {   // U, P, sig and the candidate weight stay for the FP64 pass; its sums are that pass's
            fallback = 1.0;
            iters = 0.0;
        } else {
            M3<double> Un, Pn;
            {
                M3<double> D;
#pragma unroll
                for (int i = 0; i < 9; ++i) D.m[i] = S.m[i] - (Zn.m[i] - Uo.m[i]);
                aft = 0.5 * w2 * sqnorm9(D);
            }
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                Un.m[i] = Uo.m[i] + (S.m[i] - Zn.m[i]);
                Pn.m[i] = Zn.m[i] - Un.m[i];
            }
            st9(U, n_ep, e, Un);
            st9(P, n_ep, e, Pn);
            if (update_cand) {
                const double k = kst[e];
                double cs = k;
                if (use_hessian) cs = nh_candidate_stiffness(x, mat_of(m, e), k * m.beta_min, k * m.beta_max);
                const double cw = v * cs;
                cand[e] = cw;
                const double a = w2 / cw, b = 1.0 / a;
                ratio = a > b ? a : b;
            }
        }
    }
    // slot 9 (max evaluations per element) of a deferred element is the FP64 pass's
    store_local_partials(partials, lb, bef, aft, pe, ratio, iters, fallback, neg, 0.0, evals, defer ? 0.0 : evals, 0.0,
                         shk, tail.nslots);
    reduce_tail<kLocalThreads, 11>(tail, partials);
}

// The FP64 redo of the elements an FP32 local step deferred (fb_info nonzero), listed in
// element order in `list` (count elements): the k_local FP64 prox on each, from the
// unchanged U, P, sig, with the same outputs.  Its sums go to their own slots: aft, pe,
// max ratio, FP64 iterations, inverted, FP64 evaluations (slot 8) and each element's
// evaluations of both attempts (slot 9, by max); the FP32 kernel counted the fallbacks and
// summed bef and the FP32 evaluations.  A fixed grid walks the list (i = block * 256 +
// thread, + grid * 256, ...), so every thread's sums are taken in a fixed order.
template <int PROX>
__global__ void __launch_bounds__(kLocalThreads)
k_local_fallback(int n_ep, const int* __restrict__ list, const int* __restrict__ count,
                 const double* __restrict__ symF, double* __restrict__ U, double* __restrict__ P,
                 const double* __restrict__ wsq, const double* __restrict__ vol, const double* __restrict__ kst,
                 double* __restrict__ cand, double* __restrict__ sig, MatArgs m, int prox_iters, double rtol,
                 int update_cand, int use_hessian, const float* __restrict__ fbinfo, double* __restrict__ partials,
                 float* __restrict__ tlog, RedTail tail) {
    __shared__ double shk[11][kLocalThreads / 64];
    const int n = *count;
    double aft = 0.0, pe = 0.0, ratio = 0.0, iters = 0.0, neg = 0.0, evals = 0.0, emax = 0.0;
    for (int i = blockIdx.x * kLocalThreads + threadIdx.x; i < n; i += gridDim.x * kLocalThreads) {
        const int e = list[i];
        M3<double> S, Uo;
        ld9(symF, n_ep, e, S);
        ld9(U, n_ep, e, Uo);
        const double w2 = wsq[e], v = vol[e];
        M3<double> Sin;
#pragma unroll
        for (int k = 0; k < 9; ++k) Sin.m[k] = S.m[k] + Uo.m[k];
        Sin = transpose(Sin);
        const NHParams<double> pd = mat_of(m, e);
        const double sl[3] = {sig[e], sig[n_ep + e], sig[2 * n_ep + e]};
        M3<double> Zn;
        double x[3] = {1.0, 1.0, 1.0}, it = 0.0, smin, ev = 0.0, shat[3];
        bool ok;
        elem_prox<double, PROX>(Sin, sl, pd, w2 / v, prox_iters, rtol, Zn, x, it, ok, smin, ev, shat);
        if (tlog) log_tuple(tlog, e, shat, w2 / v, pd, x);
        pe += v * nh_psi(x, pd);
        neg += (x[0] < 0.0 || x[1] < 0.0 || x[2] < 0.0) ? 1.0 : 0.0;
        sig[e] = x[0];
        sig[n_ep + e] = x[1];
        sig[2 * n_ep + e] = x[2];
        M3<double> D, Un, Pn;
#pragma unroll
        for (int k = 0; k < 9; ++k) D.m[k] = S.m[k] - (Zn.m[k] - Uo.m[k]);
        aft += 0.5 * w2 * sqnorm9(D);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            Un.m[k] = Uo.m[k] + (S.m[k] - Zn.m[k]);
            Pn.m[k] = Zn.m[k] - Un.m[k];
        }
        st9(U, n_ep, e, Un);
        st9(P, n_ep, e, Pn);
        if (update_cand) {
            const double k = kst[e];
            double cs = k;
            if (use_hessian) cs = nh_candidate_stiffness(x, pd, k * m.beta_min, k * m.beta_max);
            const double cw = v * cs;
            cand[e] = cw;
            const double a = w2 / cw, b = 1.0 / a;
            ratio = fmax(ratio, a > b ? a : b);
        }
        iters += it;
        evals += ev;
        emax = fmax(emax, static_cast<double>(fbinfo[e]) - 1.0 + ev);
    }
    store_local_partials(partials, blockIdx.x, 0.0, aft, pe, ratio, iters, 0.0, neg, 0.0, evals, emax, 0.0, shk,
                         tail.nslots);
    reduce_tail<kLocalThreads, 11>(tail, partials);
}

// Learned warm start (WRAPD_PROX_LEARNED): after the SVD of S_in, a wave in which some
// element's previous solution is a poor start (relative optimality residual above
// the context's learned_gate) evaluates the MLP
// (device/wd_mlp.hpp) on its 64 elements; each element's Newton prox then starts from
// whichever of the prediction and the previous solution has the smaller residual |grad|,
// and a prediction already within the Newton tolerance is taken as is.  Everything else
// (reductions, FP64 fallback, candidate weights) is k_local's.
// |grad| of the prox objective at x (infinite outside x > 0 or when not finite); f and g
// are kept for the Newton prox that may start there.
template <typename T>
__device__ __forceinline__ T prox_residual(const ProxProblem<T>& P, const T x[3], T& f, T g[3]) {
    g[0] = g[1] = g[2] = T(0);
    f = T(0);
    if (!(x[0] > T(0) && x[1] > T(0) && x[2] > T(0))) return T(INFINITY);
    f = P.gradient(x, g);
    const T r = wsqrt(dot3(g, g));
    return r == r ? r : T(INFINITY);
}

template <int PREC>
__global__ void __launch_bounds__(kLocalThreads, local_min_blocks(PREC, WRAPD_PROX_LEARNED))
k_local_learned(int n_e, int n_ep, const double* __restrict__ symF, double* __restrict__ U, double* __restrict__ P,
                const double* __restrict__ wsq, const double* __restrict__ vol, const double* __restrict__ kst,
                double* __restrict__ cand, double* __restrict__ sig, MatArgs m, int prox_iters, double rtol,
                double fallback_sigma, int update_cand, int use_hessian, double* __restrict__ partials,
                float* __restrict__ tlog, float* __restrict__ fbinfo, MlpDevice mlp, double gate, RedTail tail) {
    using T = typename std::conditional<PREC == WRAPD_FP32, float, double>::type;
    __shared__ double shk[11][kLocalThreads / 64];   // the block reductions' wave results
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int e = lb * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = e < n_e;
    double bef = 0.0, aft = 0.0, pe = 0.0, ratio = 0.0, iters = 0.0, fallback = 0.0, neg = 0.0, evals = 0.0;
    double accepted = 0.0, used = 0.0;
    bool defer = false;   // an FP32 result unusable here: redone in FP64 by k_local_fallback
    M3<double> S, Uo, Po;
    double w2 = 1.0, v = 1.0;
    NHParams<double> pd{1.0, 1.0};
    double sl[3] = {1.0, 1.0, 1.0};
#pragma unroll
    for (int i = 0; i < 9; ++i) S.m[i] = Uo.m[i] = Po.m[i] = 0.0;
    S.m[0] = S.m[4] = S.m[8] = 1.0;
    if (valid) {
        ld9(symF, n_ep, e, S);
        ld9(U, n_ep, e, Uo);
        ld9(P, n_ep, e, Po);
        w2 = wsq[e];
        v = vol[e];
        pd = mat_of(m, e);
        sl[0] = sig[e]; sl[1] = sig[n_ep + e]; sl[2] = sig[2 * n_ep + e];
        M3<double> D;
#pragma unroll
        for (int i = 0; i < 9; ++i) D.m[i] = S.m[i] - Po.m[i];
        bef = 0.5 * w2 * sqnorm9(D);
    }
    M3<double> Sin;
#pragma unroll
    for (int i = 0; i < 9; ++i) Sin.m[i] = S.m[i] + Uo.m[i];
    Sin = transpose(Sin);
    const M3<T> St = cvt<T>(Sin);
    T shat[3];
    M3<T> Ut, Vt;
    svd3(St, shat, Ut, Vt);
    ProxProblem<T> PP;
    PP.p.mu = static_cast<T>(pd.mu);
    PP.p.lambda = static_cast<T>(pd.lambda);
    PP.x0[0] = shat[0]; PP.x0[1] = shat[1]; PP.x0[2] = shat[2];
    PP.wov = static_cast<T>(w2 / v);
    const T xl[3] = {static_cast<T>(sl[0]), static_cast<T>(sl[1]), static_cast<T>(sl[2])};
    T fl, gl[3];
    const T rl = prox_residual(PP, xl, fl, gl);
    const T xln = wsqrt(dot3(xl, xl));
    const bool poor = valid && !(rl <= static_cast<T>(gate) * (PP.p.mu + PP.wov) * (xln > T(1) ? xln : T(1)));
    float dpred[kMlpOut] = {0.0f, 0.0f, 0.0f};
    // wave-uniform: the MLP is a whole-wave product, over the 16-element blocks with a poor start
    const unsigned long long poor_lanes = __ballot(poor);
    if (poor_lanes) {
        float feat[kMlpIn];
        mlp_features(mlp.norm, static_cast<float>(shat[0]), static_cast<float>(shat[1]), static_cast<float>(shat[2]),
                     static_cast<float>(w2 / v), static_cast<float>(pd.mu), static_cast<float>(pd.lambda), feat);
        mlp_wave_predict(mlp, feat, lane, dpred, poor_lanes);   // every lane of the wave takes part
    }
    if (valid) {
        const T xp[3] = {shat[0] + static_cast<T>(dpred[0]), shat[1] + static_cast<T>(dpred[1]),
                         shat[2] + static_cast<T>(dpred[2])};
        T fp = T(0), gp[3] = {T(0), T(0), T(0)};
        const T rp = poor ? prox_residual(PP, xp, fp, gp) : T(INFINITY);
        const bool use_pred = rp < rl;
        T x[3] = {use_pred ? xp[0] : xl[0], use_pred ? xp[1] : xl[1], use_pred ? xp[2] : xl[2]};
        const T xn = wsqrt(dot3(x, x));
        used = use_pred ? 1.0 : 0.0;
        accepted = (use_pred && rp <= static_cast<T>(rtol) * (PP.p.mu + PP.wov) * (xn > T(1) ? xn : T(1))) ? 1.0 : 0.0;
        T gn = T(0);
        // the start's objective and gradient are the residual's (when the start is valid)
        const bool have = use_pred || rl < T(INFINITY);
        const T g0[3] = {use_pred ? gp[0] : gl[0], use_pred ? gp[1] : gl[1], use_pred ? gp[2] : gl[2]};
        const int it = have ? prox_newton_from(PP, x, use_pred ? fp : fl, g0, prox_iters, static_cast<T>(rtol), gn)
                            : prox_newton(PP, x, prox_iters, static_cast<T>(rtol), gn);
        iters = it < 0 ? 0.0 : static_cast<double>(it);
        evals = iters;
        const M3<T> Zt = udvt(Vt, x, Ut);
        M3<double> Zn;
        bool ok = true;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            Zn.m[i] = static_cast<double>(Zt.m[i]);
            ok &= isfinite(Zn.m[i]);
        }
        double xd[3] = {static_cast<double>(x[0]), static_cast<double>(x[1]), static_cast<double>(x[2])};
        double shd[3] = {static_cast<double>(shat[0]), static_cast<double>(shat[1]), static_cast<double>(shat[2])};
        // an unusable FP32 result is redone in FP64 by k_local_fallback (Newton, 1e-12)
        defer = PREC == WRAPD_FP32 &&
                (!ok || static_cast<double>(shat[2]) < fallback_sigma || xd[0] < 0.0 || xd[1] < 0.0 || xd[2] < 0.0);
        if (PREC == WRAPD_FP32) fbinfo[e] = defer ? static_cast<float>(evals + 1.0) : 0.0f;
        if (defer) {
            fallback = 1.0;
            iters = 0.0;
        } else {
            pe = v * nh_psi(xd, pd);
            neg = (xd[0] < 0.0 || xd[1] < 0.0 || xd[2] < 0.0) ? 1.0 : 0.0;
            sig[e] = xd[0];
            sig[n_ep + e] = xd[1];
            sig[2 * n_ep + e] = xd[2];
            if (tlog) log_tuple(tlog, e, shd, w2 / v, pd, xd);
            M3<double> D, Un, Pn;
#pragma unroll
            for (int i = 0; i < 9; ++i) D.m[i] = S.m[i] - (Zn.m[i] - Uo.m[i]);
            aft = 0.5 * w2 * sqnorm9(D);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                Un.m[i] = Uo.m[i] + (S.m[i] - Zn.m[i]);
                Pn.m[i] = Zn.m[i] - Un.m[i];
            }
            st9(U, n_ep, e, Un);
            st9(P, n_ep, e, Pn);
            if (update_cand) {
                const double k = kst[e];
                double cs = k;
                if (use_hessian) cs = nh_candidate_stiffness(xd, pd, k * m.beta_min, k * m.beta_max);
                const double cw = v * cs;
                cand[e] = cw;
                const double a = w2 / cw, b = 1.0 / a;
                ratio = a > b ? a : b;
            }
        }
    }
    store_local_partials(partials, lb, bef, aft, pe, ratio, iters, fallback, neg, accepted, evals, defer ? 0.0 : evals,
                         used, shk, tail.nslots);
    reduce_tail<kLocalThreads, 11>(tail, partials);
}

// Rotation-unaware local step (src/TetElements.cpp:193-254); the cache holds Ft = F^T.
template <int MODEL>
__global__ void __launch_bounds__(kLocalThreads)
k_local_ua(int n_e, int n_ep, const double* __restrict__ Ftc, double* __restrict__ U, double* __restrict__ P,
           const double* __restrict__ wsq, const double* __restrict__ vol, const double* __restrict__ kst,
           double* __restrict__ cand, double* __restrict__ sig, MatArgs m, int prox_iters, int update_cand,
           int use_hessian) {
    const int e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (e >= n_e) return;
    M3<double> Dix, Uo;
    ld9(Ftc, n_ep, e, Dix);
    ld9(U, n_ep, e, Uo);
    M3<double> Ft;
#pragma unroll
    for (int i = 0; i < 9; ++i) Ft.m[i] = Dix.m[i] + Uo.m[i];
    const M3<double> F = transpose(Ft);
    double s[3];
    M3<double> Us, Vs;
    svd3(F, s, Us, Vs);
    double x[3] = {1.0, 1.0, 1.0};
    if (MODEL == 1) {
        const NHParams<double> pd = mat_of(m, e);
        ProxProblem<double> Pp;
        Pp.p = pd;
        Pp.x0[0] = s[0]; Pp.x0[1] = s[1]; Pp.x0[2] = s[2];
        Pp.wov = wsq[e] / vol[e];
        x[0] = sig[e]; x[1] = sig[n_ep + e]; x[2] = sig[2 * n_ep + e];
        int evals = 0;
        prox_lbfgs(Pp, x, prox_iters, evals);
        sig[e] = x[0]; sig[n_ep + e] = x[1]; sig[2 * n_ep + e] = x[2];
        Ft = udvt(Vs, x, Us);
    } else {
        const M3<double> R = matmul(Us, transpose(Vs));
        M3<double> Fr;
#pragma unroll
        for (int i = 0; i < 9; ++i) Fr.m[i] = 0.5 * (F.m[i] + R.m[i]);
        Ft = transpose(Fr);
    }
    M3<double> Un, Pn;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        Un.m[i] = Uo.m[i] + (Dix.m[i] - Ft.m[i]);
        Pn.m[i] = Ft.m[i] - Un.m[i];
    }
    st9(U, n_ep, e, Un);
    st9(P, n_ep, e, Pn);
    if (update_cand) {
        const double k = kst[e];
        double cs = k;
        if (use_hessian) cs = nh_candidate_stiffness(x, mat_of(m, e), k * m.beta_min, k * m.beta_max);
        cand[e] = vol[e] * cs;
    }
}

__global__ void __launch_bounds__(kElemThreads)
k_update_w(int n_e, int n_ep, double* __restrict__ U, double* __restrict__ P, double* __restrict__ wsq,
           const double* __restrict__ cand) {
    const int e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (e >= n_e) return;
    const double r = wsq[e] / cand[e];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const double u = U[i * n_ep + e];
        const double z = P[i * n_ep + e] + u;
        const double un = r * u;
        U[i * n_ep + e] = un;
        P[i * n_ep + e] = z - un;
    }
    wsq[e] = cand[e];
}

__global__ void __launch_bounds__(kElemThreads)
k_max_ratio(int n_e, const double* __restrict__ wsq, const double* __restrict__ cand, double* __restrict__ partials,
            RedTail tail) {
    __shared__ double sh[kElemThreads / 64];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int e = lb * blockDim.x + threadIdx.x;
    double r = 0.0;
    if (e < n_e) {
        const double a = wsq[e] / cand[e], b = 1.0 / a;
        r = a > b ? a : b;
    }
    const double t = block_max<kElemThreads>(r, sh);
    if (threadIdx.x == 0 && tail.nslots > 0) put_partial(partials, lb, 0, t);
    reduce_tail<kElemThreads, 1>(tail, partials);
}

// A(slot) = sum over (element, a, b) contributions of w_e^2 d_a . d_b, in element order
__global__ void __launch_bounds__(kElemThreads)
k_assemble_A(int nnz, int n_ep, const int* __restrict__ cptr, const int* __restrict__ cidx,
             const double* __restrict__ Di, const double* __restrict__ wsq, double* __restrict__ Ax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    double acc = 0.0;
    for (int q = cptr[i]; q < cptr[i + 1]; ++q) {
        const int c = cidx[q];
        const int e = c >> 4, a = (c >> 2) & 3, b = c & 3;
        const double dd = ldD(Di, n_ep, 0, a, e) * ldD(Di, n_ep, 0, b, e) +
                          ldD(Di, n_ep, 1, a, e) * ldD(Di, n_ep, 1, b, e) +
                          ldD(Di, n_ep, 2, a, e) * ldD(Di, n_ep, 2, b, e);
        acc += wsq[e] * dd;
    }
    Ax[i] = acc;
}

// ---------------------------------------------------------------------------------
static MatArgs mat_args(wrapd_ctx* c) {
    MatArgs m;
    m.mu_e = c->mu_e;
    m.lam_e = c->lam_e;
    m.mu = c->mat.mu;
    m.lam = c->mat.lambda;
    m.beta_min = c->mat.beta_min;
    m.beta_max = c->mat.beta_max;
    return m;
}

// the device SVD on n given matrices (row-major), in T arithmetic; for checking svd3
template <typename T>
__global__ void __launch_bounds__(kElemThreads)
k_svd_batch(int n, const double* __restrict__ F, double* __restrict__ sig, double* __restrict__ U,
            double* __restrict__ V) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    M3<T> f, u, v;
#pragma unroll
    for (int i = 0; i < 9; ++i) f.m[i] = static_cast<T>(F[9ll * e + i]);
    T s[3];
    svd3(f, s, u, v);
#pragma unroll
    for (int i = 0; i < 3; ++i) sig[3ll * e + i] = static_cast<double>(s[i]);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        U[9ll * e + i] = static_cast<double>(u.m[i]);
        V[9ll * e + i] = static_cast<double>(v.m[i]);
    }
}

void launch_svd_batch(wrapd_ctx* c, int n, int precision, const double* F, double* sig, double* U, double* V) {
    if (precision == WRAPD_FP32)
        hipLaunchKernelGGL(k_svd_batch<float>, dim3(blocks_for(n, kElemThreads)), dim3(kElemThreads), 0, c->stream, n,
                           F, sig, U, V);
    else
        hipLaunchKernelGGL(k_svd_batch<double>, dim3(blocks_for(n, kElemThreads)), dim3(kElemThreads), 0, c->stream,
                           n, F, sig, U, V);
}

void launch_setup(wrapd_ctx* c, const int* d_tv, const double* d_rest, int* d_bad) {
    ProfScope ps(c, "setup");
    hipLaunchKernelGGL(k_setup, dim3(blocks_for(c->n_ep, kElemThreads)), dim3(kElemThreads), 0, c->stream, c->n_e,
                       c->n_ep, reinterpret_cast<const int4*>(d_tv), d_rest, c->Di, c->vol, c->kst, c->wsq, c->sig,
                       mat_args(c), c->mat.model, c->mat.beta_static, d_bad);
}

void launch_init_duals(wrapd_ctx* c, int ra) {
    ProfScope ps(c, "init_duals");
    const int nb = blocks_for(c->n_e, kElemThreads);
    auto k = ra ? k_init_duals<1> : k_init_duals<0>;
    hipLaunchKernelGGL(k, dim3(nb), dim3(kElemThreads), 0, c->stream, c->n_e, c->n_ep, c->n_f,
                       reinterpret_cast<const int4*>(c->vidx), c->Di, c->xall, c->vol, c->U, c->P, mat_args(c),
                       c->mat.model, c->partials, take_tail(c, nb, true));
    finish_reduction(c);
}

template <int MODE>
static void defo_mode(wrapd_ctx* c, int use_hessian) {
    const int nb = blocks_for(c->n_e, kElemThreads);
    auto k = c->mat.model == 1 ? k_defo<MODE, 1> : k_defo<MODE, 0>;
    const RedTail t = take_tail(c, nb, true);   // used by the modes that reduce (value, energy)
    hipLaunchKernelGGL(k, dim3(nb), dim3(kElemThreads), 0, c->stream, c->n_e, c->n_ep, c->n_f,
                       reinterpret_cast<const int4*>(c->vidx), c->Di, c->xall, c->P, c->wsq, c->vol, c->kst,
                       c->symF, c->gstage, reinterpret_cast<const int4*>(c->cpos), c->cand, mat_args(c), use_hessian,
                       c->partials, t);
    finish_reduction(c);
}

void launch_defo(wrapd_ctx* c, int mode, int use_hessian) {
    ProfScope ps(c, (mode & DM_GRAD) ? "defo_grad" : ((mode & DM_BVEC) ? "defo_b" : "defo"));
    switch (mode) {
        case DM_GRAD | DM_VALUE: defo_mode<DM_GRAD | DM_VALUE>(c, use_hessian); break;
        case DM_GRAD | DM_VALUE | DM_WRITE_SYMF | DM_ENERGY:
            defo_mode<DM_GRAD | DM_VALUE | DM_WRITE_SYMF | DM_ENERGY>(c, use_hessian); break;
        case DM_VALUE | DM_WRITE_SYMF | DM_ENERGY: defo_mode<DM_VALUE | DM_WRITE_SYMF | DM_ENERGY>(c, use_hessian); break;
        case DM_VALUE: defo_mode<DM_VALUE>(c, use_hessian); break;
        case DM_WRITE_SYMF | DM_ENERGY: defo_mode<DM_WRITE_SYMF | DM_ENERGY>(c, use_hessian); break;
        case DM_CAND: defo_mode<DM_CAND>(c, use_hessian); break;
        case DM_CAND | DM_DIRECT: defo_mode<DM_CAND | DM_DIRECT>(c, use_hessian); break;
        case DM_FT_CACHE | DM_ENERGY: defo_mode<DM_FT_CACHE | DM_ENERGY>(c, use_hessian); break;
        case DM_FT_CACHE | DM_CAND: defo_mode<DM_FT_CACHE | DM_CAND>(c, use_hessian); break;
        case DM_BVEC: defo_mode<DM_BVEC>(c, use_hessian); break;
        default: break;
    }
}

template <int MODEL, int PREC, int PROX>
static void local_inst(wrapd_ctx* c, const wrapd_local_args* a, float* tlog) {
    const int nb = blocks_for(c->n_e, kLocalThreads);
    hipLaunchKernelGGL((k_local<MODEL, PREC, PROX>), dim3(nb), dim3(kLocalThreads), 0, c->stream, c->n_e, c->n_ep,
                       c->symF, c->U, c->P, c->wsq, c->vol, c->kst, c->cand, c->sig, mat_args(c), a->prox_iters,
                       a->newton_rtol, a->fallback_sigma, a->update_cand, a->cand_use_hessian, c->partials, tlog,
                       PREC == WRAPD_FP32 ? c->fb_info : nullptr, take_tail(c, nb, true));
}

template <int PREC>
static void local_learned_inst(wrapd_ctx* c, const wrapd_local_args* a, float* tlog) {
    const int nb = blocks_for(c->n_e, kLocalThreads);
    hipLaunchKernelGGL((k_local_learned<PREC>), dim3(nb), dim3(kLocalThreads), 0, c->stream, c->n_e, c->n_ep,
                       c->symF, c->U, c->P, c->wsq, c->vol, c->kst, c->cand, c->sig, mat_args(c), a->prox_iters,
                       a->newton_rtol, a->fallback_sigma, a->update_cand, a->cand_use_hessian, c->partials, tlog,
                       PREC == WRAPD_FP32 ? c->fb_info : nullptr, mlp_device(c), c->learned_gate, take_tail(c, nb, true));
}

constexpr int kFallbackBlocks = 256;   // the FP64 pass's fixed grid

int prepare_local_fallback(wrapd_ctx* c) {
    size_t bytes = 0;
    if (check_hip(c, hipcub::DeviceSelect::Flagged(nullptr, bytes, hipcub::CountingInputIterator<int>(0), c->fb_info,
                                                   c->fb_list, c->fb_count, c->n_e, c->stream),
                  "fallback list size"))
        return WRAPD_ERR_HIP;
    if (c->fb_tmp) { (void)hipFree(c->fb_tmp); c->fb_tmp = nullptr; }
    c->fb_tmp_bytes = std::max<size_t>(bytes, 1);
    return check_hip(c, hipMalloc(&c->fb_tmp, c->fb_tmp_bytes), "hipMalloc(fallback list)") ? WRAPD_ERR_HIP : WRAPD_OK;
}

// The local step: the element kernel and, for FP32 steps, the compaction of the deferred
// elements and their FP64 redo (its sums armed into S_LOCAL_FB), all in the local-step
// phase.  A wide-grid element kernel finishes its sums in a separate launch (finish_reduction).
static void launch_local_kernel(wrapd_ctx* c, const wrapd_local_args* a, float* tlog);
void launch_local(wrapd_ctx* c, const wrapd_local_args* a) {
    ProfScope ps(c, "local_step");
    // tuple log: the whole step is appended when it fits
    float* tlog = nullptr;
    if (c->mlp_logging && c->mat.model == 1 && c->mlp_log_count + c->n_e <= c->mlp_log_cap) {
        tlog = c->mlp_log + 9ll * c->mlp_log_count;
        c->mlp_log_count += c->n_e;
    }
    launch_local_kernel(c, a, tlog);
    finish_reduction(c);
    c->local_fb = c->mat.model == 1 && a->precision == WRAPD_FP32;
    if (!c->local_fb) return;
    (void)hipcub::DeviceSelect::Flagged(c->fb_tmp, c->fb_tmp_bytes, hipcub::CountingInputIterator<int>(0), c->fb_info,
                                        c->fb_list, c->fb_count, c->n_e, c->stream);
    arm_tail(c, 11, kLocalMaxMask, S_LOCAL_FB);
    const bool lbfgs = a->prox_mode == WRAPD_PROX_LBFGS;
    auto k = lbfgs ? k_local_fallback<WRAPD_PROX_LBFGS> : k_local_fallback<WRAPD_PROX_NEWTON>;
    hipLaunchKernelGGL(k, dim3(kFallbackBlocks), dim3(kLocalThreads), 0, c->stream, c->n_ep, c->fb_list, c->fb_count,
                       c->symF, c->U, c->P, c->wsq, c->vol, c->kst, c->cand, c->sig, mat_args(c), a->prox_iters,
                       lbfgs ? a->newton_rtol : 1e-12, a->update_cand, a->cand_use_hessian, c->fb_info, c->partials,
                       tlog, take_tail(c, kFallbackBlocks, false));
}

static void launch_local_kernel(wrapd_ctx* c, const wrapd_local_args* a, float* tlog) {
    if (c->mat.model != 1) { local_inst<0, WRAPD_FP64, WRAPD_PROX_LBFGS>(c, a, nullptr); return; }
    const bool f32 = a->precision == WRAPD_FP32;
    if (a->prox_mode == WRAPD_PROX_LEARNED) {
        if (f32) local_learned_inst<WRAPD_FP32>(c, a, tlog);
        else local_learned_inst<WRAPD_FP64>(c, a, tlog);
        return;
    }
    const bool newton = a->prox_mode != WRAPD_PROX_LBFGS;
    if (!f32 && !newton) local_inst<1, WRAPD_FP64, WRAPD_PROX_LBFGS>(c, a, tlog);
    else if (!f32 && newton) local_inst<1, WRAPD_FP64, WRAPD_PROX_NEWTON>(c, a, tlog);
    else if (f32 && newton) local_inst<1, WRAPD_FP32, WRAPD_PROX_NEWTON>(c, a, tlog);
    else local_inst<1, WRAPD_FP32, WRAPD_PROX_LBFGS>(c, a, tlog);
}

void launch_local_unaware(wrapd_ctx* c, const wrapd_local_args* a) {
    ProfScope ps(c, "local_step_ua");
    const int nb = blocks_for(c->n_e, kLocalThreads);
    auto k = c->mat.model == 1 ? k_local_ua<1> : k_local_ua<0>;
    hipLaunchKernelGGL(k, dim3(nb), dim3(kLocalThreads), 0, c->stream, c->n_e, c->n_ep, c->symF, c->U, c->P, c->wsq,
                       c->vol, c->kst, c->cand, c->sig, mat_args(c), a->prox_iters, a->update_cand,
                       a->cand_use_hessian);
}

void launch_update_weights(wrapd_ctx* c) {
    ProfScope ps(c, "update_weights");
    hipLaunchKernelGGL(k_update_w, dim3(blocks_for(c->n_e, kElemThreads)), dim3(kElemThreads), 0, c->stream, c->n_e,
                       c->n_ep, c->U, c->P, c->wsq, c->cand);
}

void launch_max_ratio(wrapd_ctx* c) {
    ProfScope ps(c, "max_ratio");
    const int nb = blocks_for(c->n_e, kElemThreads);
    hipLaunchKernelGGL(k_max_ratio, dim3(nb), dim3(kElemThreads), 0, c->stream, c->n_e, c->wsq, c->cand, c->partials,
                       take_tail(c, nb, true));
    finish_reduction(c);
}

void launch_assemble_A(wrapd_ctx* c) {
    ProfScope ps(c, "assemble_A");
    hipLaunchKernelGGL(k_assemble_A, dim3(blocks_for(c->nnzA, kElemThreads)), dim3(kElemThreads), 0, c->stream,
                       c->nnzA, c->n_ep, c->a_cptr, c->a_cidx, c->Di, c->wsq, c->Ax);
}

}  // namespace wd
