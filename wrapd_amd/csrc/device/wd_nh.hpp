// Neo-Hookean stretch density in singular values and the per-element stretch prox.
//
// Density (src/TetElements.hpp:61-77, 93-144): with I1 = |s|^2, J = s0 s1 s2, J0 = 0.1,
//   psi = mu/2 (I1 - 3) - mu l + lambda/2 l^2,  l = log J for J > J0, else the quadratic
//   extension f(J) = log J0 + (J - J0)/J0 - ((J/J0) - 1)^2 / 2.
// The prox minimises psi(s) + (w^2/vol)/2 |s - s_hat|^2 (src/TetElements.cpp:302-318).
//
// Two prox solvers:
//   prox_lbfgs  - the reference's solver step for step: L-BFGS with memory 8, gamma
//                 scaling, steepest-descent reset on a non-descent direction, cubic
//                 backtracking (c = 1e-4, steps clamped to [0.1a, 0.5a]), stop when the
//                 pre-step gradient norm or the step is below 1e-8, at most `max_iters`
//                 iterations (deps/mcloptlib/include/MCL/LBFGS.hpp:52-152,
//                 Backtracking.hpp:111-175, src/TetElements.hpp:40-42).
//   prox_newton - Newton on the same objective with the PSD-projected density Hessian
//                 (src/TetElements.hpp:109-149) plus w^2/vol I; full steps once they are
//                 small (where an Armijo test only sees rounding noise), Armijo
//                 backtracking before that; stops on a gradient-norm tolerance relative
//                 to the objective's scale or on a step below the resolution of x.
#pragma once

#include "wd_math.hpp"

namespace wd {

template <typename T>
struct NHParams {
    T mu, lambda;
};

template <typename T>
WD_HD T nh_fJ(T J) {
    const T J0 = T(0.1);
    const T t = (J / J0) - T(1);
    return wlog(J0) + ((J - J0) / J0) - (T(0.5) * (t * t));
}

// Density without the negative-singular-value barrier (what the energy reports use).
template <typename T>
WD_HD T nh_psi(const T x[3], NHParams<T> p) {
    const T J0 = T(0.1);
    const T I1 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    const T J = x[0] * x[1] * x[2];
    T e = T(0.5) * p.mu * (I1 - T(3));
    const T l = J > J0 ? wlog(J) : nh_fJ(J);
    e += -p.mu * l + T(0.5) * p.lambda * l * l;
    return e;
}

template <typename T>
WD_HD T nh_energy_density(const T x[3], NHParams<T> p) {
    if (x[0] < T(0) || x[1] < T(0) || x[2] < T(0)) return T(1e15);
    return nh_psi(x, p);
}

template <typename T>
WD_HD void nh_grad_density(const T x[3], NHParams<T> p, T g[3]) {
    const T J0 = T(0.1);
    const T J = x[0] * x[1] * x[2];
    T c;
    if (J > J0) {
        c = p.lambda * wlog(J) - p.mu;
    } else {
        const T f = nh_fJ(J);
        const T df = T(1) / J0 - (J - J0) / (J0 * J0);
        c = (p.lambda * f - p.mu) * df * J;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) g[i] = p.mu * x[i] + c * (T(1) / x[i]);
}

// Density Hessian before the PSD projection.
template <typename T>
WD_HD void nh_hess_density_raw(const T x[3], NHParams<T> p, M3<T>& H) {
    const T J0 = T(0.1);
    const T J = x[0] * x[1] * x[2];
    if (J > J0) {
        const T lJ = wlog(J);
        const T ct = p.mu - p.lambda * (lJ - T(1));
        H(0, 0) = p.mu + ct / (x[0] * x[0]);
        H(1, 1) = p.mu + ct / (x[1] * x[1]);
        H(2, 2) = p.mu + ct / (x[2] * x[2]);
        H(0, 1) = H(1, 0) = p.lambda / (x[0] * x[1]);
        H(0, 2) = H(2, 0) = p.lambda / (x[0] * x[2]);
        H(1, 2) = H(2, 1) = p.lambda / (x[1] * x[2]);
    } else {
        const T f = nh_fJ(J);
        const T df = T(1) / J0 - (J - J0) / (J0 * J0);
        const T lfm = p.lambda * f - p.mu;
        const T a12 = x[1] * x[2], a02 = x[0] * x[2], a01 = x[0] * x[1];
        H(0, 0) = p.mu - lfm * ((a12 / J0) * (a12 / J0)) + p.lambda * ((df * a12) * (df * a12));
        H(1, 1) = p.mu - lfm * ((a02 / J0) * (a02 / J0)) + p.lambda * ((df * a02) * (df * a02));
        H(2, 2) = p.mu - lfm * ((a01 / J0) * (a01 / J0)) + p.lambda * ((df * a01) * (df * a01));
        const T ct = lfm * (df - J / (J0 * J0)) + p.lambda * J * (df * df);
        H(0, 1) = H(1, 0) = x[2] * ct;
        H(0, 2) = H(2, 0) = x[1] * ct;
        H(1, 2) = H(2, 1) = x[0] * ct;
    }
}

// Candidate stiffness of the reweighting rule (src/TetElements.cpp:406-420):
// clamp(lambda_max(makePD(H)), lo, hi); lambda_max(makePD(H)) = max(lambda_max(H), 0)
// and lo > 0, so the projection never changes the clamped result.
template <typename T>
WD_HD T nh_candidate_stiffness(const T x[3], NHParams<T> p, T lo, T hi) {
    M3<T> H;
    nh_hess_density_raw(x, p, H);
    T k = sym3_lambda_max(H);
    if (k < lo) k = lo;
    if (k > hi) k = hi;
    return k;
}

// ---------------------------------------------------------------------------------
template <typename T>
struct ProxProblem {
    NHParams<T> p;
    T x0[3];
    T wov;   // w^2 / vol
    WD_HD T value(const T x[3]) const {
        const T d0 = x[0] - x0[0], d1 = x[1] - x0[1], d2 = x[2] - x0[2];
        return nh_energy_density(x, p) + (wov * T(0.5)) * (d0 * d0 + d1 * d1 + d2 * d2);
    }
    WD_HD T gradient(const T x[3], T g[3]) const {
        nh_grad_density(x, p, g);
#pragma unroll
        for (int i = 0; i < 3; ++i) g[i] += wov * (x[i] - x0[i]);
        return value(x);
    }
};

template <typename T>
WD_HD T dot3(const T* a, const T* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <typename T>
WD_HD T cubic_step(T fx0, T gtp, T fxa, T alpha, T fxp, T alphap) {
    const T mult = T(1) / (alpha * alpha * alphap * alphap * (alpha - alphap));
    const T A00 = alphap * alphap, A01 = -alpha * alpha;
    const T A10 = -alphap * alphap * alphap, A11 = alpha * alpha * alpha;
    const T B0 = fxa - fx0 - alpha * gtp, B1 = fxp - fx0 - alphap * gtp;
    const T r0 = (mult * A00) * B0 + (mult * A01) * B1;
    const T r1 = (mult * A10) * B0 + (mult * A11) * B1;
    if (wabs(r0) <= T(0)) return -gtp / (T(2) * r1);
    const T d = wsqrt(r1 * r1 - T(3) * r0 * gtp);
    return (-r1 + d) / (T(3) * r0);
}

// Returns the accepted step, or -1 when the search exhausts its iteration budget.
template <typename T>
WD_HD T backtracking_cubic(const T x[3], const T p[3], const ProxProblem<T>& P, T alpha0, int& evals) {
    const T decrease = T(1e-4);
    const int max_iters = 100000;
    if (wsqrt(dot3(p, p)) <= Eps<T>::v) return decrease;
    T alpha = alpha0, g[3];
    const T fx0 = P.gradient(x, g);
    const T gtp = dot3(g, p);
    T fxp = fx0, alphap = alpha;
    int iter = 0;
    for (; iter < max_iters; ++iter) {
        const T xa[3] = {x[0] + alpha * p[0], x[1] + alpha * p[1], x[2] + alpha * p[2]};
        const T fxa = P.value(xa);
        ++evals;
        if (fxa <= fx0 + alpha * decrease * gtp) break;
        const T at = iter == 0 ? (gtp / (T(2) * (fx0 + gtp - fxa))) : cubic_step(fx0, gtp, fxa, alpha, fxp, alphap);
        fxp = fxa;
        alphap = alpha;
        const T lo = T(0.1) * alpha, hi = T(0.5) * alpha;
        alpha = at < lo ? lo : (at > hi ? hi : at);
        // A NaN step never satisfies the Armijo test again (every later trial value and
        // interpolation is NaN), so the loop above would only exhaust max_iters and fail:
        // return that failure now.
        if (alpha != alpha) return T(-1);
    }
    if (iter >= max_iters) return T(-1);
    return alpha;
}

// L-BFGS(8) exactly as the reference prox; returns iterations (or -1 on line-search failure).
template <typename T>
WD_HD int prox_lbfgs(const ProxProblem<T>& P, T x[3], int max_iters_in, int& evals) {
    constexpr int M = 8;
    T s[M][3], y[M][3], alpha[M], rho[M];
    T grad[3], q[3], grad_old[3], x_old[3];
    P.gradient(x, grad);
    T gamma_k = T(1), alpha_init = T(1);
    int global_iter = 0, max_iters = max_iters_in;
    for (int k = 0; k < max_iters; ++k) {
#pragma unroll
        for (int i = 0; i < 3; ++i) { x_old[i] = x[i]; grad_old[i] = grad[i]; q[i] = grad[i]; }
        global_iter++;
        const int iter = k < M ? k : M;
#pragma unroll
        for (int i = M - 1; i >= 0; --i) {
            if (i < iter) {
                rho[i] = T(1) / dot3(s[i], y[i]);
                alpha[i] = rho[i] * dot3(s[i], q);
#pragma unroll
                for (int c = 0; c < 3; ++c) q[c] = q[c] - alpha[i] * y[i][c];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) q[c] = gamma_k * q[c];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            if (i < iter) {
                const T beta = rho[i] * dot3(q, y[i]);
#pragma unroll
                for (int c = 0; c < 3; ++c) q[c] = q[c] + (alpha[i] - beta) * s[i][c];
            }
        }
        const T dir = dot3(q, grad);
        if (dir <= T(0)) {
#pragma unroll
            for (int c = 0; c < 3; ++c) q[c] = grad[c];
            max_iters -= k;
            k = 0;
            T gi = wabs(grad[0]);
            gi = wabs(grad[1]) > gi ? wabs(grad[1]) : gi;
            gi = wabs(grad[2]) > gi ? wabs(grad[2]) : gi;
            const T inv = T(1) / gi;
            alpha_init = inv < T(1) ? inv : T(1);
        }
        const T mq[3] = {-q[0], -q[1], -q[2]};
        const T rate = backtracking_cubic(x, mq, P, alpha_init, evals);
        if (rate <= T(0)) return -1;
        T x_last[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) { x_last[c] = x[c]; x[c] -= rate * q[c]; }
        {
            const T gn = wsqrt(dot3(grad, grad));
            const T d0 = x_last[0] - x[0], d1 = x_last[1] - x[1], d2 = x_last[2] - x[2];
            if (gn < T(1e-8) || wsqrt(d0 * d0 + d1 * d1 + d2 * d2) < T(1e-8)) break;
        }
        P.gradient(x, grad);
        T st[3], yt[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) { st[c] = x[c] - x_old[c]; yt[c] = grad[c] - grad_old[c]; }
        if (k < M) {
#pragma unroll
            for (int i = 0; i < M; ++i)
                if (i == k) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) { s[i][c] = st[c]; y[i][c] = yt[c]; }
                }
        } else {
#pragma unroll
            for (int i = 0; i < M - 1; ++i)
#pragma unroll
                for (int c = 0; c < 3; ++c) { s[i][c] = s[i + 1][c]; y[i][c] = y[i + 1][c]; }
#pragma unroll
            for (int c = 0; c < 3; ++c) { s[M - 1][c] = st[c]; y[M - 1][c] = yt[c]; }
        }
        const T denom = dot3(yt, yt);
        if (wabs(denom) <= T(0)) break;
        gamma_k = dot3(st, yt) / denom;
        alpha_init = T(1);
    }
    return global_iter;
}

// Solve a symmetric positive-definite 3x3 system H d = g (Cholesky).  Returns false if
// H is not numerically positive definite.
template <typename T>
WD_HD bool spd_solve3(const M3<T>& H, const T g[3], T d[3]) {
    const T l00s = H(0, 0);
    if (!(l00s > T(0))) return false;
    const T l00 = wsqrt(l00s);
    const T l10 = H(1, 0) / l00, l20 = H(2, 0) / l00;
    const T l11s = H(1, 1) - l10 * l10;
    if (!(l11s > T(0))) return false;
    const T l11 = wsqrt(l11s);
    const T l21 = (H(2, 1) - l20 * l10) / l11;
    const T l22s = H(2, 2) - l20 * l20 - l21 * l21;
    if (!(l22s > T(0))) return false;
    const T l22 = wsqrt(l22s);
    const T z0 = g[0] / l00;
    const T z1 = (g[1] - l10 * z0) / l11;
    const T z2 = (g[2] - l20 * z0 - l21 * z1) / l22;
    d[2] = z2 / l22;
    d[1] = (z1 - l21 * d[2]) / l11;
    d[0] = (z0 - l10 * d[1] - l20 * d[2]) / l00;
    return true;
}

// Newton's method on the prox objective.  The density Hessian is made positive
// semi-definite by lifting its negative eigenvalues to zero through a diagonal shift
// (max(0, -lambda_min) I), which with the +w^2/vol I term keeps every step a descent
// direction.  Convergence: |grad| <= rtol * (mu + w^2/vol) * max(1, |x|).
// Returns the iteration count, or -1 when the iterate would leave the feasible region
// without making progress.
// f0, g0: the objective and gradient at the start x (P.gradient), when the caller has them.
template <typename T>
WD_HD int prox_newton_from(const ProxProblem<T>& P, T x[3], T f0, const T g0[3], int max_iters, T rtol,
                           T& gnorm_out) {
    const T scale = (P.p.mu + P.wov);
    T g[3] = {g0[0], g0[1], g0[2]};
    T f = f0;
    int it = 0;
    for (; it < max_iters; ++it) {
        const T gn = wsqrt(dot3(g, g));
        const T xn = wsqrt(dot3(x, x));
        gnorm_out = gn;
        if (gn <= rtol * scale * (xn > T(1) ? xn : T(1))) return it;
        M3<T> H;
        nh_hess_density_raw(x, P.p, H);
        M3<T> Hw = H;
        Hw(0, 0) += P.wov;
        Hw(1, 1) += P.wov;
        Hw(2, 2) += P.wov;
        T d[3];
        if (!spd_solve3(Hw, g, d)) {
            // indefinite density Hessian: lift its spectrum to >= 0 (the projection the
            // reference's makePD performs, src/Math.hpp:243-261, here as a shift by the
            // exact smallest eigenvalue) before adding w^2/vol
            M3<T> Hn = H;
#pragma unroll
            for (int i = 0; i < 9; ++i) Hn.m[i] = -H.m[i];
            const T lmin = -sym3_lambda_max(Hn);
            const T shift = lmin < T(0) ? -lmin * (T(1) + T(16) * Eps<T>::v) : T(0);
            Hw(0, 0) += shift;
            Hw(1, 1) += shift;
            Hw(2, 2) += shift;
            if (!spd_solve3(Hw, g, d)) {
                const T inv = T(1) / scale;
                d[0] = g[0] * inv; d[1] = g[1] * inv; d[2] = g[2] * inv;
            }
        }
        // Near the minimiser the predicted decrease g.d falls below the rounding error of f
        // (f sums terms far larger than itself), so an Armijo test there only sees noise:
        // steps of at most 1e-4 |x| (FP64) or 1e-2 |x| (FP32) are taken whole (Newton's
        // quadratic region), and a step below the resolution of x ends the iteration.
        const T xmax = wmax(T(1), wmax(wabs(x[0]), wmax(wabs(x[1]), wabs(x[2]))));
        const T dmax = wmax(wabs(d[0]), wmax(wabs(d[1]), wabs(d[2])));
        const T full_step = sizeof(T) == 4 ? T(1e-2) : T(1e-4);
        if (dmax <= full_step * xmax && x[0] - d[0] > T(0) && x[1] - d[1] > T(0) && x[2] - d[2] > T(0)) {
#pragma unroll
            for (int c = 0; c < 3; ++c) x[c] -= d[c];
            f = P.gradient(x, g);
            if (dmax <= T(4) * Eps<T>::v * xmax) {
                gnorm_out = wsqrt(dot3(g, g));
                return it + 1;
            }
            continue;
        }
        // Armijo backtracking; infeasible points cost 1e15 through the barrier.  The test
        // allows a few ulps of |f| for the same rounding reason.
        T a = T(1);
        const T gtd = dot3(g, d);
        const T fslack = T(8) * Eps<T>::v * wabs(f);
        T xn3[3];
        T fn = f;
        bool ok = false;
        for (int ls = 0; ls < 40; ++ls) {
#pragma unroll
            for (int c = 0; c < 3; ++c) xn3[c] = x[c] - a * d[c];
            fn = P.value(xn3);
            if (fn <= f - T(1e-4) * a * gtd + fslack) { ok = true; break; }
            a *= T(0.5);
        }
        if (!ok) return it;   // no further decrease possible at this precision
#pragma unroll
        for (int c = 0; c < 3; ++c) x[c] = xn3[c];
        f = P.gradient(x, g);
    }
    gnorm_out = wsqrt(dot3(g, g));
    return it;
}

template <typename T>
WD_HD int prox_newton(const ProxProblem<T>& P, T x[3], int max_iters, T rtol, T& gnorm_out) {
    T g[3];
    const T f = P.gradient(x, g);
    return prox_newton_from(P, x, f, g, max_iters, rtol, gnorm_out);
}

}  // namespace wd
