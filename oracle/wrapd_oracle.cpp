// Oracle: CPU FP64 restatement of the reference WRAPD ADMM path.
//
// TEST INFRASTRUCTURE ONLY (see wrapd_oracle.h).  Written from the mathematics of
// the reference sources; every function cites the reference lines it restates.
// The reference is compiled where it lies into oracle/_ref by oracle/Makefile and
// the tests under tests/ pin this restatement against it (golden vectors in
// tests/golden/) and against closed forms.
//
// Deliberate differences from the reference, none of which changes the exact
// arithmetic result beyond round-off:
//   * 3x3 SVD by one-sided Jacobi (Golub & Van Loan 8.6.3) instead of Eigen's
//     JacobiSVD with QR preconditioning; same sign convention (src/Math.hpp:147-160).
//   * sparse LDL^T (Davis, ACM TOMS 849) with reverse Cuthill-McKee ordering
//     instead of Eigen SimplicialLDLT with AMD (deps/mcloptlib/include/MCL/LinearSolver.hpp:279-298).
//   * eigenvalues of the 3x3 prox Hessian by symmetric Jacobi instead of
//     Eigen EigenSolver (src/TetElements.cpp:416).
//   * process exit paths (NaN decrement, negative prox singular value) set an error
//     status instead of calling exit(0).
#include "wrapd_oracle.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------
// 3x3 helpers (row-major)
struct M3 {
    double m[9];
    double& operator()(int r, int c) { return m[3 * r + c]; }
    double operator()(int r, int c) const { return m[3 * r + c]; }
};

M3 zero3() { M3 a; std::memset(a.m, 0, sizeof(a.m)); return a; }
M3 eye3() { M3 a = zero3(); a(0, 0) = a(1, 1) = a(2, 2) = 1.0; return a; }
M3 tr(const M3& a) { M3 b; for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) b(r, c) = a(c, r); return b; }
M3 mul(const M3& a, const M3& b) {
    M3 c;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int j = 0; j < 3; ++j) s += a(r, j) * b(j, k);
            c(r, k) = s;
        }
    return c;
}
M3 add(const M3& a, const M3& b) { M3 c; for (int i = 0; i < 9; ++i) c.m[i] = a.m[i] + b.m[i]; return c; }
M3 sub(const M3& a, const M3& b) { M3 c; for (int i = 0; i < 9; ++i) c.m[i] = a.m[i] - b.m[i]; return c; }
M3 scale(double s, const M3& a) { M3 c; for (int i = 0; i < 9; ++i) c.m[i] = s * a.m[i]; return c; }
double sqnorm(const M3& a) { double s = 0; for (int i = 0; i < 9; ++i) s += a.m[i] * a.m[i]; return s; }
double det3(const M3& a) {
    return a(0, 0) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1))
         - a(0, 1) * (a(1, 0) * a(2, 2) - a(1, 2) * a(2, 0))
         + a(0, 2) * (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0));
}
M3 inv3(const M3& a) {
    M3 c;
    c(0, 0) = a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1);
    c(0, 1) = a(0, 2) * a(2, 1) - a(0, 1) * a(2, 2);
    c(0, 2) = a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1);
    c(1, 0) = a(1, 2) * a(2, 0) - a(1, 0) * a(2, 2);
    c(1, 1) = a(0, 0) * a(2, 2) - a(0, 2) * a(2, 0);
    c(1, 2) = a(0, 2) * a(1, 0) - a(0, 0) * a(1, 2);
    c(2, 0) = a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0);
    c(2, 1) = a(0, 1) * a(2, 0) - a(0, 0) * a(2, 1);
    c(2, 2) = a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0);
    double d = a(0, 0) * c(0, 0) + a(0, 1) * c(1, 0) + a(0, 2) * c(2, 0);
    return scale(1.0 / d, c);
}
// V diag(s) W^T
M3 udvt(const M3& U, const double s[3], const M3& V) {
    M3 c;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k)
            c(r, k) = U(r, 0) * s[0] * V(k, 0) + U(r, 1) * s[1] * V(k, 1) + U(r, 2) * s[2] * V(k, 2);
    return c;
}

// ---------------------------------------------------------------------------------
// SVD with the reference sign convention (src/Math.hpp:140-161): sigma sorted by
// magnitude descending, det U = det V = +1, the sign carried by sigma[2].
// One-sided Jacobi on the columns of F: A = F V has orthogonal columns a_i = s_i u_i.
void svd3(const M3& F, double s[3], M3& U, M3& V) {
    M3 A = F;
    V = eye3();
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int r = 0; r < 3; ++r) {
                    alpha += A(r, p) * A(r, p);
                    beta += A(r, q) * A(r, q);
                    gamma += A(r, p) * A(r, q);
                }
                if (gamma == 0.0 || std::fabs(gamma) <= 1e-17 * std::sqrt(alpha * beta)) continue;
                rotated = true;
                double zeta = (beta - alpha) / (2.0 * gamma);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                double c = 1.0 / std::sqrt(1.0 + t * t);
                double sn = c * t;
                for (int r = 0; r < 3; ++r) {
                    double ap = A(r, p), aq = A(r, q);
                    A(r, p) = c * ap - sn * aq;
                    A(r, q) = sn * ap + c * aq;
                    double vp = V(r, p), vq = V(r, q);
                    V(r, p) = c * vp - sn * vq;
                    V(r, q) = sn * vp + c * vq;
                }
            }
        if (!rotated) break;
    }
    double nrm[3];
    for (int i = 0; i < 3; ++i) nrm[i] = std::sqrt(A(0, i) * A(0, i) + A(1, i) * A(1, i) + A(2, i) * A(2, i));
    int ord[3] = {0, 1, 2};
    std::sort(ord, ord + 3, [&](int a, int b) { return nrm[a] > nrm[b]; });
    M3 As, Vs;
    for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r) { As(r, j) = A(r, ord[j]); Vs(r, j) = V(r, ord[j]); }
    double n0 = nrm[ord[0]], n1 = nrm[ord[1]];
    double u0[3], u1[3], u2[3];
    if (n0 <= 0.0) {
        U = eye3();
        s[0] = s[1] = s[2] = 0.0;
        V = Vs;
        if (det3(V) < 0) { for (int r = 0; r < 3; ++r) V(r, 2) = -V(r, 2); }
        return;
    }
    for (int r = 0; r < 3; ++r) u0[r] = As(r, 0) / n0;
    // second column: Gram-Schmidt against u0; if (near) rank one, any unit vector orthogonal to u0
    double d01 = u0[0] * As(0, 1) + u0[1] * As(1, 1) + u0[2] * As(2, 1);
    for (int r = 0; r < 3; ++r) u1[r] = As(r, 1) - d01 * u0[r];
    double m1 = std::sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    if (m1 <= 1e-300 || n1 <= 1e-14 * n0) {
        int k = 0;
        if (std::fabs(u0[1]) < std::fabs(u0[k])) k = 1;
        if (std::fabs(u0[2]) < std::fabs(u0[k])) k = 2;
        double e[3] = {0, 0, 0};
        e[k] = 1.0;
        for (int r = 0; r < 3; ++r) u1[r] = e[r] - u0[k] * u0[r];
        m1 = std::sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    }
    for (int r = 0; r < 3; ++r) u1[r] /= m1;
    u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
    u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
    u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
    for (int r = 0; r < 3; ++r) { U(r, 0) = u0[r]; U(r, 1) = u1[r]; U(r, 2) = u2[r]; }
    s[0] = n0;
    s[1] = u1[0] * As(0, 1) + u1[1] * As(1, 1) + u1[2] * As(2, 1);
    s[2] = u2[0] * As(0, 2) + u2[1] * As(1, 2) + u2[2] * As(2, 2);   // signed
    V = Vs;
    if (det3(V) < 0.0) {
        for (int r = 0; r < 3; ++r) V(r, 2) = -V(r, 2);
        s[2] = -s[2];
    }
}

// Symmetric 3x3 eigenvalues (ascending) and eigenvectors (columns) by cyclic Jacobi.
void sym_eig3(const M3& Ain, double ev[3], M3* Q = nullptr) {
    M3 A = Ain, V = eye3();
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = A(0, 1) * A(0, 1) + A(0, 2) * A(0, 2) + A(1, 2) * A(1, 2);
        double dia = A(0, 0) * A(0, 0) + A(1, 1) * A(1, 1) + A(2, 2) * A(2, 2);
        if (off <= 1e-34 * dia || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double apq = A(p, q);
                if (apq == 0.0) continue;
                double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < 3; ++k) {   // A <- J^T A J
                    double akp = A(k, p), akq = A(k, q);
                    A(k, p) = c * akp - sn * akq;
                    A(k, q) = sn * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    double apk = A(p, k), aqk = A(q, k);
                    A(p, k) = c * apk - sn * aqk;
                    A(q, k) = sn * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = V(k, p), vkq = V(k, q);
                    V(k, p) = c * vkp - sn * vkq;
                    V(k, q) = sn * vkp + c * vkq;
                }
            }
    }
    int ord[3] = {0, 1, 2};
    std::sort(ord, ord + 3, [&](int a, int b) { return A(a, a) < A(b, b); });
    for (int i = 0; i < 3; ++i) ev[i] = A(ord[i], ord[i]);
    if (Q) for (int i = 0; i < 3; ++i) for (int r = 0; r < 3; ++r) (*Q)(r, i) = V(r, ord[i]);
}

// makePD (src/Math.hpp:243-261): clamp negative eigenvalues to zero.
void make_pd(M3& H) {
    double ev[3];
    M3 Q;
    sym_eig3(H, ev, &Q);
    if (ev[0] >= 0.0) return;
    for (int i = 0; i < 3; ++i) if (ev[i] < 0.0) ev[i] = 0.0; else break;
    H = udvt(Q, ev, Q);
}

// ---------------------------------------------------------------------------------
// Neo-Hookean with the quadratic-log extension below J0 (src/TetElements.hpp:50-162)
struct NH {
    double mu, lambda, J0 = 0.1;
    double fJ(double J) const { return std::log(J0) + ((J - J0) / J0) - (0.5 * std::pow(((J / J0) - 1.0), 2.0)); }
    // energy_density without the negative-sigma barrier; what the energy reports use
    // (src/TetElements.cpp:326-336, 616-630, 655-668)
    double psi(const double x[3]) const {
        double I1 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        double J = x[0] * x[1] * x[2];
        double e = 0.5 * mu * (I1 - 3.0);
        if (J > J0) {
            double lJ = std::log(J);
            e += -mu * lJ + 0.5 * lambda * lJ * lJ;
        } else {
            double f = fJ(J);
            e += -mu * f + 0.5 * lambda * f * f;
        }
        return e;
    }
    // energy_density (src/TetElements.hpp:61-77): barrier 1e15 for any negative sigma
    double energy_density(const double x[3]) const {
        if (x[0] < 0. || x[1] < 0. || x[2] < 0.) return 1.e15;
        return psi(x);
    }
    // gradient of the density (src/TetElements.hpp:93-104)
    void grad_density(const double x[3], double g[3]) const {
        double J = x[0] * x[1] * x[2];
        if (J > J0) {
            double c = lambda * std::log(J) - mu;
            for (int i = 0; i < 3; ++i) g[i] = mu * x[i] + c * (1.0 / x[i]);
        } else {
            double f = fJ(J);
            double df = 1.0 / J0 - (J - J0) / (J0 * J0);
            double c = (lambda * f - mu) * df * J;
            for (int i = 0; i < 3; ++i) g[i] = mu * x[i] + c * (1.0 / x[i]);
        }
    }
    // hessian_density (src/TetElements.hpp:109-144), then makePD
    void hessian_density(const double x[3], M3& H) const {
        double J = x[0] * x[1] * x[2];
        if (J > J0) {
            double lJ = std::log(J);
            double ct = mu - lambda * (lJ - 1.0);
            H(0, 0) = mu + ct / (x[0] * x[0]);
            H(1, 1) = mu + ct / (x[1] * x[1]);
            H(2, 2) = mu + ct / (x[2] * x[2]);
            H(0, 1) = H(1, 0) = lambda / (x[0] * x[1]);
            H(0, 2) = H(2, 0) = lambda / (x[0] * x[2]);
            H(1, 2) = H(2, 1) = lambda / (x[1] * x[2]);
        } else {
            double f = fJ(J);
            double df = 1.0 / J0 - (J - J0) / (J0 * J0);
            double lfm = lambda * f - mu;
            H(0, 0) = mu - lfm * std::pow(x[1] * x[2] / J0, 2.0) + lambda * std::pow(df * x[1] * x[2], 2.0);
            H(1, 1) = mu - lfm * std::pow(x[0] * x[2] / J0, 2.0) + lambda * std::pow(df * x[0] * x[2], 2.0);
            H(2, 2) = mu - lfm * std::pow(x[0] * x[1] / J0, 2.0) + lambda * std::pow(df * x[0] * x[1], 2.0);
            double ct = lfm * (df - J / (J0 * J0)) + lambda * J * std::pow(df, 2.0);
            H(0, 1) = H(1, 0) = x[2] * ct;
            H(0, 2) = H(2, 0) = x[1] * ct;
            H(1, 2) = H(2, 1) = x[0] * ct;
        }
        make_pd(H);
    }
};

// The prox objective: density + (w^2/vol)/2 |x - x0|^2 (src/TetElements.hpp:87-107)
struct Prox {
    NH nh;
    double x0[3];
    double wov;
    double value(const double x[3]) const {
        double d0 = x[0] - x0[0], d1 = x[1] - x0[1], d2 = x[2] - x0[2];
        return nh.energy_density(x) + (wov * 0.5) * (d0 * d0 + d1 * d1 + d2 * d2);
    }
    double gradient(const double x[3], double g[3]) const {
        nh.grad_density(x, g);
        for (int i = 0; i < 3; ++i) g[i] += wov * (x[i] - x0[i]);
        return value(x);
    }
    // src/TetElements.hpp:40-42
    static bool converged(const double a[3], const double b[3], const double g[3]) {
        double gn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
        return gn < 1e-8 || std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) < 1e-8;
    }
};

inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Cubic interpolation step (deps/mcloptlib/include/MCL/Backtracking.hpp:161-175)
double bt_cubic_interp(double fx0, double gtp, double fxa, double alpha, double fxp, double alphap) {
    double mult = 1.0 / (alpha * alpha * alphap * alphap * (alpha - alphap));
    double A00 = alphap * alphap, A01 = -alpha * alpha;
    double A10 = -alphap * alphap * alphap, A11 = alpha * alpha * alpha;
    double B0 = fxa - fx0 - alpha * gtp, B1 = fxp - fx0 - alphap * gtp;
    double r0 = (mult * A00) * B0 + (mult * A01) * B1;
    double r1 = (mult * A10) * B0 + (mult * A11) * B1;
    if (std::fabs(r0) <= 0.0) return -gtp / (2.0 * r1);
    double d = std::sqrt(r1 * r1 - 3.0 * r0 * gtp);
    return (-r1 + d) / (3.0 * r0);
}

// BacktrackingCubic::search (deps/mcloptlib/include/MCL/Backtracking.hpp:111-145) with the
// Minimizer defaults ls_max_iters = 1e5, ls_decrease = 1e-4 (Minimizer.hpp:58-62)
double bt_cubic_search(const double x[3], const double p[3], const Prox& P, double alpha0) {
    const double decrease = 1e-4;
    const int max_iters = 100000;
    if (std::sqrt(dot3(p, p)) <= DBL_EPSILON) return decrease;
    double alpha = alpha0, g[3];
    double fx0 = P.gradient(x, g);
    double gtp = dot3(g, p);
    double fxp = fx0, alphap = alpha;
    int iter = 0;
    for (; iter < max_iters; ++iter) {
        double xa[3] = {x[0] + alpha * p[0], x[1] + alpha * p[1], x[2] + alpha * p[2]};
        double fxa = P.value(xa);
        double fx0_fxa = fx0 + alpha * decrease * gtp;
        if (fxa <= fx0_fxa) break;
        double at = iter == 0 ? (gtp / (2.0 * (fx0 + gtp - fxa)))
                              : bt_cubic_interp(fx0, gtp, fxa, alpha, fxp, alphap);
        fxp = fxa;
        alphap = alpha;
        double lo = 0.1 * alpha, hi = 0.5 * alpha;
        alpha = at < lo ? lo : (at > hi ? hi : at);
    }
    if (iter >= max_iters) return -1;
    return alpha;
}

// LBFGS<double,3,8>::minimize (deps/mcloptlib/include/MCL/LBFGS.hpp:52-152)
int prox_lbfgs(const Prox& P, double x[3], int max_iters_in) {
    const int M = 8;
    double s[M][3], y[M][3], alpha[M], rho[M];
    double grad[3], q[3], grad_old[3], x_old[3], x_last[3];
    P.gradient(x, grad);
    double gamma_k = 1.0, alpha_init = 1.0;
    int global_iter = 0, max_iters = max_iters_in;
    for (int k = 0; k < max_iters; ++k) {
        for (int i = 0; i < 3; ++i) { x_old[i] = x[i]; grad_old[i] = grad[i]; q[i] = grad[i]; }
        global_iter++;
        int iter = std::min(M, k);
        for (int i = iter - 1; i >= 0; --i) {
            rho[i] = 1.0 / dot3(s[i], y[i]);
            alpha[i] = rho[i] * dot3(s[i], q);
            for (int c = 0; c < 3; ++c) q[c] = q[c] - alpha[i] * y[i][c];
        }
        for (int c = 0; c < 3; ++c) q[c] = gamma_k * q[c];
        for (int i = 0; i < iter; ++i) {
            double beta = rho[i] * dot3(q, y[i]);
            for (int c = 0; c < 3; ++c) q[c] = q[c] + (alpha[i] - beta) * s[i][c];
        }
        double dir = dot3(q, grad);
        if (dir <= 0) {
            for (int c = 0; c < 3; ++c) q[c] = grad[c];
            max_iters -= k;
            k = 0;
            double gi = std::max(std::fabs(grad[0]), std::max(std::fabs(grad[1]), std::fabs(grad[2])));
            alpha_init = std::min(1.0, 1.0 / gi);
        }
        double mq[3] = {-q[0], -q[1], -q[2]};
        double rate = bt_cubic_search(x, mq, P, alpha_init);
        if (rate <= 0) return -1;
        for (int c = 0; c < 3; ++c) { x_last[c] = x[c]; x[c] -= rate * q[c]; }
        if (Prox::converged(x_last, x, grad)) break;
        P.gradient(x, grad);
        double st[3], yt[3];
        for (int c = 0; c < 3; ++c) { st[c] = x[c] - x_old[c]; yt[c] = grad[c] - grad_old[c]; }
        if (k < M) {
            for (int c = 0; c < 3; ++c) { s[k][c] = st[c]; y[k][c] = yt[c]; }
        } else {
            for (int i = 0; i < M - 1; ++i)
                for (int c = 0; c < 3; ++c) { s[i][c] = s[i + 1][c]; y[i][c] = y[i + 1][c]; }
            for (int c = 0; c < 3; ++c) { s[M - 1][c] = st[c]; y[M - 1][c] = yt[c]; }
        }
        double denom = dot3(yt, yt);
        if (std::fabs(denom) <= 0) break;
        gamma_k = dot3(st, yt) / denom;
        alpha_init = 1.0;
    }
    return global_iter;
}

// ---------------------------------------------------------------------------------
// Sparse LDL^T (Davis 2005, "Algorithm 849") on P A P^T with reverse Cuthill-McKee P.
struct LDLT {
    int n = 0;
    std::vector<int> perm, pinv;            // new -> old, old -> new
    std::vector<int> Cp, Ci;                 // upper triangle of C = P A P^T, CSC
    std::vector<int> amap;                   // C entry -> A entry
    std::vector<int> parent, Lp, Li, Lnz;
    std::vector<double> Lx, D;
    bool ok = false;

    void analyze(int nn, const std::vector<int>& Ap, const std::vector<int>& Aj) {
        n = nn;
        // symmetric adjacency for RCM
        std::vector<std::vector<int>> adj(n);
        for (int i = 0; i < n; ++i)
            for (int p = Ap[i]; p < Ap[i + 1]; ++p) {
                int j = Aj[p];
                if (j != i) { adj[i].push_back(j); adj[j].push_back(i); }
            }
        std::vector<int> deg(n);
        for (int i = 0; i < n; ++i) deg[i] = static_cast<int>(adj[i].size());
        std::vector<char> seen(n, 0);
        std::vector<int> order;
        order.reserve(n);
        auto bfs = [&](int start, std::vector<int>& out, bool mark) {
            std::vector<char> loc(n, 0);
            std::vector<int> level;
            level.push_back(start);
            loc[start] = 1;
            size_t head = 0;
            std::vector<int> q = {start};
            while (head < q.size()) {
                int v = q[head++];
                std::vector<int> nb;
                for (int w : adj[v]) if (!loc[w] && !seen[w]) { loc[w] = 1; nb.push_back(w); }
                std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b] || (deg[a] == deg[b] && a < b); });
                for (int w : nb) q.push_back(w);
            }
            out = q;
            if (mark) for (int v : q) seen[v] = 1;
        };
        for (int root = 0; root < n; ++root) {
            if (seen[root]) continue;
            // pseudo-peripheral start: minimum degree in the component, then BFS twice
            std::vector<int> comp;
            bfs(root, comp, false);
            int start = comp[0];
            for (int v : comp) if (deg[v] < deg[start]) start = v;
            for (int pass = 0; pass < 2; ++pass) {
                std::vector<int> ord;
                bfs(start, ord, false);
                start = ord.back();
            }
            std::vector<int> ord;
            bfs(start, ord, true);
            order.insert(order.end(), ord.begin(), ord.end());
        }
        std::reverse(order.begin(), order.end());
        perm = order;
        pinv.assign(n, 0);
        for (int k = 0; k < n; ++k) pinv[perm[k]] = k;
        // upper triangle of C in CSC with map back to A entries
        std::vector<int> cnt(n + 1, 0);
        for (int i = 0; i < n; ++i)
            for (int p = Ap[i]; p < Ap[i + 1]; ++p) {
                int a = pinv[i], b = pinv[Aj[p]];
                cnt[std::max(a, b) + 1]++;
            }
        for (int k = 0; k < n; ++k) cnt[k + 1] += cnt[k];
        Cp = cnt;
        Ci.assign(Cp[n], 0);
        amap.assign(Cp[n], 0);
        std::vector<int> fill(Cp.begin(), Cp.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int p = Ap[i]; p < Ap[i + 1]; ++p) {
                int a = pinv[i], b = pinv[Aj[p]];
                int col = std::max(a, b), row = std::min(a, b);
                Ci[fill[col]] = row;
                amap[fill[col]] = p;
                fill[col]++;
            }
        // symbolic (ldl_symbolic)
        parent.assign(n, -1);
        Lnz.assign(n, 0);
        std::vector<int> flag(n, 0);
        for (int k = 0; k < n; ++k) {
            parent[k] = -1;
            flag[k] = k;
            for (int p = Cp[k]; p < Cp[k + 1]; ++p) {
                int i = Ci[p];
                if (i < k) {
                    for (; flag[i] != k; i = parent[i]) {
                        if (parent[i] == -1) parent[i] = k;
                        Lnz[i]++;
                        flag[i] = k;
                    }
                }
            }
        }
        Lp.assign(n + 1, 0);
        for (int k = 0; k < n; ++k) Lp[k + 1] = Lp[k] + Lnz[k];
        Li.assign(Lp[n], 0);
        Lx.assign(Lp[n], 0.0);
        D.assign(n, 0.0);
    }

    bool factorize(const std::vector<double>& Ax) {
        std::vector<double> Y(n, 0.0);
        std::vector<int> pattern(n), flag(n);
        for (int k = 0; k < n; ++k) {
            Y[k] = 0.0;
            int top = n;
            flag[k] = k;
            Lnz[k] = 0;
            for (int p = Cp[k]; p < Cp[k + 1]; ++p) {
                int i = Ci[p];
                Y[i] += Ax[amap[p]];
                int len = 0;
                for (; flag[i] != k; i = parent[i]) { pattern[len++] = i; flag[i] = k; }
                while (len > 0) pattern[--top] = pattern[--len];
            }
            D[k] = Y[k];
            Y[k] = 0.0;
            for (; top < n; top++) {
                int i = pattern[top];
                double yi = Y[i];
                Y[i] = 0.0;
                int p2 = Lp[i] + Lnz[i];
                for (int p = Lp[i]; p < p2; ++p) Y[Li[p]] -= Lx[p] * yi;
                double lki = yi / D[i];
                D[k] -= lki * yi;
                Li[p2] = k;
                Lx[p2] = lki;
                Lnz[i]++;
            }
            if (D[k] == 0.0) { ok = false; return false; }
        }
        ok = true;
        return true;
    }

    void solve(const double* b, double* x, int stride) const {
        std::vector<double> y(n);
        for (int k = 0; k < n; ++k) y[k] = b[stride * perm[k]];
        for (int j = 0; j < n; ++j)
            for (int p = Lp[j]; p < Lp[j + 1]; ++p) y[Li[p]] -= Lx[p] * y[j];
        for (int j = 0; j < n; ++j) y[j] /= D[j];
        for (int j = n - 1; j >= 0; --j)
            for (int p = Lp[j]; p < Lp[j + 1]; ++p) y[j] -= Lx[p] * y[Li[p]];
        for (int k = 0; k < n; ++k) x[stride * perm[k]] = y[k];
    }
};

using Clock = std::chrono::high_resolution_clock;
double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// ---------------------------------------------------------------------------------
struct Oracle {
    oracle_settings_t st;
    int n_v = 0, n_e = 0, n_f = 0, n_fix = 0;
    std::vector<double> X, X_init, X_soln;           // (n_v,3)
    std::vector<int> tets, pins;                     // sorted pins
    std::vector<int> free_idx, fix_idx;              // per vertex
    // element data (TetElements, src/TetElements.hpp:211-242)
    std::vector<M3> edges_inv, Z, Ui, cF, csymF, cU, cV, Dfix_xfix;
    std::vector<double> Di;                          // (n_e, 3, 4) row-major
    std::vector<double> DtD;                         // (n_e, 4, 4)
    std::vector<double> vol, wsq, cand_wsq, mu_e, lam_e, k_e;
    std::vector<double> last_sigma, csigma;          // (n_e, 3)
    std::vector<int> nfree;
    double curr_pe = 0.0;
    int cand_count = 0;                               // the function-static counter (src/TetElements.cpp:404)
    // global step state
    std::vector<double> x_free, x_fix, soln_free, prev_x_free;
    std::vector<int> Ap, Aj;                          // upper triangle CSR over free vertices
    std::vector<double> Ax;
    std::vector<int> a_slot;                          // (n_e, 16) slot or -1
    LDLT ldlt;
    // SplitLBFGS state (deps/mcloptlib/include/MCL/SplitLBFGS.hpp:48-62)
    std::vector<std::vector<double>> hs, hy;
    std::vector<double> lalpha, lrho, q, grad, grad_old, x_old, x_last;
    double m_tol = 1e-15;
    int m_k = 0;
    double obj_bef = 0, obj_aft = 0;
    // AlgorithmData (src/AlgorithmData.hpp:104-145)
    int iter = 0;
    double init_x_err = 1.0, last_ls_dal = 0.0, init_auglag = 0.0;
    std::vector<double> objectives, x_errors, acc_time, residuals, ls_dal, gs_dal;
    std::vector<int> reweighted, flips, inner_iters;
    double local_ms = 0, global_ms = 0, refactor_ms = 0;
    int error = 0;   // 1: NaN decrement, 2: negative prox sigma, 3: singular A, 4: line search failure

    bool nh() const { return st.model == 1; }
    const double* d(int e) const { return &Di[12 * e]; }   // column k of Di: d(e)[r*4+k]

    // TetElements ctor (src/TetElements.cpp:36-143) + create_tets_from_mesh (TetElements.hpp:250-288)
    int build(const double* mu_arr, const double* lam_arr) {
        edges_inv.resize(n_e); Z.resize(n_e); Ui.resize(n_e); cF.resize(n_e); csymF.resize(n_e);
        cU.resize(n_e); cV.resize(n_e); Dfix_xfix.assign(n_e, zero3());
        Di.assign(12 * n_e, 0.0); DtD.assign(16 * n_e, 0.0);
        vol.assign(n_e, 0); wsq.assign(n_e, 0); cand_wsq.assign(n_e, 0);
        mu_e.assign(n_e, st.mu); lam_e.assign(n_e, st.lambda); k_e.assign(n_e, 1.0);
        last_sigma.assign(3 * n_e, 1.0); csigma.assign(3 * n_e, 0.0); nfree.assign(n_e, 0);
        for (int e = 0; e < n_e; ++e) {
            const int* t = &tets[4 * e];
            M3 E;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) E(r, c) = X[3 * t[c + 1] + r] - X[3 * t[0] + r];
            edges_inv[e] = inv3(E);
            vol[e] = det3(E) / 6.0;
            if (vol[e] < 0.) return -1;
            for (int i = 0; i < 4; ++i) if (free_idx[t[i]] >= 0) nfree[e]++;
            const M3& Ei = edges_inv[e];
            double* D = &Di[12 * e];
            for (int r = 0; r < 3; ++r) {
                D[r * 4 + 0] = -Ei(0, r) - Ei(1, r) - Ei(2, r);
                for (int i = 0; i < 3; ++i) D[r * 4 + i + 1] = Ei(i, r);
            }
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    double s = 0;
                    for (int r = 0; r < 3; ++r) s += D[r * 4 + a] * D[r * 4 + b];
                    DtD[16 * e + 4 * a + b] = s;
                }
            if (mu_arr) mu_e[e] = mu_arr[e];
            if (lam_arr) lam_e[e] = lam_arr[e];
            double k = nh() ? 2.0066 * mu_e[e] + 1.0122 * lam_e[e] : 1.0;
            k_e[e] = k;
            wsq[e] = st.beta_static * k * vol[e];
        }
        return 0;
    }

    // F from full positions (src/TetElements.cpp:150-156, 366-373)
    M3 F_from_X(int e, const std::vector<double>& P) const {
        const int* t = &tets[4 * e];
        const double* D = d(e);
        M3 F;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0;
                for (int k = 0; k < 4; ++k) s += D[r * 4 + k] * P[3 * t[k] + c];
                F(c, r) = s;
            }
        return F;
    }

    // update_fix_cache (src/TetElements.cpp:547-562)
    void update_fix_cache() {
        for (int e = 0; e < n_e; ++e) {
            if (nfree[e] == 4) continue;
            const int* t = &tets[4 * e];
            const double* D = d(e);
            M3 Ft = zero3();
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    double s = 0;
                    for (int k = 0; k < 4; ++k) {
                        int j = fix_idx[t[k]];
                        if (j >= 0) s += D[r * 4 + k] * x_fix[3 * j + c];
                    }
                    Ft(r, c) = s;
                }
            Dfix_xfix[e] = Ft;
        }
    }

    // update_defo_cache (src/TetElements.cpp:564-596)
    void update_defo_cache(const std::vector<double>& xf, bool polar = true) {
        for (int e = 0; e < n_e; ++e) {
            const int* t = &tets[4 * e];
            const double* D = d(e);
            M3 Ft;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    double s = 0;
                    for (int k = 0; k < 4; ++k) {
                        int j = free_idx[t[k]];
                        if (j >= 0) s += D[r * 4 + k] * xf[3 * j + c];
                    }
                    Ft(r, c) = s;
                }
            if (nfree[e] < 4) Ft = add(Ft, Dfix_xfix[e]);
            cF[e] = tr(Ft);
        }
        if (!polar) return;
        for (int e = 0; e < n_e; ++e) {
            svd3(cF[e], &csigma[3 * e], cU[e], cV[e]);
            csymF[e] = udvt(cV[e], &csigma[3 * e], cV[e]);
        }
    }

    NH nh_of(int e) const { NH m; m.mu = mu_e[e]; m.lambda = lam_e[e]; return m; }

    // update_all_candidate_weights (src/TetElements.cpp:391-432)
    int update_all_candidate_weights(const std::vector<double>& sig) {
        if (!nh()) return -1;
        for (int e = 0; e < n_e; ++e) {
            double wsq_rest = 2.0066 * mu_e[e] + 1.0122 * lam_e[e];
            if (cand_count > st.rw_delay) {
                double wsq_min = wsq_rest * st.beta_min, wsq_max = wsq_rest * st.beta_max;
                M3 K;
                nh_of(e).hessian_density(&sig[3 * e], K);
                double ev[3];
                sym_eig3(K, ev);
                double kmax = ev[2];
                double cs = kmax;
                if (kmax < wsq_min) cs = wsq_min;
                if (kmax > wsq_max) cs = wsq_max;
                cand_wsq[e] = vol[e] * cs;
            } else {
                cand_wsq[e] = vol[e] * wsq_rest;
            }
        }
        cand_count++;
        return 0;
    }

    // update_actual_weights (src/TetElements.cpp:435-442)
    void update_actual_weights() {
        for (int e = 0; e < n_e; ++e) {
            Ui[e] = scale(wsq[e] / cand_wsq[e], Ui[e]);
            wsq[e] = cand_wsq[e];
        }
    }

    // Elements::max_wsq_ratio (src/Elements.hpp:56-69)
    double max_wsq_ratio() const {
        double mx = 0.;
        for (int e = 0; e < n_e; ++e) {
            double a = wsq[e] / cand_wsq[e], b = 1.0 / a;
            if (a > mx) mx = a;
            if (b > mx) mx = b;
        }
        return mx;
    }

    // midupdate_weights(X) / () (src/TetElements.cpp:360-388)
    void midupdate_weights_X(const std::vector<double>& P) {
        if (!st.reweighting) return;
        std::vector<double> sig(3 * n_e);
        for (int e = 0; e < n_e; ++e) {
            M3 U, V;
            svd3(F_from_X(e, P), &sig[3 * e], U, V);
        }
        update_all_candidate_weights(sig);
        update_actual_weights();
    }
    void midupdate_weights_cache() {
        if (!st.reweighting) return;
        update_all_candidate_weights(csigma);
        update_actual_weights();
    }

    // get_A_triplets / update_A_coeffs (src/TetElements.cpp:483-510)
    void build_A_pattern() {
        std::vector<long long> keys;
        keys.reserve(10 * n_e);
        for (int e = 0; e < n_e; ++e) {
            const int* t = &tets[4 * e];
            for (int i = 0; i < 4; ++i) {
                int fi = free_idx[t[i]];
                if (fi < 0) continue;
                for (int j = 0; j < 4; ++j) {
                    int fj = free_idx[t[j]];
                    if (fj >= 0 && fi <= fj) keys.push_back(static_cast<long long>(fi) * n_f + fj);
                }
            }
        }
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        Ap.assign(n_f + 1, 0);
        Aj.resize(keys.size());
        for (size_t p = 0; p < keys.size(); ++p) {
            Ap[keys[p] / n_f + 1]++;
            Aj[p] = static_cast<int>(keys[p] % n_f);
        }
        for (int i = 0; i < n_f; ++i) Ap[i + 1] += Ap[i];
        a_slot.assign(16 * n_e, -1);
        for (int e = 0; e < n_e; ++e) {
            const int* t = &tets[4 * e];
            for (int i = 0; i < 4; ++i) {
                int fi = free_idx[t[i]];
                if (fi < 0) continue;
                for (int j = 0; j < 4; ++j) {
                    int fj = free_idx[t[j]];
                    if (fj >= 0 && fi <= fj) {
                        auto it = std::lower_bound(Aj.begin() + Ap[fi], Aj.begin() + Ap[fi + 1], fj);
                        a_slot[16 * e + 4 * i + j] = static_cast<int>(it - Aj.begin());
                    }
                }
            }
        }
        Ax.assign(Aj.size(), 0.0);
    }
    void assemble_A() {
        std::fill(Ax.begin(), Ax.end(), 0.0);
        for (int e = 0; e < n_e; ++e)
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    int s = a_slot[16 * e + 4 * i + j];
                    if (s >= 0) Ax[s] += wsq[e] * DtD[16 * e + 4 * i + j];
                }
    }
    bool refactor(bool first) {
        if (first) ldlt.analyze(n_f, Ap, Aj);
        if (!ldlt.factorize(Ax)) { error = 3; return false; }
        return true;
    }
    void solve3(const std::vector<double>& b, std::vector<double>& x) {
        x.assign(3 * n_f, 0.0);
        for (int c = 0; c < 3; ++c) ldlt.solve(&b[c], &x[c], 3);
    }

    // initialize_dual_vars (src/TetElements.cpp:145-183)
    void initialize_dual_vars(const std::vector<double>& P) {
        for (int e = 0; e < n_e; ++e) {
            M3 F = F_from_X(e, P);
            if (st.rot_aware) {
                double s[3];
                M3 U, V;
                svd3(F, s, U, V);
                Z[e] = tr(udvt(V, s, V));
            } else {
                Z[e] = tr(F);
            }
            Ui[e] = zero3();
        }
        if (st.rot_aware) curr_pe = polar_potential_energy();
    }

    // polar_potential_energy (src/TetElements.cpp:638-675)
    double polar_potential_energy() const {
        double pe = 0;
        for (int e = 0; e < n_e; ++e) {
            if (nh()) {
                double s[3];
                M3 U, V;
                svd3(tr(Z[e]), s, U, V);
                pe += vol[e] * nh_of(e).psi(s);
            } else {
                pe += 0.5 * 1.0 * vol[e] * sqnorm(sub(Z[e], eye3()));
            }
        }
        return pe;
    }

    // penalty_energy(false) (src/TetElements.cpp:513-524)
    double penalty_energy() const {
        double se = 0;
        for (int e = 0; e < n_e; ++e) se += 0.5 * wsq[e] * sqnorm(sub(csymF[e], sub(Z[e], Ui[e])));
        return se;
    }

    // potential_energy_x (src/TetElements.cpp:598-635)
    double potential_energy_x() const {
        double pe = 0;
        for (int e = 0; e < n_e; ++e) {
            if (nh()) {
                pe += vol[e] * nh_of(e).psi(&csigma[3 * e]);
            } else {
                M3 R = mul(cU[e], tr(cV[e]));
                pe += 0.5 * 1.0 * vol[e] * sqnorm(sub(cF[e], R));
            }
        }
        return pe;
    }

    // inversion_count (src/TetElements.cpp:444-468); the cache always holds the SVD of F
    int inversion_count() const {
        int c = 0;
        for (int e = 0; e < n_e; ++e)
            if (csigma[3 * e] < 0. || csigma[3 * e + 1] < 0. || csigma[3 * e + 2] < 0.) c++;
        return c;
    }

    // polar_update (src/TetElements.cpp:256-356)
    double polar_update(bool update_cand) {
        double bef_pe = curr_pe;
        double pe = 0, bef = 0, aft = 0;
        std::vector<double> tsig(3 * n_e, 0.0);
        if (!nh()) {
            for (int e = 0; e < n_e; ++e) {
                M3 S = add(csymF[e], Ui[e]);
                bef += 0.5 * wsq[e] * sqnorm(sub(csymF[e], sub(Z[e], Ui[e])));
                S = scale(0.5, add(S, eye3()));
                pe += 0.5 * 1.0 * vol[e] * sqnorm(sub(csymF[e], eye3()));
                aft += 0.5 * wsq[e] * sqnorm(sub(csymF[e], sub(S, Ui[e])));
                Z[e] = S;
                Ui[e] = add(Ui[e], sub(csymF[e], S));
            }
        } else {
            std::vector<M3> S(n_e);
            for (int e = 0; e < n_e; ++e) {
                S[e] = tr(add(csymF[e], Ui[e]));
                bef += 0.5 * wsq[e] * sqnorm(sub(csymF[e], sub(Z[e], Ui[e])));
            }
            for (int e = 0; e < n_e; ++e) {
                double sh[3];
                M3 U, V;
                svd3(S[e], sh, U, V);
                Prox P;
                P.nh = nh_of(e);
                for (int i = 0; i < 3; ++i) P.x0[i] = sh[i];
                P.wov = wsq[e] / vol[e];
                double x[3] = {last_sigma[3 * e], last_sigma[3 * e + 1], last_sigma[3 * e + 2]};
                prox_lbfgs(P, x, st.prox_iters);
                for (int i = 0; i < 3; ++i) { last_sigma[3 * e + i] = x[i]; tsig[3 * e + i] = x[i]; }
                if (x[0] < 0. || x[1] < 0. || x[2] < 0.) error = 2;
                S[e] = udvt(V, x, U);
                pe += vol[e] * P.nh.psi(x);
            }
            for (int e = 0; e < n_e; ++e) {
                aft += 0.5 * wsq[e] * sqnorm(sub(csymF[e], sub(S[e], Ui[e])));
                Z[e] = S[e];
                Ui[e] = add(Ui[e], sub(csymF[e], S[e]));
            }
        }
        curr_pe = pe;
        if (update_cand) update_all_candidate_weights(tsig);
        return aft - bef + pe - bef_pe;
    }

    // update (non-polar local step, src/TetElements.cpp:193-254)
    void update_nonpolar(bool update_cand) {
        std::vector<double> tsig(3 * n_e, 0.0);
        for (int e = 0; e < n_e; ++e) {
            M3 Dix = tr(cF[e]);
            M3 Ft = add(Dix, Ui[e]);
            M3 F = tr(Ft);
            double s[3];
            M3 U, V;
            svd3(F, s, U, V);
            if (!nh()) {
                M3 R = mul(U, tr(V));
                Ft = tr(scale(0.5, add(F, R)));
            } else {
                Prox P;
                P.nh = nh_of(e);
                for (int i = 0; i < 3; ++i) P.x0[i] = s[i];
                P.wov = wsq[e] / vol[e];
                double x[3] = {last_sigma[3 * e], last_sigma[3 * e + 1], last_sigma[3 * e + 2]};
                prox_lbfgs(P, x, st.prox_iters);
                for (int i = 0; i < 3; ++i) { last_sigma[3 * e + i] = x[i]; tsig[3 * e + i] = x[i]; }
                Ft = udvt(V, x, U);
            }
            Ui[e] = add(Ui[e], sub(Dix, Ft));
            Z[e] = Ft;
        }
        if (update_cand) update_all_candidate_weights(tsig);
    }

    // add_to_b (src/TetElements.cpp:470-481)
    std::vector<double> get_b() const {
        std::vector<double> b(3 * n_f, 0.0);
        for (int e = 0; e < n_e; ++e) {
            const int* t = &tets[4 * e];
            M3 R = sub(sub(Z[e], Ui[e]), Dfix_xfix[e]);
            const double* D = d(e);
            for (int i = 0; i < 4; ++i) {
                int fi = free_idx[t[i]];
                if (fi < 0) continue;
                for (int c = 0; c < 3; ++c) {
                    double s = 0;
                    for (int r = 0; r < 3; ++r) s += D[r * 4 + i] * R(r, c);
                    b[3 * fi + c] += wsq[e] * s;
                }
            }
        }
        return b;
    }

    // global_obj_value (src/TetElements.cpp:678-688)
    double global_obj_value(const std::vector<double>& xf) {
        update_defo_cache(xf);
        double val = 0;
        for (int e = 0; e < n_e; ++e) {
            M3 P = tr(sub(Z[e], Ui[e]));
            val += (0.5 * wsq[e]) * sqnorm(sub(csymF[e], P));
        }
        return val;
    }

    // gradient from the cache (src/TetElements.cpp:742-787; :698-729 for the value)
    double grad_from_cache(std::vector<double>& g) const {
        g.assign(3 * n_f, 0.0);
        double val = 0;
        for (int e = 0; e < n_e; ++e) {
            const double* s = &csigma[3 * e];
            M3 P = tr(sub(Z[e], Ui[e]));
            M3 Q = mul(tr(cV[e]), mul(P, cV[e]));
            M3 K;
            K(0, 0) = Q(0, 0) / s[0];
            K(1, 1) = Q(1, 1) / s[1];
            K(2, 2) = Q(2, 2) / s[2];
            K(0, 1) = K(1, 0) = (Q(0, 1) + Q(1, 0)) / (s[0] + s[1]);
            K(0, 2) = K(2, 0) = (Q(0, 2) + Q(2, 0)) / (s[0] + s[2]);
            K(1, 2) = K(2, 1) = (Q(1, 2) + Q(2, 1)) / (s[1] + s[2]);
            M3 US = cU[e];
            for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) US(r, c) *= s[c];
            M3 M = sub(cF[e], mul(mul(US, K), tr(cV[e])));
            const int* t = &tets[4 * e];
            const double* D = d(e);
            for (int k = 0; k < 4; ++k) {
                int fk = free_idx[t[k]];
                if (fk < 0) continue;
                for (int c = 0; c < 3; ++c) {
                    double acc = 0;
                    for (int r = 0; r < 3; ++r) acc += D[r * 4 + k] * (wsq[e] * M(c, r));
                    g[3 * fk + c] += acc;
                }
            }
            val += (0.5 * wsq[e]) * sqnorm(sub(csymF[e], P));
        }
        return val;
    }
    double global_obj_grad(const std::vector<double>& xf, std::vector<double>& g) {
        update_defo_cache(xf);
        return grad_from_cache(g);
    }

    // Backtracking::search on n x 3 (deps/mcloptlib/include/MCL/Backtracking.hpp:71-98)
    double global_armijo(const std::vector<double>& x, const std::vector<double>& p, double alpha0) {
        const double decrease = 1e-4, tau = 0.7;
        double pn = 0;
        for (double v : p) pn += v * v;
        if (std::sqrt(pn) <= DBL_EPSILON) return decrease;
        std::vector<double> g;
        double fx0 = global_obj_grad(x, g);
        double gtp = 0;
        for (size_t i = 0; i < p.size(); ++i) gtp += g[i] * p[i];
        double alpha = alpha0;
        int it = 0;
        std::vector<double> xa(x.size());
        for (; it < 100000; ++it) {
            for (size_t i = 0; i < x.size(); ++i) xa[i] = x[i] + alpha * p[i];
            double fxa = global_obj_value(xa);
            if (fxa <= fx0 + alpha * decrease * gtp) break;
            alpha *= tau;
        }
        if (it >= 100000) return -1;
        return alpha;
    }

    // SplitLBFGS::minimize (deps/mcloptlib/include/MCL/SplitLBFGS.hpp:114-256)
    int lbfgs_minimize(std::vector<double>& x) {
        const int M = st.lbfgs_m;
        const size_t n3 = x.size();
        if (m_k == 0) {
            if (static_cast<int>(hs.size()) != M) { hs.assign(M, {}); hy.assign(M, {}); }
            lalpha.assign(M, 0.0); lrho.assign(M, 0.0);
            q.assign(n3, 0.0); grad.assign(n3, 0.0); grad_old.assign(n3, 0.0);
            x_old.assign(n3, 0.0); x_last.assign(n3, 0.0);
        }
        grad_from_cache(grad);
        double alpha_init = 1.0;
        int global_iter = 0;
        int max_iters = st.global_max_iters;
        double dg = 0;
        for (size_t i = 0; i < n3; ++i) dg += (grad[i] - grad_old[i]) * (grad[i] - grad_old[i]);
        if (dg < 1.e-13) return global_iter;
        global_iter = 1;
        bool breaking_out = false;
        for (; global_iter <= max_iters; ++global_iter) {
            x_old = x;
            grad_old = grad;
            q = grad;
            int it = std::min(M, m_k);
            for (int i = it - 1; i >= 0; --i) {
                double sy = 0, sq = 0;
                for (size_t j = 0; j < n3; ++j) sy += hs[i][j] * hy[i][j];
                lrho[i] = 1.0 / sy;
                for (size_t j = 0; j < n3; ++j) sq += hs[i][j] * q[j];
                lalpha[i] = lrho[i] * sq;
                for (size_t j = 0; j < n3; ++j) q[j] = q[j] - lalpha[i] * hy[i][j];
            }
            {
                std::vector<double> r;
                solve3(q, r);
                q.swap(r);
            }
            for (int i = 0; i < it; ++i) {
                double yq = 0;
                for (size_t j = 0; j < n3; ++j) yq += hy[i][j] * q[j];
                double beta = lrho[i] * yq;
                for (size_t j = 0; j < n3; ++j) q[j] = q[j] + hs[i][j] * (lalpha[i] - beta);
            }
            double dir = 0;
            for (size_t j = 0; j < n3; ++j) dir += grad[j] * q[j];
            if (std::isnan(dir)) { error = 1; return global_iter; }
            if (dir <= 0) {
                q = grad;
                max_iters -= m_k;
                m_k = 0;
                double gi = 0;
                for (double v : grad) gi = std::max(gi, std::fabs(v));
                alpha_init = std::min(1.0, 1.0 / gi);
            }
            double rate = 1.0;
            if (m_tol > 1.e-12 && st.global_ls == 1) {
                std::vector<double> mq(n3);
                for (size_t j = 0; j < n3; ++j) mq[j] = -q[j];
                rate = global_armijo(x, mq, alpha_init);
            }
            if (rate <= 0) { error = 4; return -1; }
            x_last = x;
            for (size_t j = 0; j < n3; ++j) x[j] -= rate * q[j];
            global_obj_grad(x, grad);
            std::vector<double> s_t(n3), y_t(n3);
            for (size_t j = 0; j < n3; ++j) { s_t[j] = x[j] - x_old[j]; y_t[j] = grad[j] - grad_old[j]; }
            if (m_k < M) {
                hs[m_k] = s_t;
                hy[m_k] = y_t;
            } else {
                hs.erase(hs.begin()); hs.push_back(s_t);
                hy.erase(hy.begin()); hy.push_back(y_t);
            }
            if (dir > 0 && (dir < st.kappa * m_tol || dir < st.kappa * 1.e-12)) {
                m_k += 1;
                x_old = x;
                x_last = x;
                breaking_out = true;
                break;
            }
            double denom = 0;
            for (size_t j = 0; j < n3; ++j) denom += y_t[j] * y_t[j];
            if (std::fabs(denom) <= 0) { breaking_out = true; break; }
            alpha_init = 1.0;
            m_k++;
        }
        m_k = std::min(M, m_k);
        return breaking_out ? global_iter : global_iter - 1;
    }

    // ---- ADMM (src/ADMM.cpp) --------------------------------------------------------
    // the element state ADMM::solve sets up (fixed positions and their cache, weights at
    // X_init, x = X_init), without the linear system: for element-level checks on meshes
    // too large for the CPU factorization
    int setup_elements() {
        x_fix.assign(3 * n_fix, 0.0);
        for (int j = 0; j < n_fix; ++j)
            for (int c = 0; c < 3; ++c) x_fix[3 * j + c] = X_init[3 * pins[j] + c];
        update_fix_cache();
        midupdate_weights_X(X_init);
        X = X_init;
        x_free.assign(3 * n_f, 0.0);
        for (int v = 0; v < n_v; ++v) {
            const int f = free_idx[v];
            if (f >= 0)
                for (int c = 0; c < 3; ++c) x_free[3 * f + c] = X[3 * v + c];
        }
        return 0;
    }

    int initialize() {
        x_fix.assign(3 * n_fix, 0.0);
        for (int j = 0; j < n_fix; ++j)
            for (int c = 0; c < 3; ++c) x_fix[3 * j + c] = X_init[3 * pins[j] + c];
        update_fix_cache();
        // GlobalStepBehavior ctor (src/GlobalStepBehavior.hpp:78-96)
        midupdate_weights_X(X_init);
        build_A_pattern();
        assemble_A();
        // RotAwareGS / UnawareGS ctor: compute (analyze + factorize)
        if (!refactor(true)) return -3;
        m_k = 0;
        X = X_init;   // Solver::initialize -> set_initial_pose (src/Solver.cpp:54)
        return 0;
    }

    double residual_norm(const std::vector<double>& a, const std::vector<double>& b) const {
        double s = 0;
        for (size_t i = 0; i < a.size(); ++i) s += (a[i] - b[i]) * (a[i] - b[i]);
        return std::sqrt(s);
    }

    int setup_solve() {
        x_free.assign(3 * n_f, 0.0);
        soln_free.assign(3 * n_f, 0.0);
        for (int v = 0; v < n_v; ++v) {
            int f = free_idx[v];
            if (f < 0) continue;
            for (int c = 0; c < 3; ++c) { x_free[3 * f + c] = X[3 * v + c]; soln_free[3 * f + c] = X_soln[3 * v + c]; }
        }
        update_fix_cache();
        update_defo_cache(x_free);
        initialize_dual_vars(X);
        prev_x_free = x_free;
        iter = 0;
        // log_initial_data (src/AlgorithmData.cpp:50-64)
        objectives.assign(1, potential_energy_x());
        flips.assign(1, inversion_count());
        x_errors.assign(1, 1.0);
        reweighted.assign(1, 1);
        acc_time.assign(1, 0.0);
        residuals.clear(); ls_dal.clear(); gs_dal.clear(); inner_iters.clear();
        init_x_err = residual_norm(prev_x_free, soln_free);
        if (st.reweighting) {
            // initialize_weights (src/GlobalStepBehavior.hpp:131-143, 223-230)
            midupdate_weights_cache();
            assemble_A();
            if (!refactor(false)) return -3;
            m_k = 0;
        }
        if (st.rot_aware) init_auglag = polar_potential_energy() + penalty_energy();
        return 0;
    }

    bool converged() const {
        bool no_flips = flips[iter] == 0 && flips[iter - 1] == 0;
        if (no_flips && residuals[iter - 1] < st.tol) return true;
        if (st.soln_provided && x_errors[iter] < 1.e-4) return true;
        return false;
    }

    int iterate() {
        prev_x_free = x_free;
        bool upd = st.reweighting != 0;
        auto t0 = Clock::now();
        if (st.rot_aware) {
            last_ls_dal = polar_update(upd);
        } else {
            update_nonpolar(upd);
        }
        local_ms += ms_since(t0);
        update_defo_cache(x_free);
        // global step (src/GlobalStepBehavior.hpp:164-184, 246-258)
        double gs_delta = 1.e8;
        if (st.rot_aware) {
            obj_bef = global_obj_value(x_free);
            t0 = Clock::now();
            m_tol = std::fabs(last_ls_dal);
            int its = lbfgs_minimize(x_free);
            global_ms += ms_since(t0);
            inner_iters.push_back(its);
            obj_aft = global_obj_value(x_free);
            gs_delta = obj_aft - obj_bef;
        } else {
            t0 = Clock::now();
            std::vector<double> b = get_b();
            solve3(b, x_free);
            update_defo_cache(x_free, false);
            global_ms += ms_since(t0);
            update_defo_cache(x_free, true);
            inner_iters.push_back(1);
        }
        // reweighting (src/GlobalStepBehavior.hpp:145-162, 232-244)
        t0 = Clock::now();
        bool rw = false;
        if (st.reweighting) {
            rw = max_wsq_ratio() > st.gamma;
            if (rw) {
                update_actual_weights();
                assemble_A();
                if (!refactor(false)) return -3;
                m_k = 0;
            }
        }
        refactor_ms += ms_since(t0);
        reweighted.push_back(rw ? 1 : 0);
        ls_dal.push_back(std::fabs(st.rot_aware ? last_ls_dal : 0.0));
        gs_dal.push_back(std::fabs(gs_delta));
        // update_per_iter_data (src/AlgorithmData.cpp:66-78)
        double obj = potential_energy_x();
        flips.push_back(inversion_count());
        x_errors.push_back(residual_norm(x_free, soln_free) / init_x_err);
        objectives.push_back(obj);
        residuals.push_back(std::fabs(objectives[iter + 1] - objectives[iter]) / (std::fabs(objectives[iter + 1]) + 1.0));
        acc_time.push_back((local_ms + global_ms + refactor_ms) / 1000.0);
        iter += 1;
        if (error) return -error;
        return converged() ? 1 : 0;
    }

    void finish() {
        for (int v = 0; v < n_v; ++v) {
            int f = free_idx[v], j = fix_idx[v];
            for (int c = 0; c < 3; ++c) X[3 * v + c] = f >= 0 ? x_free[3 * f + c] : x_fix[3 * j + c];
        }
    }
};

Oracle* O(void* h) { return static_cast<Oracle*>(h); }

}  // namespace

extern "C" {

void oracle_default_settings(oracle_settings_t* s) {
    // Settings ctor, GStep ctor, ModelSettings ctor, Lame::preset(0.45)
    s->model = 1;
    const double E = 1e7, nu = 0.45;
    s->mu = E / (2.0 * (1.0 + nu));
    s->lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    s->beta_static = 1.0; s->beta_min = 0.1; s->beta_max = 10.0;
    s->rot_aware = 1; s->reweighting = 1; s->rw_delay = 50; s->gamma = 1.5; s->tol = 1e-11;
    s->max_admm_iters = 5000; s->global_max_iters = 100; s->global_ls = 0; s->kappa = 1.0;
    s->lbfgs_m = 6; s->prox_iters = 20; s->soln_provided = 0;
}

void oracle_svd3(const double* F, double* sigma, double* U, double* V) {
    M3 f, u, v;
    std::memcpy(f.m, F, sizeof(f.m));
    svd3(f, sigma, u, v);
    std::memcpy(U, u.m, sizeof(u.m));
    std::memcpy(V, v.m, sizeof(v.m));
}

double oracle_nh_psi(const double* s, double mu, double lambda) {
    NH m; m.mu = mu; m.lambda = lambda;
    return m.energy_density(s);
}

void oracle_nh_hess(const double* s, double mu, double lambda, double* H) {
    NH m; m.mu = mu; m.lambda = lambda;
    M3 h;
    m.hessian_density(s, h);
    std::memcpy(H, h.m, sizeof(h.m));
}

double oracle_hess_lambda_max(const double* s, double mu, double lambda) {
    NH m; m.mu = mu; m.lambda = lambda;
    M3 h;
    m.hessian_density(s, h);
    double ev[3];
    sym_eig3(h, ev);
    return ev[2];
}

void oracle_sym_eig3(const double* A, double* evals) {
    M3 a;
    std::memcpy(a.m, A, sizeof(a.m));
    sym_eig3(a, evals);
}

int oracle_prox_lbfgs(const double* sig_hat, double wov, double mu, double lambda, int max_iters,
                      double* x) {
    Prox P;
    P.nh.mu = mu; P.nh.lambda = lambda;
    for (int i = 0; i < 3; ++i) P.x0[i] = sig_hat[i];
    P.wov = wov;
    return prox_lbfgs(P, x, max_iters);
}

void* oracle_create(const oracle_settings_t* s, int n_v, const double* rest, const double* init,
                    const double* soln, int n_e, const int* tets, int n_pins, const int* pins,
                    const double* mu_e, const double* lambda_e) {
    Oracle* o = new Oracle();
    o->st = *s;
    o->n_v = n_v; o->n_e = n_e;
    o->X.assign(rest, rest + 3 * n_v);
    o->X_init.assign(init, init + 3 * n_v);
    o->X_soln.assign(soln ? soln : rest, (soln ? soln : rest) + 3 * n_v);
    o->tets.assign(tets, tets + 4 * n_e);
    // binding::add_tetmesh (samples/utils/AddMeshes.hpp:89-110): sorted pins -> index maps
    o->pins.assign(pins, pins + n_pins);
    std::sort(o->pins.begin(), o->pins.end());
    o->free_idx.assign(n_v, -1);
    o->fix_idx.assign(n_v, -1);
    int j = 0, k = 0;
    for (int i = 0; i < n_v; ++i) {
        if (j < n_pins && i == o->pins[j]) { o->fix_idx[i] = j; j++; }
        else { o->free_idx[i] = k; k++; }
    }
    o->n_f = k;
    o->n_fix = j;
    if (o->build(mu_e, lambda_e) != 0) { delete o; return nullptr; }
    return o;
}

void oracle_destroy(void* h) { delete O(h); }
int oracle_initialize(void* h) { return O(h)->initialize(); }
int oracle_setup_elements(void* h) { return O(h)->setup_elements(); }
int oracle_setup_solve(void* h) { return O(h)->setup_solve(); }
int oracle_iterate(void* h) { return O(h)->iterate(); }
int oracle_finish(void* h) { O(h)->finish(); return 0; }

int oracle_solve(void* h) {
    Oracle* o = O(h);
    int r = o->setup_solve();
    if (r != 0) return r;
    for (int s = 0; s < o->st.max_admm_iters; ++s) {
        r = o->iterate();
        if (r != 0) break;
    }
    o->finish();
    return r < 0 ? r : o->iter;
}

int oracle_num_iters(void* h) { return O(h)->iter; }
int oracle_n_free(void* h) { return O(h)->n_f; }

int oracle_get_log(void* h, const char* name, double* out, int max_len) {
    Oracle* o = O(h);
    std::string n(name);
    auto cp = [&](const auto& v) {
        int m = std::min(max_len, static_cast<int>(v.size()));
        for (int i = 0; i < m; ++i) out[i] = static_cast<double>(v[i]);
        return static_cast<int>(v.size());
    };
    if (n == "objectives") return cp(o->objectives);
    if (n == "reweighted") return cp(o->reweighted);
    if (n == "inner_iters") return cp(o->inner_iters);
    if (n == "ls_dal") return cp(o->ls_dal);
    if (n == "gs_dal") return cp(o->gs_dal);
    if (n == "flips") return cp(o->flips);
    if (n == "x_errors") return cp(o->x_errors);
    if (n == "residuals") return cp(o->residuals);
    if (n == "accumulated_time_s") return cp(o->acc_time);
    return -1;
}

void oracle_get_x(void* h, double* X) { std::memcpy(X, O(h)->X.data(), sizeof(double) * O(h)->X.size()); }
void oracle_get_x_free(void* h, double* x) { std::memcpy(x, O(h)->x_free.data(), sizeof(double) * O(h)->x_free.size()); }
void oracle_set_x_free(void* h, const double* x) {
    Oracle* o = O(h);
    o->x_free.assign(x, x + 3 * o->n_f);
}

int oracle_get_elem(void* h, const char* name, double* out) {
    Oracle* o = O(h);
    std::string n(name);
    auto m3 = [&](const std::vector<M3>& v) { for (int e = 0; e < o->n_e; ++e) std::memcpy(out + 9 * e, v[e].m, 72); return 0; };
    auto sc = [&](const std::vector<double>& v) { std::memcpy(out, v.data(), sizeof(double) * v.size()); return 0; };
    if (n == "U") return m3(o->Ui);
    if (n == "Z") return m3(o->Z);
    if (n == "F") return m3(o->cF);
    if (n == "symF") return m3(o->csymF);
    if (n == "wsq") return sc(o->wsq);
    if (n == "cand_wsq") return sc(o->cand_wsq);
    if (n == "sigma_last") return sc(o->last_sigma);
    if (n == "sigma") return sc(o->csigma);
    if (n == "vol") return sc(o->vol);
    if (n == "Di") return sc(o->Di);
    if (n == "mu") return sc(o->mu_e);
    if (n == "lambda") return sc(o->lam_e);
    return -1;
}

int oracle_set_elem(void* h, const char* name, const double* in) {
    Oracle* o = O(h);
    std::string n(name);
    auto m3 = [&](std::vector<M3>& v) { for (int e = 0; e < o->n_e; ++e) std::memcpy(v[e].m, in + 9 * e, 72); return 0; };
    auto sc = [&](std::vector<double>& v) { std::memcpy(v.data(), in, sizeof(double) * v.size()); return 0; };
    if (n == "U") return m3(o->Ui);
    if (n == "Z") return m3(o->Z);
    if (n == "wsq") return sc(o->wsq);
    if (n == "cand_wsq") return sc(o->cand_wsq);
    if (n == "sigma_last") return sc(o->last_sigma);
    return -1;
}

double oracle_get_scalar(void* h, const char* name) {
    Oracle* o = O(h);
    std::string n(name);
    if (n == "curr_pe") return o->curr_pe;
    if (n == "last_ls_dal") return o->last_ls_dal;
    if (n == "cand_count") return o->cand_count;
    if (n == "m_k") return o->m_k;
    if (n == "error") return o->error;
    if (n == "init_auglag") return o->init_auglag;
    if (n == "local_ms") return o->local_ms;
    if (n == "global_ms") return o->global_ms;
    if (n == "refactor_ms") return o->refactor_ms;
    return NAN;
}

void oracle_set_scalar(void* h, const char* name, double v) {
    Oracle* o = O(h);
    std::string n(name);
    if (n == "curr_pe") o->curr_pe = v;
    else if (n == "cand_count") o->cand_count = static_cast<int>(v);
}

double oracle_local_step(void* h, int update_cand) { return O(h)->polar_update(update_cand != 0); }
void oracle_update_defo_cache(void* h) { O(h)->update_defo_cache(O(h)->x_free); }

double oracle_global_value(void* h, const double* x_free) {
    Oracle* o = O(h);
    std::vector<double> x(x_free, x_free + 3 * o->n_f);
    return o->global_obj_value(x);
}

double oracle_global_grad(void* h, const double* x_free, double* grad) {
    Oracle* o = O(h);
    std::vector<double> x(x_free, x_free + 3 * o->n_f), g;
    double v = o->global_obj_grad(x, g);
    std::memcpy(grad, g.data(), sizeof(double) * g.size());
    return v;
}

double oracle_potential_energy(void* h) { return O(h)->potential_energy_x(); }
int oracle_inversion_count(void* h) { return O(h)->inversion_count(); }

int oracle_get_A(void* h, int* rows, int* cols, double* vals, int max_nnz) {
    Oracle* o = O(h);
    int nnz = static_cast<int>(o->Aj.size());
    if (rows) {
        int m = std::min(nnz, max_nnz);
        int p = 0;
        for (int i = 0; i < o->n_f && p < m; ++i)
            for (int q = o->Ap[i]; q < o->Ap[i + 1] && p < m; ++q, ++p) {
                rows[p] = i; cols[p] = o->Aj[q]; vals[p] = o->Ax[q];
            }
    }
    return nnz;
}

int oracle_solve3(void* h, const double* b, double* x) {
    Oracle* o = O(h);
    std::vector<double> bb(b, b + 3 * o->n_f), xx;
    o->solve3(bb, xx);
    std::memcpy(x, xx.data(), sizeof(double) * xx.size());
    return 0;
}

}  // extern "C"
