// 3x3 linear algebra for one element per lane (host+device, float or double).
//
// Matrices are row-major m[3*r+c] in registers.  The SVD follows the sign
// convention of the reference's math::svd (src/Math.hpp:140-161): singular values
// sorted by magnitude in descending order, det U = det V = +1, and the sign of
// det F carried by s[2].  Algorithm: cyclic Jacobi eigen-decomposition of F^T F
// gives V (a product of rotations), then a Givens QR of F V gives U and the signed
// singular values on the diagonal of R (McAdams, Selle, Tamstorf, Teran & Sifakis,
// "Computing the Singular Value Decomposition of 3x3 matrices with minimal
// branching and elementary floating point operations", 2011), with exact rather
// than approximate rotations.
#pragma once

#include <hip/hip_runtime.h>

#define WD_HD __host__ __device__ __forceinline__

namespace wd {

template <typename T>
struct M3 {
    T m[9];
    WD_HD T& operator()(int r, int c) { return m[3 * r + c]; }
    WD_HD T operator()(int r, int c) const { return m[3 * r + c]; }
};

template <typename T> WD_HD T wsqrt(T x) { return sqrt(x); }
template <typename T> WD_HD T wlog(T x) { return log(x); }
template <typename T> WD_HD T wabs(T x) { return x < T(0) ? -x : x; }
template <typename T> WD_HD T wmax(T a, T b) { return a > b ? a : b; }

template <typename T>
WD_HD M3<T> eye3() {
    M3<T> a;
#pragma unroll
    for (int i = 0; i < 9; ++i) a.m[i] = T(0);
    a.m[0] = a.m[4] = a.m[8] = T(1);
    return a;
}

template <typename T>
WD_HD M3<T> transpose(const M3<T>& a) {
    M3<T> b;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) b(r, c) = a(c, r);
    return b;
}

template <typename T>
WD_HD M3<T> matmul(const M3<T>& a, const M3<T>& b) {
    M3<T> c;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) c(r, k) = a(r, 0) * b(0, k) + a(r, 1) * b(1, k) + a(r, 2) * b(2, k);
    return c;
}

// A^T B
template <typename T>
WD_HD M3<T> matmul_tn(const M3<T>& a, const M3<T>& b) {
    M3<T> c;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) c(r, k) = a(0, r) * b(0, k) + a(1, r) * b(1, k) + a(2, r) * b(2, k);
    return c;
}

// A diag(s) B^T
template <typename T>
WD_HD M3<T> udvt(const M3<T>& a, const T s[3], const M3<T>& b) {
    M3<T> c;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            c(r, k) = a(r, 0) * s[0] * b(k, 0) + a(r, 1) * s[1] * b(k, 1) + a(r, 2) * s[2] * b(k, 2);
    return c;
}

template <typename T>
WD_HD T sqnorm9(const M3<T>& a) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) s += a.m[i] * a.m[i];
    return s;
}

template <typename T>
WD_HD T det3(const M3<T>& a) {
    return a(0, 0) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) - a(0, 1) * (a(1, 0) * a(2, 2) - a(1, 2) * a(2, 0)) +
           a(0, 2) * (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0));
}

// Symmetric Jacobi rotation of the (p,q) plane on S (full symmetric storage) and
// accumulation into the columns of V.
template <typename T, int P, int Q>
WD_HD void jacobi_rotate(T S[3][3], M3<T>& V) {
    const T apq = S[P][Q];
    if (apq == T(0)) return;
    const T theta = (S[Q][Q] - S[P][P]) / (T(2) * apq);
    const T t = (theta >= T(0) ? T(1) : T(-1)) / (wabs(theta) + wsqrt(theta * theta + T(1)));
    const T c = T(1) / wsqrt(t * t + T(1));
    const T s = t * c;
    constexpr int R = 3 - P - Q;   // the remaining index
    const T spp = S[P][P], sqq = S[Q][Q], srp = S[R][P], srq = S[R][Q];
    S[P][P] = spp - t * apq;
    S[Q][Q] = sqq + t * apq;
    S[P][Q] = S[Q][P] = T(0);
    S[R][P] = S[P][R] = c * srp - s * srq;
    S[R][Q] = S[Q][R] = s * srp + c * srq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T vp = V(k, P), vq = V(k, Q);
        V(k, P) = c * vp - s * vq;
        V(k, Q) = s * vp + c * vq;
    }
}

template <typename T> struct JacobiSweeps;
template <> struct JacobiSweeps<double> { static constexpr int max = 12; };
template <> struct JacobiSweeps<float> { static constexpr int max = 8; };

template <typename T> struct Eps;
template <> struct Eps<double> { static constexpr double v = 2.220446049250313e-16; };
template <> struct Eps<float> { static constexpr float v = 1.1920929e-07f; };

// Givens rotation (c, s) with c*a + s*b = r >= 0 and -s*a + c*b = 0.
template <typename T>
WD_HD void givens(T a, T b, T& c, T& s) {
    const T r = wsqrt(a * a + b * b);
    if (r == T(0)) { c = T(1); s = T(0); return; }
    c = a / r;
    s = b / r;
}

template <typename T>
WD_HD void svd3(const M3<T>& F, T sig[3], M3<T>& U, M3<T>& V) {
    // S = F^T F
    T S[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = r; c < 3; ++c) {
            const T v = F(0, r) * F(0, c) + F(1, r) * F(1, c) + F(2, r) * F(2, c);
            S[r][c] = v;
            S[c][r] = v;
        }
    V = eye3<T>();
    const T scale2 = S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2];
    for (int sweep = 0; sweep < JacobiSweeps<T>::max; ++sweep) {
        const T off = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
        if (off <= Eps<T>::v * Eps<T>::v * T(1e-4) * scale2 || off == T(0)) break;
        jacobi_rotate<T, 0, 1>(S, V);
        jacobi_rotate<T, 0, 2>(S, V);
        jacobi_rotate<T, 1, 2>(S, V);
    }
    // sort eigenvalues descending; swaps negate one column so det V stays +1
    T e0 = S[0][0], e1 = S[1][1], e2 = S[2][2];
    auto swapcols = [&](int a, int b) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const T t = V(k, a);
            V(k, a) = -V(k, b);
            V(k, b) = t;
        }
    };
    if (e0 < e1) { T t = e0; e0 = e1; e1 = t; swapcols(0, 1); }
    if (e0 < e2) { T t = e0; e0 = e2; e2 = t; swapcols(0, 2); }
    if (e1 < e2) { T t = e1; e1 = e2; e2 = t; swapcols(1, 2); }
    // B = F V, then QR by Givens rotations: U^T B = R
    M3<T> B = matmul(F, V);
    U = eye3<T>();
    T c, s;
    // zero B(1,0) with rows (0,1)
    givens(B(0, 0), B(1, 0), c, s);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T b0 = B(0, k), b1 = B(1, k);
        B(0, k) = c * b0 + s * b1;
        B(1, k) = -s * b0 + c * b1;
        const T u0 = U(k, 0), u1 = U(k, 1);
        U(k, 0) = c * u0 + s * u1;
        U(k, 1) = -s * u0 + c * u1;
    }
    // zero B(2,0) with rows (0,2)
    givens(B(0, 0), B(2, 0), c, s);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T b0 = B(0, k), b2 = B(2, k);
        B(0, k) = c * b0 + s * b2;
        B(2, k) = -s * b0 + c * b2;
        const T u0 = U(k, 0), u2 = U(k, 2);
        U(k, 0) = c * u0 + s * u2;
        U(k, 2) = -s * u0 + c * u2;
    }
    // zero B(2,1) with rows (1,2)
    givens(B(1, 1), B(2, 1), c, s);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T b1 = B(1, k), b2 = B(2, k);
        B(1, k) = c * b1 + s * b2;
        B(2, k) = -s * b1 + c * b2;
        const T u1 = U(k, 1), u2 = U(k, 2);
        U(k, 1) = c * u1 + s * u2;
        U(k, 2) = -s * u1 + c * u2;
    }
    sig[0] = B(0, 0);
    sig[1] = B(1, 1);
    sig[2] = B(2, 2);
}

// Largest eigenvalue of a symmetric 3x3 matrix (trigonometric closed form, Smith 1961).
template <typename T>
WD_HD T sym3_lambda_max(const M3<T>& A) {
    const T p1 = A(0, 1) * A(0, 1) + A(0, 2) * A(0, 2) + A(1, 2) * A(1, 2);
    const T q = (A(0, 0) + A(1, 1) + A(2, 2)) / T(3);
    const T d0 = A(0, 0) - q, d1 = A(1, 1) - q, d2 = A(2, 2) - q;
    const T p2 = d0 * d0 + d1 * d1 + d2 * d2 + T(2) * p1;
    if (p2 <= T(0)) return q;
    const T p = wsqrt(p2 / T(6));
    M3<T> B = A;
    B(0, 0) = d0; B(1, 1) = d1; B(2, 2) = d2;
    const T ip = T(1) / p;
#pragma unroll
    for (int i = 0; i < 9; ++i) B.m[i] *= ip;
    T r = det3(B) / T(2);
    r = r < T(-1) ? T(-1) : (r > T(1) ? T(1) : r);
    const T phi = acos(r) / T(3);
    return q + T(2) * p * cos(phi);
}

}  // namespace wd
