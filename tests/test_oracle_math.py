"""Pins of the oracle against what the mathematics fixes (independent of the reference):
SVD against LAPACK, the NH density against finite differences, the prox KKT condition,
the sparse factorization against a dense solve, A against a brute-force dense assembly,
the global gradient against finite differences of the global objective, the ARAP local
step against its closed form, and the SplitLBFGS two-loop against the dense BFGS matrix.
"""
import numpy as np
import pytest

import oracle_lib as ol
from wrapd_amd import meshgen

rng = np.random.default_rng(7)


def lame(E=1e6, nu=0.45):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


@pytest.mark.parametrize("i", range(40))
def test_svd_against_lapack(i):
    F = rng.standard_normal((3, 3)) if i % 2 else np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    if i % 5 == 0:
        F[:, 0] *= -1
    s, U, V = ol.svd3(F)
    ref = np.linalg.svd(F, compute_uv=False)
    assert np.allclose(np.abs(s), ref, rtol=1e-13, atol=1e-14)
    assert s[0] >= s[1] >= abs(s[2])
    assert np.isclose(np.linalg.det(U), 1.0) and np.isclose(np.linalg.det(V), 1.0)
    assert np.sign(s[2]) == np.sign(np.linalg.det(F))
    assert np.allclose(U @ np.diag(s) @ V.T, F, atol=1e-13)
    assert np.allclose(U.T @ U, np.eye(3), atol=1e-13)


def _psi_fd_grad(s, mu, lam, h=1e-6):
    g = np.zeros(3)
    for i in range(3):
        e = np.zeros(3)
        e[i] = h
        g[i] = (ol.nh_psi(s + e, mu, lam) - ol.nh_psi(s - e, mu, lam)) / (2 * h)
    return g


@pytest.mark.parametrize("branch", ["log", "extension"])
def test_nh_gradient_and_hessian_finite_differences(branch):
    mu, lam = lame()
    for _ in range(20):
        s = rng.uniform(0.7, 1.4, 3) if branch == "log" else rng.uniform(0.2, 0.4, 3)
        # gradient through the prox with w = 0 at x0 = s: grad = psi'(s)
        x, _ = ol.prox_lbfgs(s, 0.0, mu, lam, s.copy(), max_iters=0)
        g_fd = _psi_fd_grad(s, mu, lam)
        # the prox objective's gradient at s with w=0 is psi'(s); check via the density Hessian
        H = ol.nh_hess(s, mu, lam)
        H_fd = np.zeros((3, 3))
        h = 1e-6
        for j in range(3):
            e = np.zeros(3)
            e[j] = h
            H_fd[:, j] = (_psi_fd_grad(s + e, mu, lam, 1e-5) - _psi_fd_grad(s - e, mu, lam, 1e-5)) / (2 * h)
        w, Q = np.linalg.eigh(0.5 * (H_fd + H_fd.T))
        H_pd = Q @ np.diag(np.maximum(w, 0)) @ Q.T   # makePD (src/Math.hpp:243-261)
        assert np.allclose(H, H_pd, rtol=2e-4, atol=2e-4 * np.abs(H).max())
        assert np.all(np.isfinite(g_fd))


def test_prox_kkt_residual():
    """sigma* minimises psi + w/2 |s - s_hat|^2: psi'(s*) + w (s* - s_hat) = 0."""
    for _ in range(50):
        E = 10 ** rng.uniform(5, 7)
        mu, lam = lame(E)
        k = 2.0066 * mu + 1.0122 * lam
        s_hat = np.sort(rng.uniform(0.6, 1.6, 3))[::-1]
        wov = k * 10 ** rng.uniform(-1, 1)
        x, it = ol.prox_lbfgs(s_hat, wov, mu, lam, np.ones(3), max_iters=200)
        g = _psi_fd_grad(x, mu, lam, 1e-7) + wov * (x - s_hat)
        assert np.linalg.norm(g) <= 1e-5 * (mu + wov) * np.linalg.norm(x)


def test_candidate_weight_eigenvalue():
    mu, lam = lame()
    for _ in range(50):
        s = rng.uniform(0.3, 1.8, 3)
        H = ol.nh_hess(s, mu, lam)
        assert np.isclose(ol.hess_lambda_max(s, mu, lam), np.linalg.eigvalsh(H).max(), rtol=1e-12)


def _tiny():
    return meshgen.small_beam(4, 2, 2, seed=5, twist_deg=60.0)


def test_A_brute_force_and_solve():
    prob = _tiny()
    o = ol.Oracle(prob)
    o.initialize()
    r, c, v = o.A_coo()
    n = o.n_f
    A = np.zeros((n, n))
    A[r, c] = v
    A = A + np.triu(A, 1).T
    # brute force: sum_e w^2 D_e^T D_e over free vertices
    free = np.full(prob.n_verts, -1)
    fr = np.setdiff1d(np.arange(prob.n_verts), prob.pins)
    free[fr] = np.arange(len(fr))
    Di, wsq = o.elem("Di"), o.elem("wsq")
    B = np.zeros((n, n))
    for e, t in enumerate(prob.tets):
        DtD = Di[e].T @ Di[e]
        for a in range(4):
            for b in range(4):
                if free[t[a]] >= 0 and free[t[b]] >= 0:
                    B[free[t[a]], free[t[b]]] += wsq[e] * DtD[a, b]
    assert np.allclose(A, B, rtol=1e-13, atol=1e-12 * np.abs(B).max())
    assert np.all(np.linalg.eigvalsh(A) > 0)
    rhs = rng.standard_normal((n, 3))
    assert np.allclose(o.solve3(rhs), np.linalg.solve(A, rhs), rtol=1e-10, atol=1e-12)


def test_global_gradient_finite_differences():
    prob = _tiny()
    o = ol.Oracle(prob, tol=0.0)
    o.initialize()
    o.setup_solve()
    o.iterate()
    o.local_step(update_cand=False)
    x = o.x_free()
    v0, g = o.global_grad(x)
    d = rng.standard_normal(x.shape)
    h = 1e-6
    fd = (o.global_value(x + h * d) - o.global_value(x - h * d)) / (2 * h)
    assert abs(fd - np.sum(g * d)) <= 1e-6 * max(1.0, abs(fd))


def test_arap_local_step_closed_form():
    prob = meshgen.small_beam(4, 2, 2, seed=6, twist_deg=60.0, model="arap", reweighting=0)
    o = ol.Oracle(prob, tol=0.0, model=0, reweighting=0)
    o.initialize()
    o.setup_solve()
    o.iterate()
    S, U = o.elem("symF"), o.elem("U")
    o.local_step(update_cand=False)
    Z = o.elem("Z")
    assert np.allclose(Z, 0.5 * (S + U + np.eye(3)), atol=1e-14)
    assert np.allclose(o.elem("U"), U + S - Z, atol=1e-14)


def test_two_loop_equals_dense_bfgs():
    """The L-BFGS two-loop with H0 applied is the dense BFGS inverse update applied to g."""
    n = 6
    H0 = np.linalg.inv(np.diag(rng.uniform(1, 3, n)))
    S = [rng.standard_normal(n) for _ in range(4)]
    Y = [s + 0.1 * rng.standard_normal(n) for s in S]
    g = rng.standard_normal(n)
    # two-loop (deps/mcloptlib/include/MCL/SplitLBFGS.hpp:152-177)
    q = g.copy()
    alpha = [0.0] * 4
    for i in range(3, -1, -1):
        alpha[i] = (S[i] @ q) / (S[i] @ Y[i])
        q = q - alpha[i] * Y[i]
    q = H0 @ q
    for i in range(4):
        beta = (Y[i] @ q) / (S[i] @ Y[i])
        q = q + S[i] * (alpha[i] - beta)
    # dense: H_{k+1} = (I - rho s y^T) H_k (I - rho y s^T) + rho s s^T
    H = H0.copy()
    for s, y in zip(S, Y):
        rho = 1.0 / (s @ y)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    assert np.allclose(q, H @ g, rtol=1e-12, atol=1e-12)
    # the Gram-matrix form the device uses (wrapd_amd/csrc/wd_vec.hip) gives the same vector
    G = np.array([[S[i] @ Y[j] for j in range(4)] for i in range(4)])
    a = np.zeros(4)
    for i in range(3, -1, -1):
        a[i] = (S[i] @ g - sum(a[j] * G[i, j] for j in range(i + 1, 4))) / G[i, i]
    r = H0 @ (g - sum(a[j] * Y[j] for j in range(4)))
    c = np.zeros(4)
    for i in range(4):
        b = (Y[i] @ r + sum(c[j] * G[j, i] for j in range(i))) / G[i, i]
        c[i] = a[i] - b
    r = r + sum(c[j] * S[j] for j in range(4))
    assert np.allclose(r, H @ g, rtol=1e-12, atol=1e-12)


def test_known_answer_rest_is_fixed_point():
    """Starting at rest with pins at rest, energy stays zero and positions stay put."""
    prob = meshgen.small_beam(6, 2, 2, seed=8, twist_deg=0.0)
    o = ol.Oracle(prob, tol=0.0)
    o.run(5)
    # psi(1,1,1) = 0; round-off in F leaves mu * vol * O(1e-15) per element
    assert np.max(np.abs(o.log("objectives"))) < 1e-9
    assert np.allclose(o.x(), prob.rest, atol=1e-12)
