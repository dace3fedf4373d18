"""The learned prox warm start on the GPU (wrapd_amd/csrc/wd_mlp.hip, k_local_learned).

Inference and the forward/backward gradient against the CPU restatement
(oracle/mlp_oracle.py, same BF16 rounding points, FP32 accumulation); training on tuples
the solver logs from its own run; the trained network's gain: fewer Newton iterations
in a local step started from it, the same prox solutions, and a learned + FP32 run that
ends at the FP64 fixed point.
"""
import os
import sys

import numpy as np
import pytest

from wrapd_amd import meshgen
from wrapd_amd.solver import Device, Solver

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(11)


def _net():
    p = np.zeros(mo.N_PARAMS, np.float32)
    p[0:384] = rng.uniform(-0.29, 0.29, 384)
    p[448:4544] = rng.uniform(-0.21, 0.21, 4096)
    p[4608:4800] = rng.uniform(-0.3, 0.3, 192)
    p[384:448] = rng.normal(0, 0.1, 64)
    p[4544:4608] = rng.normal(0, 0.1, 64)
    p[4800:4803] = rng.normal(0, 0.1, 3)
    norm = np.array([1.2, 1.0, 0.8, 2.0, 2.2, 13.0, 0.3, 0.2, 0.2, 1.0, 0.5, 1.0, 0.0, 0.0, 0.0, 0.05, 0.05, 0.05],
                    np.float32)
    return p, norm


def _tuples(n):
    s = np.sort(rng.uniform(0.5, 1.6, (n, 3)), axis=1)[:, ::-1]
    mu = 10 ** rng.uniform(4.5, 6.5, n)
    lam = mu * rng.uniform(2.0, 12.0, n)
    wov = mu * 10 ** rng.uniform(-1, 1.5, n)
    x = s + rng.normal(0, 0.05, (n, 3))
    return np.column_stack([s, wov, mu, lam, x]).astype(np.float32)


def test_inference_matches_oracle():
    p, norm = _net()
    d = Device()
    d.mlp_set_params(p, norm)
    t = _tuples(1000)   # not a multiple of 64: the last wave is partly idle
    out = d.mlp_infer(t[:, :6])
    ref = mo.predict(p, norm, t[:, :6])
    dev = np.abs(out - ref) / norm[15:18]   # in standardised output units
    print(f"inference vs oracle: max {dev.max():.2e}, median {np.median(dev):.2e}")
    assert np.median(dev) < 1e-4
    assert dev.max() < 2e-2   # a hidden activation may round to the other BF16 neighbour


def test_gradient_matches_oracle():
    p, norm = _net()
    d = Device()
    d.mlp_set_params(p, norm)
    t = _tuples(777)
    loss, g = d.mlp_grad(t)
    loss_o, g_o = mo.loss_grad(p, norm, t, bf=True)
    _, g_x = mo.loss_grad(p, norm, t, bf=False)
    print(f"loss {loss:.6e} oracle {loss_o:.6e}; |g - g_oracle|/|g| "
          f"{np.linalg.norm(g - g_o) / np.linalg.norm(g_o):.2e}, vs exact {np.linalg.norm(g - g_x) / np.linalg.norm(g_x):.2e}")
    assert abs(loss - loss_o) <= 1e-3 * loss_o
    assert np.linalg.norm(g - g_o) <= 1e-2 * np.linalg.norm(g_o)
    for name, (a, b) in dict(W1=(0, 384), b1=(384, 448), W2=(448, 4544), b2=(4544, 4608), W3=(4608, 4800),
                             b3=(4800, 4803)).items():
        assert np.linalg.norm(g[a:b] - g_o[a:b]) <= 2e-2 * np.linalg.norm(g_o[a:b]) + 1e-7, name
    assert np.linalg.norm(g - g_x) <= 5e-2 * np.linalg.norm(g_x)


def _prob_a():
    return meshgen.small_beam(20, 4, 4, seed=21, twist_deg=150.0, axial=-0.2, youngs=1e6)


def _prob_b():
    return meshgen.small_beam(18, 4, 4, seed=22, twist_deg=120.0, axial=-0.15, youngs=1e6)


@pytest.fixture(scope="module")
def trained():
    prob = _prob_a()
    s = Solver(prob, tol=0.0)
    d = s.device()
    d.mlp_log_start(prob.n_elems * 60)
    s.run(60)
    n = d.mlp_log_count()
    assert n == prob.n_elems * 60
    d.mlp_log_stop()
    t = d.mlp_log_get(0, n)
    # logged tuples are prox solutions: psi'(x) + w^2/vol (x - s_hat) = 0 (mu-scaled)
    sub = t[:: max(1, n // 2000)].astype(np.float64)
    x, s_hat, wov, mu, lam = sub[:, 6:9], sub[:, 0:3], sub[:, 3], sub[:, 4], sub[:, 5]
    J = np.prod(x, axis=1)
    g = mu[:, None] * x + ((lam * np.log(J) - mu)[:, None]) / x + wov[:, None] * (x - s_hat)
    assert np.median(np.linalg.norm(g, axis=1) / (mu + wov)) < 1e-6
    d.mlp_init(seed=5)
    d.mlp_augment(seed=7)     # w^2/vol redrawn over the logged range, targets re-solved
    d.mlp_fit_norm()
    hist = d.mlp_train(600, batch=4096, lr=2e-3, seed=9)
    params, norm = d.mlp_get_params()
    s.close()
    return dict(params=params, norm=norm, hist=hist, tuples=t)


def test_training_reduces_loss(trained):
    h = trained["hist"]
    print(f"training loss {h[0]:.3e} -> {h[-20:].mean():.3e}")
    assert np.all(np.isfinite(h))
    assert h[-20:].mean() < 0.1 * h[:5].mean()


def test_learned_start_saves_newton_iterations(trained):
    prob = _prob_b()
    s = Solver(prob, tol=0.0, prox_mode=1)
    s.initialize()
    s.setup()
    d = s.device()
    snap = {k: d.elem(k) for k in ("U", "P", "sigma_last", "cand_wsq")}

    def step(prox):
        for k, v in snap.items():
            d.set_elem(k, v)
        o = d.local_step(precision="fp64", prox=prox, prox_iters=30, update_cand=False, newton_rtol=1e-12)
        return o, d.elem("sigma_last"), d.elem("U")

    o_n, sig_n, U_n = step("newton")
    d.mlp_set_params(trained["params"], trained["norm"])
    o_l, sig_l, U_l = step("learned")
    print(f"first local step of a new beam: Newton iterations {o_n['prox_iters_total']:.0f} from the previous "
          f"solution, {o_l['prox_iters_total']:.0f} with the learned start (used for {o_l['learned_used']:.0f} of "
          f"{prob.n_elems} elements, accepted without Newton {o_l['learned_accepted']:.0f})")
    assert o_l["prox_iters_total"] < 0.8 * o_n["prox_iters_total"]
    assert np.max(np.abs(sig_l - sig_n)) < 1e-9
    assert np.linalg.norm(U_l - U_n) <= 1e-9 * np.linalg.norm(U_n)
    s.close()


def test_learned_mixed_run_reaches_fp64_fixed_point(trained):
    prob = _prob_b()
    ref = Solver(prob, tol=1e-12)
    n_ref = ref.run(3000)
    fast = Solver(prob, tol=1e-12, precision=1, prox_mode=2, fp64_switch=1e-8)
    fast.device().mlp_set_params(trained["params"], trained["norm"])
    n_fast = fast.run(3000)
    assert n_ref < 3000 and n_fast < 3000
    xr, xf = ref.x(), fast.x()
    err = np.linalg.norm(xf - xr) / np.linalg.norm(xr)
    print(f"FP64 {n_ref} iterations, learned + FP32 {n_fast}: x rel err {err:.2e}")
    assert err <= 1e-6


def test_adam_steps_match_oracle():
    """Three full-batch Adam steps on the GPU (k_mlp_adam after the forward/backward)
    against oracle/mlp_oracle.py's adam applied to the same gradients: each step's gradient
    is read first with wrapd_mlp_grad at the current parameters (the same kernels the step
    runs, so the same bits), and the oracle carries the moments in float64."""
    p, norm = _net()
    d = Device()
    d.mlp_set_params(p, norm)
    t = _tuples(3000)
    lr = 1e-2
    pc = p.astype(np.float64)
    m1 = np.zeros_like(pc)
    m2 = np.zeros_like(pc)
    for step in range(1, 4):
        _, g = d.mlp_grad(t)                       # the gradient this step will use
        pc, m1, m2 = mo.adam(pc, g.astype(np.float64), m1, m2, step, lr)
        d.mlp_train_full(t, 1, lr=lr)
        pg, _ = d.mlp_get_params()
        dev = np.abs(pg - pc)
        print(f"step {step}: max |p_gpu - p_oracle| {dev.max():.2e} (update ~{lr:g})")
        # float32 Adam: the update's relative rounding is a few 1e-7, the parameters round to
        # float32 (|p| < 1) each step
        assert dev.max() <= 2e-6 * lr * step + 3e-7
