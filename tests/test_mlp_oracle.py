"""Pins of the MLP restatement (oracle/mlp_oracle.py), without a GPU.

The float64 gradient of the loss is checked against central finite differences, the
BF16-rounded (device) gradient against the float64 one to BF16 precision, GELU against
its closed form, the BF16 rounding against exhaustive float32 -> bfloat16 known values,
and a few Adam steps on a quadratic against the update's closed form.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

rng = np.random.default_rng(3)


def _net(scale=0.3):
    p = rng.standard_normal(mo.N_PARAMS) * scale
    norm = np.concatenate([[1.2, 1.0, 0.8, 2.0, 2.2, 13.0], [0.3, 0.2, 0.2, 1.0, 0.5, 1.0], [0.0, 0.0, 0.0],
                           [0.05, 0.05, 0.05]])
    return p, norm


def _tuples(n):
    s = np.sort(rng.uniform(0.5, 1.6, (n, 3)), axis=1)[:, ::-1]
    mu = 10 ** rng.uniform(4.5, 6.5, n)
    lam = mu * rng.uniform(2.0, 12.0, n)
    wov = mu * 10 ** rng.uniform(-1, 1.5, n)
    x = s + rng.normal(0, 0.05, (n, 3))
    return np.column_stack([s, wov, mu, lam, x])


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.0, 3.1415927, 1e-3], np.float32)
    got = mo.bf16(vals)
    # 1.00390625 is the half-way point between 1.0 and 1.0078125: ties to even -> 1.0;
    # 1.005859375 is above the half-way point -> 1.0078125
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == np.float32(1.0078125) and got[3] == np.float32(1.0078125)
    assert got[4] == -2.0
    assert abs(got[5] - 3.140625) < 1e-7
    assert abs(got[6] / 1e-3 - 1) < 2 ** -8


def test_gelu_closed_form():
    x = np.linspace(-4, 4, 81)
    assert np.allclose(mo.gelu(x), 0.5 * x * (1 + np.tanh(np.sqrt(2 / np.pi) * (x + 0.044715 * x ** 3))))
    h = 1e-6
    assert np.allclose(mo.gelu_d(x), (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h), atol=1e-8)


def test_float64_gradient_matches_finite_differences():
    p, norm = _net()
    t = _tuples(40)
    _, g = mo.loss_grad(p, norm, t, bf=False)
    idx = rng.choice(mo.N_PARAMS, 60, replace=False)
    for name in ("W1", "b1", "W2", "b2", "W3", "b3"):   # at least one of every block
        idx = np.append(idx, mo.OFF[name])
    h = 1e-6
    for i in idx:
        e = np.zeros(mo.N_PARAMS)
        e[i] = h
        fd = (mo.loss_grad(p + e, norm, t, bf=False)[0] - mo.loss_grad(p - e, norm, t, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_gradient_close_to_exact():
    p, norm = _net()
    t = _tuples(256)
    l64, g64 = mo.loss_grad(p, norm, t, bf=False)
    l16, g16 = mo.loss_grad(p, norm, t, bf=True)
    assert abs(l16 - l64) <= 0.02 * l64
    assert np.linalg.norm(g16 - g64) <= 0.03 * np.linalg.norm(g64)


def test_adam_closed_form():
    p = np.array([1.0, -2.0])
    m1 = m2 = np.zeros(2)
    g = np.array([0.5, -0.25])
    p1, m1, m2 = mo.adam(p, g, m1, m2, 1, 0.1)
    # first step: m_hat = g, v_hat = g^2 -> step = lr * sign(g) (up to eps)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-6)


@pytest.mark.parametrize("n", [1, 17, 64])
def test_predict_shapes_and_residual_form(n):
    p, norm = _net()
    t = _tuples(n)
    out = mo.predict(p, norm, t[:, :6])
    assert out.shape == (n, 3) and np.all(np.isfinite(out))
    y = mo.forward(p, mo.features(t[:, :6], norm))["y"]
    assert np.allclose(out, t[:, :3] + y * norm[15:18] + norm[12:15], atol=1e-6)
