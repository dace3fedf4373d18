"""CPU restatement of the learned prox warm start (test infrastructure only).

The network of wrapd_amd/csrc/device/wd_mlp.hpp and wrapd_amd/csrc/wd_mlp.hip in numpy:
forward, mean-squared loss, backward and Adam, with the kernels' BF16 rounding points
(standardised inputs, weights, hidden activations, output and hidden-layer gradient
operands) and FP32 accumulation, or with `bf16=False` in float64 and no rounding (the
exact gradient, pinned by finite differences in tests/test_mlp_oracle.py).  The network
has no counterpart in the reference (it extends the stretch prox of
src/TetElements.cpp:302-318); its targets are prox solutions, checked in the tests by the
prox optimality condition.

Parameters (4803 floats): W1[64][6] b1[64] W2[64][64] b2[64] W3[3][64] b3[3].
Normalisation (18 floats): in_mean[6] in_std[6] out_mean[3] out_std[3].
Tuples (9 floats): s0 s1 s2 w^2/vol mu lambda | x0 x1 x2.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 6, 64, 3
OFF = dict(W1=0, b1=384, W2=448, b2=4544, W3=4608, b3=4800)
N_PARAMS = 4803


def unpack(p):
    p = np.asarray(p)
    return dict(W1=p[0:384].reshape(64, 6), b1=p[384:448], W2=p[448:4544].reshape(64, 64), b2=p[4544:4608],
                W3=p[4608:4800].reshape(3, 64), b3=p[4800:4803])


def pack(d):
    return np.concatenate([d["W1"].ravel(), d["b1"], d["W2"].ravel(), d["b2"], d["W3"].ravel(), d["b3"]])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def gelu(x):
    c, a = 0.7978845608028654, 0.044715
    return 0.5 * x * (1.0 + np.tanh(c * (x + a * x ** 3)))


def gelu_d(x):
    c, a = 0.7978845608028654, 0.044715
    t = np.tanh(c * (x + a * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * c * (1.0 + 3.0 * a * x * x)


def features(inp, norm, dtype=np.float32):
    """inp (n, 6): s0 s1 s2 w^2/vol mu lambda -> standardised features (n, 6)."""
    inp = np.asarray(inp, dtype=np.float64)
    raw = np.column_stack([inp[:, 0], inp[:, 1], inp[:, 2], np.log(inp[:, 3] / inp[:, 4]),
                           np.log(inp[:, 5] / inp[:, 4]), np.log(inp[:, 4])]).astype(dtype)
    norm = np.asarray(norm, dtype=dtype)
    return (raw - norm[0:6]) / norm[6:12]


def targets(tuples, norm, dtype=np.float32):
    t = np.asarray(tuples, dtype=dtype)
    norm = np.asarray(norm, dtype=dtype)
    return ((t[:, 6:9] - t[:, 0:3]) - norm[12:15]) / norm[15:18]


def forward(params, x, bf=True):
    """x (n, 6) standardised features -> cache with y (n, 3) standardised outputs."""
    dt = np.float32 if bf else np.float64
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=dt))
    P = {k: np.asarray(v, dtype=dt) for k, v in unpack(params).items()}
    W1, W2, W3 = r(P["W1"]), r(P["W2"]), r(P["W3"])
    xb = r(np.asarray(x, dtype=dt))
    z1 = xb @ W1.T + P["b1"]
    h1 = r(gelu(z1).astype(dt))
    z2 = h1 @ W2.T + P["b2"]
    h2 = r(gelu(z2).astype(dt))
    y = h2 @ W3.T + P["b3"]
    return dict(xb=xb, z1=z1, h1=h1, z2=z2, h2=h2, y=y, W1=W1, W2=W2, W3=W3)


def predict(params, norm, inp):
    """Prox solution guesses (n, 3) for raw inputs (n, 6), as the device's inference."""
    y = forward(params, features(inp, norm))["y"]
    norm = np.asarray(norm, dtype=np.float32)
    return np.asarray(inp, dtype=np.float32)[:, 0:3] + y * norm[15:18] + norm[12:15]


def loss_grad(params, norm, tuples, bf=True):
    """Mean over samples of the squared error of the standardised outputs, and its
    gradient with respect to the parameters (float32 with the kernels' roundings)."""
    dt = np.float32 if bf else np.float64
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=dt))
    x = features(np.asarray(tuples)[:, 0:6], norm, dt)
    tn = targets(tuples, norm, dt)
    c = forward(params, x, bf)
    n = len(x)
    e = c["y"] - tn
    loss = float(np.sum(e.astype(np.float64) ** 2) / n)
    dy = (2.0 * e / n).astype(dt)
    dyb = r(dy)
    dh2 = dyb @ c["W3"]
    dz2 = (dh2 * gelu_d(c["z2"])).astype(dt)
    dz2b = r(dz2)
    dh1 = dz2b @ c["W2"]
    dz1 = (dh1 * gelu_d(c["z1"])).astype(dt)
    dz1b = r(dz1)
    g = dict(W3=dyb.T @ c["h2"], b3=dy.sum(0), W2=dz2b.T @ c["h1"], b2=dz2.sum(0), W1=dz1b.T @ c["xb"],
             b1=dz1.sum(0))
    return loss, pack(g).astype(dt)


def adam(params, grad, m1, m2, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    m1 = beta1 * m1 + (1 - beta1) * grad
    m2 = beta2 * m2 + (1 - beta2) * grad * grad
    mh = m1 / (1 - beta1 ** t)
    vh = m2 / (1 - beta2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m1, m2
