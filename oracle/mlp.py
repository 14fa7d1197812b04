"""CPU restatement of the learned diagonal preconditioner (test infrastructure only).

There is no counterpart in the reference: it preconditions Davidson residuals with the
kinetic-shift diagonal (src/features/linalg/lobpcg/common.zig:214-226).  This file restates,
in numpy FP64, what the HIP kernels of the GPU library compute, so their forward pass, loss,
gradients and optimiser step can be checked:

  features  x = (|k+G|^2/2 / e_s, lambda_b/2 / e_s, log10 |r_b|)    e_s = feature energy scale (Ha)
  network   z = w3 . silu(W2 silu(W1 x + b1) + b2) + b3             3 -> 32 -> 32 -> 1, silu(u) = u/(1+e^-u)
  scaling   s = B * exp(z)                                          B: base preconditioner factor,
            B = K_tpa(|k+G|^2 / T_b)  (Teter-Payne-Allan polynomial) or
            B = 1 / (|k+G|^2 - lambda_b), |den| >= 1e-8  (the reference's kinetic shift)
  loss      L = sum_i w_i (s_i - t_i)^2 / sum_i w_i                 weighted least squares on samples i
  Adam      m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2, p -= lr mhat/(sqrt(vhat)+eps)

Parameter layout (1217 floats): W1[32][3], b1[32], W2[32][32], b2[32], w3[32], b3.

Pinned by: finite differences of the loss (central, FP64) for the gradient; z = 0 (zero last
layer) reducing s to the base factor exactly; silu and TPA closed forms.
"""
from __future__ import annotations

import numpy as np

H = 32
N_PARAMS = 3 * H + H + H * H + H + H + 1  # 1217


def tpa(x):
    """Teter-Payne-Allan K(x) = (27 + 18x + 12x^2 + 8x^3) / (27 + 18x + 12x^2 + 8x^3 + 16x^4)."""
    p = 27.0 + 18.0 * x + 12.0 * x * x + 8.0 * x ** 3
    return p / (p + 16.0 * x ** 4)


def kinetic_shift(diag, lam):
    """1 / (diag - lambda) with |den| clamped to >= 1e-8, sign kept (common.zig:214-226)."""
    den = np.asarray(diag, dtype=np.float64) - lam
    den = np.where(np.abs(den) < 1e-8, np.where(den >= 0.0, 1e-8, -1e-8), den)
    return 1.0 / den


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    w1 = p[o:o + 3 * H].reshape(H, 3); o += 3 * H
    b1 = p[o:o + H]; o += H
    w2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return w1, b1, w2, b2, w3, b3


def init_params(seed=0):
    """Uniform(+-1/sqrt(fan_in)) hidden layers, zero output layer (so s starts as TPA)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w1 = rng.uniform(-1, 1, (H, 3)) / np.sqrt(3.0)
    b1 = rng.uniform(-1, 1, H) / np.sqrt(3.0)
    w2 = rng.uniform(-1, 1, (H, H)) / np.sqrt(H)
    b2 = rng.uniform(-1, 1, H) / np.sqrt(H)
    return np.concatenate([w1.ravel(), b1, w2.ravel(), b2, np.zeros(H), [0.0]])


def silu(u):
    return u / (1.0 + np.exp(-u))


def dsilu(u):
    sg = 1.0 / (1.0 + np.exp(-u))
    return sg * (1.0 + u * (1.0 - sg))


def forward(p, x):
    """z for samples x [n, 3]; also the intermediates for backward."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    u1 = x @ w1.T + b1
    h1 = silu(u1)
    u2 = h1 @ w2.T + b2
    h2 = silu(u2)
    z = h2 @ w3 + b3
    return z, (x, u1, h1, u2, h2)


def scaling(p, x, base):
    """s = B exp(z(x)) for the base factors B of the samples."""
    z, _ = forward(p, x)
    return np.asarray(base, dtype=np.float64) * np.exp(z)


def loss_and_grad(p, x, base, t, w):
    """Weighted least-squares loss of s = B exp(z) against targets t and its gradient."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    z, (x_, u1, h1, u2, h2) = forward(p, x)
    s = np.asarray(base, dtype=np.float64) * np.exp(z)
    wsum = float(np.sum(w))
    diff = s - t
    loss = float(np.sum(w * diff * diff)) / wsum
    dz = 2.0 * w * diff * s / wsum          # dL/dz
    g_w3 = h2.T @ dz
    g_b3 = float(np.sum(dz))
    d2 = np.outer(dz, w3) * dsilu(u2)        # dL/du2
    g_w2 = d2.T @ h1
    g_b2 = d2.sum(0)
    d1 = (d2 @ w2) * dsilu(u1)               # dL/du1
    g_w1 = d1.T @ x_
    g_b1 = d1.sum(0)
    g = np.concatenate([g_w1.ravel(), g_b1, g_w2.ravel(), g_b2, g_w3, [g_b3]])
    return loss, g


def adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (p, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
