"""The learned preconditioner's CPU oracle (oracle/mlp.py) pinned by finite differences, the TPA
and kinetic-shift closed forms, and the zero-output-layer identity; the product's initialiser
layout against it."""
import numpy as np

from dzhip import MLP_PARAMS, precond
from oracle import mlp as omlp


def _random_samples(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.stack([rng.uniform(0.0, 1.5, n), rng.uniform(-0.1, 0.2, n), rng.uniform(-2.5, -0.25, n)], 1)
    # Base factors of both kinds: TPA values in (0, 1] and kinetic shifts of either sign.
    kx = rng.uniform(0.0, 6.0, n)
    d = rng.uniform(0.0, 6.0, n)
    lam = d - rng.choice([-1.0, 1.0], n) * rng.uniform(0.05, 3.0, n)
    base = np.where(np.arange(n) % 2 == 0, omlp.tpa(kx), omlp.kinetic_shift(d, lam))
    t = base * np.exp(rng.normal(0.0, 0.3, n))
    return x, base, t, rng.uniform(0.1, 2.0, n)


def _params(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    p = omlp.init_params(seed)
    p[-(omlp.H + 1):] = rng.normal(0.0, 0.3, omlp.H + 1)
    return p


def test_mlp_gradient_finite_difference():
    x, kx, t, w = _random_samples(300, 9)
    p = _params(4)
    _, g = omlp.loss_and_grad(p, x, kx, t, w)
    rng = np.random.default_rng(0)
    for idx in rng.choice(omlp.N_PARAMS, 40, replace=False):
        e = np.zeros_like(p)
        e[idx] = 1e-6
        lp, _ = omlp.loss_and_grad(p + e, x, kx, t, w)
        lm, _ = omlp.loss_and_grad(p - e, x, kx, t, w)
        assert abs((lp - lm) / 2e-6 - g[idx]) < 1e-7 + 1e-5 * abs(g[idx])


def test_zero_output_layer_reduces_to_base_and_closed_forms():
    x, base, _, _ = _random_samples(50, 2)
    p = omlp.init_params(3)
    assert np.array_equal(omlp.scaling(p, x, base), base)
    # Kinetic shift: 1/(d - lambda), |den| clamped to 1e-8 with its sign.
    ks = omlp.kinetic_shift(np.array([2.0, 1.0, 1.0 + 1e-9, 1.0 - 1e-9, 1.0]), 1.0)
    assert ks[0] == 1.0 and ks[1] == 1e8 and ks[2] == 1e8 and ks[3] == -1e8 and ks[4] == 1e8
    # K(0) = 1, K -> 27/(16 x^4)... leading behaviour 8x^3/(16x^4) = 1/(2x) for large x.
    assert omlp.tpa(np.array([0.0]))[0] == 1.0
    assert abs(omlp.tpa(np.array([1e4]))[0] * 2e4 - 1.0) < 1e-3
    assert abs(omlp.tpa(np.array([1.0]))[0] - 65.0 / 81.0) < 1e-15


def test_silu_derivative():
    u = np.linspace(-8, 8, 101)
    num = (omlp.silu(u + 1e-6) - omlp.silu(u - 1e-6)) / 2e-6
    assert np.max(np.abs(num - omlp.dsilu(u))) < 1e-8


def test_product_init_layout():
    p = precond.init_params(0)
    assert p.shape == (MLP_PARAMS,) and p.dtype == np.float32
    assert np.all(p[-(omlp.H + 1):] == 0.0)
    w1, b1, w2, b2, w3, b3 = omlp.unpack(p)
    assert w1.shape == (32, 3) and w2.shape == (32, 32) and b3 == 0.0
