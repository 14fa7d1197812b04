"""Learned preconditioner: sample collection, training and model files (host orchestration).

The network and its kernels live in the library (csrc/precond.hip, C ABI dzhip_ham_set_precond_model,
dzhip_ham_training_samples, dzhip_mlp_train, dzhip_mlp_loss_grad).  A model is the 1217 FP32
parameters, the feature energy scale e_s (Ha) and the base preconditioner B it rescales:

    s(G, b) = B(G, b) exp(z),  z = MLP(|k+G|^2/2 / e_s, lambda_b/2 / e_s, log10|r_b| / 4)
    B = 1 / (|k+G|^2 - lambda_b)  (the reference's kinetic shift)  or  K_tpa(|k+G|^2 / <x_b|T|x_b>)

Training data are residual/correction pairs of FP64 Davidson solves that converged: for the Ritz
vector x_b of an early iteration, the correction is the part of x_b outside the converged
invariant subspace (negated) and the target is the per-coefficient real factor that maps the
residual onto it, relative to the band's best scale of B.  With a zero output layer the model is
exactly B, which is where training starts.  Samples come either from cold solves on a fixed
potential (collect) or from the warm-started solves of an SCF run (collect_scf), which is the
distribution an SCF's eigensolver sees.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import FP64, MLP_PARAMS, PRECOND_KINETIC, PRECOND_TPA, Context, Hamiltonian

HIDDEN = 32
ENERGY_SCALE = 10.0  # Ha
MODEL_DIR = Path(__file__).resolve().parent / "models"


def init_params(seed: int = 0) -> np.ndarray:
    """Hidden layers U(+-1/sqrt(fan_in)), output layer zero (the model starts as TPA)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w1 = rng.uniform(-1, 1, (HIDDEN, 3)) / np.sqrt(3.0)
    b1 = rng.uniform(-1, 1, HIDDEN) / np.sqrt(3.0)
    w2 = rng.uniform(-1, 1, (HIDDEN, HIDDEN)) / np.sqrt(HIDDEN)
    b2 = rng.uniform(-1, 1, HIDDEN) / np.sqrt(HIDDEN)
    p = np.concatenate([w1.ravel(), b1, w2.ravel(), b2, np.zeros(HIDDEN), [0.0]]).astype(np.float32)
    assert p.shape[0] == MLP_PARAMS
    return p


def collect(ham: Hamiltonian, nbands: int, snapshots: int = 16, stride: int = 1, tol: float = 1e-9,
            max_iter: int = 300, seed: int = 0, energy_scale: float = ENERGY_SCALE, capacity: int = 1 << 22,
            base: int = PRECOND_TPA):
    """Samples (6, n) of one cold FP64 solve of `ham` with the base preconditioner (weight-0
    samples dropped)."""
    vals, _, st = ham.eig_solve(nbands, tol=tol, max_iter=max_iter, precision=FP64, preconditioner=base,
                                want_vectors=False, seed=seed, collect_snapshots=snapshots)
    if not st["converged"]:
        raise RuntimeError(f"training solve did not converge: {st}")
    smp = ham.training_samples(energy_scale, stride, capacity, base=base)
    return smp[:, smp[5] > 0]


def collect_scf(system, base: int = PRECOND_KINETIC, steps: int = 12, snapshots: int = 16, tol: float = 1e-9,
                stride: int = 4, energy_scale: float = ENERGY_SCALE, capacity: int = 1 << 21, **scf_opts):
    """Samples (6, n) from the eigen-solves of the first `steps` iterations of an FP64 SCF of
    `system`: every solve starts from the previous iteration's vectors, runs the base
    preconditioner to `tol` on all bands and records its first `snapshots` iterations."""
    from . import scf as dscf

    chunks = []

    def hook(k, ham, st):
        if st["converged"]:
            smp = ham.training_samples(energy_scale, stride, capacity, base=base)
            chunks.append(smp[:, smp[5] > 0])

    drv = dscf.ScfDriver(system)
    drv.solve_hook = hook
    try:
        opts = dict(conv=0.0, eig_tol=tol, eig_max_iter=400, warmup_steps=0, nconv=0)
        opts.update(scf_opts)
        drv.run(max_iter=steps, precision=FP64, preconditioner=base, collect_snapshots=snapshots, **opts)
    finally:
        drv.close()
    if not chunks:
        raise RuntimeError("no converged training solve")
    return np.concatenate(chunks, axis=1)


def normalise_weights(samples: np.ndarray) -> np.ndarray:
    """Scale the weights to mean 1 (the loss is a mean over each minibatch)."""
    out = np.array(samples, dtype=np.float32, copy=True)
    out[5] /= max(float(np.mean(out[5])), 1e-30)
    return out


def train(ctx: Context, samples: np.ndarray, epochs: int = 30, batch: int = 8192, lr: float = 2e-3,
          seed: int = 1, init: np.ndarray | None = None):
    """Device Adam training; returns (params, loss after each epoch)."""
    p0 = init_params(seed) if init is None else np.asarray(init, dtype=np.float32)
    return ctx.mlp_train(normalise_weights(samples), p0, epochs, batch, lr, seed)


def save(path, params, energy_scale=ENERGY_SCALE, base=PRECOND_TPA, **meta):
    np.savez(path, params=np.asarray(params, dtype=np.float32), energy_scale=np.array(energy_scale),
             base=np.array(base), **meta)


def load(path):
    """(params, energy_scale, base) of a model file."""
    with np.load(path, allow_pickle=False) as d:
        base = int(d["base"]) if "base" in d.files else PRECOND_TPA
        return d["params"].astype(np.float32), float(d["energy_scale"]), base


def train_on_systems(ctx: Context, systems, nbands=None, snapshots=16, epochs=40, seed=1, lr=2e-3,
                     base=PRECOND_TPA):
    """Collect residual/correction pairs from one cold FP64 solve per k-point of each SCF system
    (bare ionic potential) and train a model from the base preconditioner.  Returns (params,
    losses, sample count)."""
    from . import scf as dscf

    chunks = []
    for system in systems:
        drv = dscf.ScfDriver(system)
        try:
            for ham in drv.hams:
                ham.set_local(drv.v_ion)
                chunks.append(collect(ham, nbands or system.nbands, snapshots=snapshots, base=base))
        finally:
            drv.close()
    samples = np.concatenate(chunks, axis=1)
    params, hist = train(ctx, samples, epochs=epochs, lr=lr, seed=seed)
    return params, hist, samples.shape[1]


def train_on_scf(ctx: Context, systems, base=PRECOND_KINETIC, steps=12, snapshots=16, epochs=40, seed=1,
                 lr=2e-3, **scf_opts):
    """Collect residual/correction pairs from the warm-started solves of a short FP64 SCF of
    each system (collect_scf) and train a model from the base preconditioner.  Returns (params,
    losses, sample count)."""
    samples = np.concatenate([collect_scf(s, base=base, steps=steps, snapshots=snapshots, **scf_opts)
                              for s in systems], axis=1)
    params, hist = train(ctx, samples, epochs=epochs, lr=lr, seed=seed)
    return params, hist, samples.shape[1]
