"""Block-Davidson eigensolver as the reference implements it (called "LOBPCG" there), FP64.

Restates (paths relative to the dft-zig source tree):
  Operator / Options              src/features/linalg/lobpcg/common.zig:12-52
  serial.solve                    src/features/linalg/lobpcg/serial.zig:55-307
  build_projected (+ Hermitize)   common.zig:231-259
  combine_columns_matrix          common.zig:113-143
  precondition (kinetic shift)    common.zig:214-226
  orthonormalize_vector (MGS)     common.zig:263-284
  orthonormalize_all              common.zig:298-312
  init_random_vectors/next_rand01 common.zig:287-319 (LCG seed 0x6f3d3cbe, serial.zig:111)

The small projected eigenproblem uses numpy.linalg.eigh (LAPACK zheevd) where the
reference calls zheev: same ascending eigenvalues; eigenvector phases may
differ, which does not change Ritz values or the spanned subspace.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

MASK64 = (1 << 64) - 1


@dataclass
class Options:
    max_iter: int = 40
    tol: float = 1e-6
    max_subspace: int = 0
    block_size: int = 0
    init_diagonal: bool = False
    init_vectors: Optional[np.ndarray] = None  # (cols, n)


class Lcg:
    """x <- x*6364136223846793005 + 1 (mod 2^64); u = (x >> 11) / 2^53 (common.zig:315-319)."""

    def __init__(self, seed: int = 0x6F3D3CBE):
        self.state = seed & MASK64

    def next01(self) -> float:
        self.state = (self.state * 6364136223846793005 + 1) & MASK64
        return (self.state >> 11) / 9007199254740992.0


def init_random_vectors(n: int, m: int, rng: Lcg) -> np.ndarray:
    v = np.empty((m, n), dtype=np.complex128)
    for col in range(m):
        for i in range(n):
            r = rng.next01() - 0.5
            im = rng.next01() - 0.5
            v[col, i] = complex(r, im)
    return v


def orthonormalize_vector(v: np.ndarray, basis: np.ndarray) -> float:
    """MGS of v against rows of basis, then normalise in place (common.zig:263-284)."""
    for j in range(basis.shape[0]):
        dot = np.vdot(basis[j], v)
        v -= dot * basis[j]
    norm = float(np.linalg.norm(v))
    if norm > 0.0:
        v /= norm
    return norm


def orthonormalize_all(v: np.ndarray, m: int, rng: Lcg) -> None:
    n = v.shape[1]
    for col in range(m):
        for _ in range(3):
            norm = orthonormalize_vector(v[col], v[:col])
            if norm > 1e-8:
                break
            for i in range(n):
                v[col, i] = complex(rng.next01() - 0.5, rng.next01() - 0.5)
        if np.linalg.norm(v[col]) < 1e-8:
            raise ValueError("InvalidMatrixSize")


def precondition(diag: np.ndarray, lam: float, r: np.ndarray) -> np.ndarray:
    """out = r / (diag - lambda), |denominator| clamped at 1e-8 (common.zig:214-226)."""
    denom = diag - lam
    small = np.abs(denom) < 1e-8
    denom = np.where(small, np.where(denom >= 0.0, 1e-8, -1e-8), denom)
    return r / denom


def build_projected(v: np.ndarray, w: np.ndarray) -> np.ndarray:
    """T = V^H W, then T = (T + T^H)/2 with a real diagonal (common.zig:231-259)."""
    t = np.conj(v) @ w.T
    t = 0.5 * (t + np.conj(t.T))
    t[np.diag_indices(t.shape[0])] = t.diagonal().real
    return t


@dataclass
class EigenResult:
    values: np.ndarray  # [nbands] ascending
    vectors: np.ndarray  # (nbands, n)
    iterations: int
    apply_calls: int
    converged: bool


def davidson(apply_batch: Callable[[np.ndarray], np.ndarray], n: int, diag: np.ndarray,
             nbands: int, opts: Options = Options()) -> EigenResult:
    """serial.solve (serial.zig:55-307), standard problem (apply_s = null)."""
    if nbands == 0 or n == 0:
        return EigenResult(np.zeros(0), np.zeros((0, n), complex), 0, 0, True)
    if diag.shape[0] != n:
        raise ValueError("InvalidMatrixSize")
    max_sub = min(n, 4 * nbands + 8) if opts.max_subspace == 0 else min(n, max(opts.max_subspace, nbands))
    block = min(max_sub, nbands + 4) if opts.block_size == 0 else min(max_sub, max(opts.block_size, nbands))
    rng = Lcg()
    m = block
    v = np.zeros((max_sub, n), dtype=np.complex128)
    w = np.zeros((max_sub, n), dtype=np.complex128)
    if opts.init_diagonal:
        order = np.argsort(diag, kind="stable")
        for col in range(m):
            v[col, order[col]] = 1.0
    else:
        v[:m] = init_random_vectors(n, m, rng)
    if opts.init_vectors is not None:
        cols = min(m, opts.init_vectors.shape[0])
        v[:cols] = opts.init_vectors[:cols]
    orthonormalize_all(v, m, rng)
    w[:m] = apply_batch(v[:m])
    calls = m
    last = None
    ritz = None
    for it in range(opts.max_iter):
        t = build_projected(v[:m], w[:m])
        evals, evecs = np.linalg.eigh(t)
        last = evals[:nbands].copy()
        c = evecs[:, :nbands]  # (m, nbands)
        ritz = (c.T @ v[:m])  # (nbands, n) == V*C
        res = (c.T @ w[:m])
        res = res - evals[:nbands, None] * ritz
        norms = np.linalg.norm(res, axis=1)
        if norms.max() < opts.tol:
            return EigenResult(last, ritz, it, calls, True)
        added = 0
        for b in range(nbands):
            if m + added >= max_sub:
                break
            if norms[b] < opts.tol:
                continue
            q = precondition(diag, float(evals[b]), res[b])
            norm = orthonormalize_vector(q, v[:m + added])
            if norm < 1e-8:
                continue
            v[m + added] = q
            added += 1
        if added > 0:
            w[m:m + added] = apply_batch(v[m:m + added])
            calls += added
        if added == 0:
            return EigenResult(last, ritz, it, calls, True)
        if m + added >= max_sub:
            v[:nbands] = ritz
            m = nbands
            orthonormalize_all(v, m, rng)
            w[:m] = apply_batch(v[:m])
            calls += m
            continue
        m += added
    return EigenResult(last, ritz, opts.max_iter, calls, False)
