"""FP64 H.psi: kinetic + local (FFT) + Kleinman-Bylander nonlocal, single and batched,
plus the dense-matrix brute force.

Restates (paths relative to the dft-zig source tree):
  apply_hamiltonian_with_workspace  src/features/scf/apply.zig:577-641
  apply_local_potential_safe (fused scatter_fft/gather_fft path)  apply.zig:811-852
  compute_atom_phase / project_nonlocal_coeff / apply_nonlocal_dij /
  backproject_nonlocal_accum / apply_nonlocal_potential_safe      apply.zig:883-1036
  batched nonlocal (phase, 2x dgemm, D, 2x dgemm + conj phase)     apply.zig:1118-1482
  apply_hamiltonian_batched                                        apply.zig:1584-1618
  apply_overlap (S = I for norm-conserving species)                apply.zig:1041-1111
  build_hamiltonian / add_nonlocal_contribution /
  compute_nonlocal_value (dense brute force)       src/features/hamiltonian/hamiltonian.zig:174-374

Blocks of vectors are numpy arrays of shape (ncols, npw): in memory that is the
reference's column-major [npw x ncols] layout with column b at [b*npw, (b+1)*npw)
(src/features/linalg/lobpcg/common.zig:54-62).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import fftconv, pw
from .projectors import Species, angular_factor


@dataclass
class Atom:
    position: np.ndarray  # bohr, Cartesian
    species: int


@dataclass
class HamContext:
    """The data an ApplyContext carries (apply.zig:303-336), flattened."""

    grid: tuple
    basis: pw.Basis
    local_r: np.ndarray  # f64 [N] FFT order, Ry
    species: list = field(default_factory=list)  # [Species]
    atoms: list = field(default_factory=list)  # [Atom]
    inv_volume: float = 1.0

    def __post_init__(self):
        self.fft_idx = pw.fft_indices(self.basis, self.grid)
        self.total = int(np.prod(self.grid))
        self._phi = [sp.phi(self.basis.kpg) for sp in self.species]

    @property
    def npw(self) -> int:
        return self.basis.n

    def phi(self, s: int) -> np.ndarray:
        return self._phi[s]

    def projector_count(self) -> int:
        return sum(self.species[a.species].m_total for a in self.atoms)


def apply_local(ctx: HamContext, x: np.ndarray) -> np.ndarray:
    """Fused local path, apply.zig:820-852: scatter*N -> inverse -> *V -> forward -> gather/N."""
    n = float(ctx.total)
    work = pw.scatter_fft(ctx.fft_idx, x, ctx.total, n)
    work = fftconv.inverse3d(work, ctx.grid)
    work = work * ctx.local_r
    work = fftconv.forward3d(work, ctx.grid)
    return pw.gather_fft(ctx.fft_idx, work, 1.0 / n)


def _apply_dij(sp: Species, coeff: np.ndarray) -> np.ndarray:
    """apply_nonlocal_dij radial form (apply.zig:940-966); coeff shape (m_total, ...)."""
    out = np.zeros_like(coeff)
    for b, lb in enumerate(sp.l_list):
        for j, lj in enumerate(sp.l_list):
            if lj != lb:
                continue
            d = sp.dij[b, j]
            if d == 0.0:
                continue
            for mi in range(sp.m_counts[b]):
                out[sp.m_offsets[b] + mi] += d * coeff[sp.m_offsets[j] + mi]
    return out


def apply_nonlocal(ctx: HamContext, x: np.ndarray) -> np.ndarray:
    """apply_nonlocal_potential_safe (apply.zig:1000-1036) for one vector."""
    out = np.zeros(ctx.npw, dtype=np.complex128)
    g = ctx.basis.cart
    for s, sp in enumerate(ctx.species):
        if sp.m_total == 0:
            continue
        phi = ctx.phi(s)
        for atom in ctx.atoms:
            if atom.species != s:
                continue
            phase = np.exp(1j * (g @ atom.position))  # expi(dot(G, tau)), apply.zig:893
            xphase = x * phase
            coeff = phi @ xphase
            coeff2 = _apply_dij(sp, coeff)
            accum = phi.T @ coeff2
            out += np.conj(phase) * accum * ctx.inv_volume
    return out


def apply_nonlocal_batched(ctx: HamContext, xb: np.ndarray) -> np.ndarray:
    """apply_nonlocal_batched (apply.zig:1449-1482); xb shape (ncols, npw)."""
    out = np.zeros_like(xb, dtype=np.complex128)
    g = ctx.basis.cart
    for s, sp in enumerate(ctx.species):
        if sp.m_total == 0:
            continue
        phi = ctx.phi(s)
        for atom in ctx.atoms:
            if atom.species != s:
                continue
            phase = np.exp(1j * (g @ atom.position))
            xphase = (xb * phase[None, :]).T  # [npw, ncols], row-major as apply.zig:1131-1139
            coeffs = phi @ xphase.real + 1j * (phi @ xphase.imag)  # 2 real dgemm, apply.zig:1143-1183
            coeffs2 = _apply_dij(sp, coeffs)
            res = phi.T @ coeffs2.real + 1j * (phi.T @ coeffs2.imag)  # apply.zig:1278-1321
            out += (np.conj(phase)[:, None] * res * ctx.inv_volume).T
    return out


def apply_hamiltonian(ctx: HamContext, x: np.ndarray) -> np.ndarray:
    """y = T x + V_loc x + V_nl x (apply.zig:577-641). Raises on length mismatch (apply.zig:584)."""
    x = np.asarray(x, dtype=np.complex128)
    if x.shape != (ctx.npw,):
        raise ValueError("InvalidMatrixSize")
    y = ctx.basis.kinetic * x
    y = y + apply_local(ctx, x)
    if ctx.species:
        y = y + apply_nonlocal(ctx, x)
    return y


def apply_hamiltonian_batched(ctx: HamContext, xb: np.ndarray) -> np.ndarray:
    """apply_hamiltonian_batched (apply.zig:1584-1618): local per column, nonlocal batched."""
    xb = np.asarray(xb, dtype=np.complex128)
    if xb.ndim != 2 or xb.shape[1] != ctx.npw:
        raise ValueError("InvalidMatrixSize")
    yb = np.empty_like(xb)
    for c in range(xb.shape[0]):
        yb[c] = ctx.basis.kinetic * xb[c] + apply_local(ctx, xb[c])
    if ctx.species:
        yb = yb + apply_nonlocal_batched(ctx, xb)
    return yb


def apply_overlap(ctx: HamContext, x: np.ndarray) -> np.ndarray:
    """S x; norm-conserving species have no q_ij, so S = I (apply.zig:1074-1111)."""
    return np.array(x, dtype=np.complex128, copy=True)


# ---------------------------------------------------------------- dense brute force

def local_potential_g(ctx: HamContext) -> np.ndarray:
    """V(q) on the full FFT-order grid with V(r) = sum_q V(q) e^{iq.r}: forward(local_r)/N."""
    return fftconv.forward3d(ctx.local_r.astype(np.complex128), ctx.grid) / ctx.total


def dense_local(ctx: HamContext) -> np.ndarray:
    """H_loc[i,j] = V((G_i - G_j) mod grid): the circular convolution the FFT path computes
    (the closed-form construction of src/tests.zig:975-1106)."""
    nx, ny, nz = ctx.grid
    vq = local_potential_g(ctx)
    b = ctx.basis
    dh = (b.h[:, None].astype(np.int64) - b.h[None, :]) % nx
    dk = (b.k[:, None].astype(np.int64) - b.k[None, :]) % ny
    dl = (b.l[:, None].astype(np.int64) - b.l[None, :]) % nz
    return vq[dh + nx * (dk + ny * dl)]


def dense_nonlocal(ctx: HamContext) -> np.ndarray:
    """add_nonlocal_contribution with the Legendre angular factor (hamiltonian.zig:259-374)."""
    b = ctx.basis
    n = b.n
    h = np.zeros((n, n), dtype=np.complex128)
    kpg = b.kpg
    gm = np.linalg.norm(kpg, axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        cosg = (kpg @ kpg.T) / np.outer(gm, gm)
    cosg[(gm[:, None] == 0.0) | (gm[None, :] == 0.0)] = 1.0
    cosg = np.clip(cosg, -1.0, 1.0)
    for s, sp in enumerate(ctx.species):
        rad = sp.radial_values(kpg)
        val = np.zeros((n, n))
        for i, li in enumerate(sp.l_list):
            ang = np.vectorize(lambda c, l=li: angular_factor(l, c))(cosg)
            for j, lj in enumerate(sp.l_list):
                if lj != li or sp.dij[i, j] == 0.0:
                    continue
                val += sp.dij[i, j] * ang * np.outer(rad[i], rad[j])
        for atom in ctx.atoms:
            if atom.species != s:
                continue
            ph = np.exp(-1j * (b.cart @ atom.position))
            h += np.outer(ph, np.conj(ph)) * ctx.inv_volume * val
    return h


def dense_hamiltonian(ctx: HamContext, with_nonlocal: bool = True) -> np.ndarray:
    """build_hamiltonian (hamiltonian.zig:174-212) with the local part supplied as local_r."""
    h = dense_local(ctx).astype(np.complex128)
    h[np.diag_indices(ctx.npw)] += ctx.basis.kinetic
    if with_nonlocal and ctx.species:
        h += dense_nonlocal(ctx)
    return h
