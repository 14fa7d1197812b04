"""Synthetic, seeded H.psi workloads (host side; numpy, FP64).

These stand in for what dft-zig's SCF driver hands to the H.psi callback for
one k-point: the G-sphere, the FFT grid, local_r, the KB projector tables and
the atoms.  No UPF files ship with the reference (its pseudo/ directory holds
only a README), so projectors are Gaussian s/p functions with closed-form
radial transforms.

Workloads:
  bench     cubic L = 20.5 bohr, ecut 15 Ha (30 Ry), grid 72^3, N_pw = 23,847, 160 bands,
            V_loc = sum of 64 Gaussians (sigma 1.5 bohr, depth U[-2,0] Ha), 64 atoms x (s + 3p),
            D = diag(+1.0, -0.5) Ha, psi_Gb = (N(0,1) + i N(0,1)) / (1 + |G|^2).
  si8       Si conventional cubic cell (a = 5.431 A, diamond sites), 20 Ry, grid 32^3,
            k = (0.25, 0.25, 0.25) fractional, 20 bands, same synthetic potential/projector recipe.
  tiny      small skewed cell, non-Gamma k, 8^3 grid, 87 plane waves, 2 atoms (s+p; s+s+d), 4 bands.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import tables

BENCH_SEED = 20261202


@dataclass
class SpeciesDesc:
    """A KB species: l per beta, closed-form radial parameters, D_ij in Ry (row-major)."""

    l_list: list
    rc: list
    dij: np.ndarray

    @property
    def m_total(self) -> int:
        return sum(2 * l + 1 for l in self.l_list)

    def radial_fns(self):
        return [tables.gaussian_radial(l, rc) for l, rc in zip(self.l_list, self.rc)]


@dataclass
class Workload:
    name: str
    cell: np.ndarray
    recip: np.ndarray
    volume: float
    grid: tuple
    ecut_ry: float
    k_cart: np.ndarray
    basis: tables.Basis
    local_r: np.ndarray
    species: list
    atom_pos: np.ndarray  # [natom, 3] bohr
    atom_species: np.ndarray  # [natom] int
    psi: np.ndarray  # (nbands, npw) complex128
    meta: dict = field(default_factory=dict)

    @property
    def npw(self) -> int:
        return self.basis.n

    @property
    def nbands(self) -> int:
        return int(self.psi.shape[0])

    @property
    def inv_volume(self) -> float:
        return 1.0 / self.volume

    def phi(self, s: int) -> np.ndarray:
        sp = self.species[s]
        return tables.build_phi(self.basis.kpg, sp.l_list, sp.radial_fns())

    def projector_count(self) -> int:
        return int(sum(self.species[s].m_total for s in self.atom_species))


def gaussian_potential(grid, cell_len, centers, depths_ry, sigma):
    """V(r) = sum_j d_j exp(-|r - c_j|^2_min-image / (2 sigma^2)) on the FFT-order grid of a cubic cell."""
    nx, ny, nz = grid
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.stack([x.ravel() * cell_len / nx, y.ravel() * cell_len / ny, z.ravel() * cell_len / nz], axis=1)
    v = np.zeros(r.shape[0])
    for c, d in zip(centers, depths_ry):
        dr = r - c[None, :]
        dr -= cell_len * np.round(dr / cell_len)
        v += d * np.exp(-np.sum(dr * dr, axis=1) / (2.0 * sigma * sigma))
    return v


def _random_psi(rng, nbands, kinetic):
    re = rng.standard_normal((nbands, kinetic.shape[0]))
    im = rng.standard_normal((nbands, kinetic.shape[0]))
    return (re + 1j * im) / (1.0 + kinetic)[None, :]


def bench(nbands: int = 160, seed: int = BENCH_SEED, grid=(72, 72, 72)) -> Workload:
    """The H.psi bench workload; `grid` (default 72^3, the bench size) may be any FFT grid that
    holds the 15 Ha sphere, for kernels specialised by grid size."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = 20.5
    cell = np.eye(3) * L
    recip = tables.reciprocal(cell)
    grid = tuple(int(v) for v in grid)
    ecut = 30.0
    centers = rng.uniform(0.0, L, size=(64, 3))
    depths = rng.uniform(-2.0, 0.0, size=64) * 2.0  # Ha -> Ry
    atom_pos = rng.uniform(0.0, L, size=(64, 3))
    basis = tables.generate_basis(recip, ecut, np.zeros(3))
    local_r = gaussian_potential(grid, L, centers, depths, 1.5)
    sp = SpeciesDesc([0, 1], [1.0, 1.0], np.diag([2.0, -1.0]))
    psi = _random_psi(rng, nbands, basis.kinetic)
    return Workload("bench", cell, recip, L ** 3, grid, ecut, np.zeros(3), basis, local_r, [sp],
                    atom_pos, np.zeros(64, dtype=np.int32), psi, {"seed": seed})


def si8(nbands: int = 20, seed: int = 8) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    a = 5.431 * tables.ANGSTROM_TO_BOHR
    cell = np.eye(3) * a
    recip = tables.reciprocal(cell)
    ecut = 20.0
    grid = tables.auto_grid(ecut, recip)
    frac = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    atom_pos = frac @ cell
    k_cart = np.array([0.25, 0.25, 0.25]) @ recip
    basis = tables.generate_basis(recip, ecut, k_cart)
    depths = np.full(8, -1.5)
    local_r = gaussian_potential(grid, a, atom_pos, depths, 1.2)
    sp = SpeciesDesc([0, 1], [0.9, 1.1], np.diag([1.6, 0.7]))
    psi = _random_psi(rng, nbands, basis.kinetic)
    return Workload("si8", cell, recip, a ** 3, grid, ecut, k_cart, basis, local_r, [sp],
                    atom_pos, np.zeros(8, dtype=np.int32), psi, {"seed": seed})


def tiny(nbands: int = 4, seed: int = 7) -> Workload:
    rng = np.random.Generator(np.random.PCG64(seed))
    cell = np.array([[7.0, 0.0, 0.0], [0.6, 6.5, 0.0], [0.3, -0.4, 7.4]])
    recip = tables.reciprocal(cell)
    vol = abs(float(np.linalg.det(cell)))
    grid = (8, 8, 8)
    ecut = 6.0
    k_cart = np.array([0.1, 0.2, -0.15]) @ recip
    basis = tables.generate_basis(recip, ecut, k_cart)
    nx, ny, nz = grid
    local_r = rng.uniform(-1.0, 0.5, size=nx * ny * nz)
    atom_pos = rng.uniform(0.0, 1.0, size=(2, 3)) @ cell
    sp0 = SpeciesDesc([0, 1], [0.8, 1.0], np.diag([1.3, -0.6]))
    sp1 = SpeciesDesc([0, 0, 2], [0.7, 1.1, 0.9],
                      np.array([[0.9, 0.2, 0.0], [0.2, -0.4, 0.0], [0.0, 0.0, 0.5]]))
    psi = _random_psi(rng, nbands, basis.kinetic)
    return Workload("tiny", cell, recip, vol, grid, ecut, k_cart, basis, local_r, [sp0, sp1],
                    atom_pos, np.array([0, 1], dtype=np.int32), psi, {"seed": seed})


BUILDERS = {"bench": bench, "si8": si8, "tiny": tiny}
