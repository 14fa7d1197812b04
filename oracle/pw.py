"""Lattice, plane-wave basis, FFT grid and G-sphere <-> grid index maps (FP64/int).

Restates (paths relative to the dft-zig source tree):
  reciprocal           src/features/math/math.zig:131-142
  next_fft_size        src/lib/fft/sizing.zig:1-15
  auto_grid/min_index  src/features/scf/pw_grid.zig:54-72
  generate (basis)     src/features/plane_wave/basis.zig:26-74
  PwGridMap.init       src/features/scf/pw_grid_map.zig:18-51
  build_fft_index_map  src/features/scf/fft_grid.zig:407-432
  index_to_freq        src/features/scf/fft_grid.zig:435-438
  build_fft_indices    src/features/scf/pw_grid_map.zig:58-79
  scatter_fft/gather_fft  src/features/scf/pw_grid_map.zig:95-117

Grid arrays are flat, in FFT order, index ``x + nx*(y + ny*z)``
(src/lib/fft/fft.zig:423,440); as numpy arrays of shape (nz, ny, nx).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

ANGSTROM_TO_BOHR = 1.0 / 0.52917721067


def reciprocal(cell: np.ndarray) -> np.ndarray:
    """Rows b_i with a_i . b_j = 2*pi*delta_ij (math.zig:131-142)."""
    a1, a2, a3 = (np.asarray(cell, dtype=np.float64)[i] for i in range(3))
    vol = float(np.dot(a1, np.cross(a2, a3)))
    s = 2.0 * math.pi / vol
    return np.stack([np.cross(a2, a3) * s, np.cross(a3, a1) * s, np.cross(a1, a2) * s])


def cell_volume(cell: np.ndarray) -> float:
    c = np.asarray(cell, dtype=np.float64)
    return abs(float(np.dot(c[0], np.cross(c[1], c[2]))))


def is_fft_size(n: int) -> bool:
    for f in (2, 3, 5):
        while n % f == 0:
            n //= f
    return n == 1


def next_fft_size(n: int) -> int:
    """Smallest 2/3/5-smooth integer >= n (sizing.zig:1-15)."""
    if n <= 1:
        return 1
    while not is_fft_size(n):
        n += 1
    return n


def auto_grid(ecut_ry: float, recip: np.ndarray, grid_scale: float = 1.0) -> tuple[int, int, int]:
    """FFT grid from the density cutoff 2*sqrt(ecut) (pw_grid.zig:54-68)."""
    scale = grid_scale if grid_scale > 0.0 else 1.0
    gmax = max(2.0, scale) * math.sqrt(ecut_ry)
    out = []
    for i in range(3):
        b = float(np.linalg.norm(recip[i]))
        n = int(math.ceil(gmax / b)) * 2 + 1
        out.append(next_fft_size(max(n, 3)))
    return tuple(out)


def min_index(n: int) -> int:
    """pw_grid.zig:70-72: lowest centred index, -floor(n/2)."""
    return -(n // 2)


def index_to_freq(i: int, n: int) -> int:
    """fft_grid.zig:435-438."""
    half = (n - 1) // 2
    return i if i <= half else i - n


@dataclass
class Basis:
    """Plane waves |k+G|^2 <= ecut_ry, ordered by loops h, k, l ascending (basis.zig:43-70)."""

    h: np.ndarray  # int32 [npw]
    k: np.ndarray
    l: np.ndarray
    cart: np.ndarray  # f64 [npw, 3]  G
    kpg: np.ndarray  # f64 [npw, 3]  k+G
    kinetic: np.ndarray  # f64 [npw]  |k+G|^2 (Ry)

    @property
    def n(self) -> int:
        return int(self.h.shape[0])


def generate_basis(recip: np.ndarray, ecut_ry: float, k_cart=(0.0, 0.0, 0.0)) -> Basis:
    """plane_wave.generate (basis.zig:26-74)."""
    recip = np.asarray(recip, dtype=np.float64)
    k_cart = np.asarray(k_cart, dtype=np.float64)
    min_b = min(float(np.linalg.norm(recip[i])) for i in range(3))
    gmax = math.sqrt(ecut_ry)
    max_n = int(math.ceil(gmax / min_b)) + 1
    r = np.arange(-max_n, max_n + 1, dtype=np.int64)
    # Loop order h (outer), k, l (inner) == C-order meshgrid with indexing="ij".
    hh, kk, ll = np.meshgrid(r, r, r, indexing="ij")
    hh, kk, ll = hh.ravel(), kk.ravel(), ll.ravel()
    # Same operation order as basis.zig:47-56: (b1*h + b2*k) + b3*l.
    g = (recip[0][None, :] * hh[:, None].astype(np.float64)
         + recip[1][None, :] * kk[:, None].astype(np.float64)) \
        + recip[2][None, :] * ll[:, None].astype(np.float64)
    kpg = k_cart[None, :] + g
    kin = kpg[:, 0] * kpg[:, 0] + kpg[:, 1] * kpg[:, 1] + kpg[:, 2] * kpg[:, 2]
    sel = kin <= ecut_ry
    return Basis(
        h=hh[sel].astype(np.int32), k=kk[sel].astype(np.int32), l=ll[sel].astype(np.int32),
        cart=g[sel].copy(), kpg=kpg[sel].copy(), kinetic=kin[sel].copy(),
    )


def centered_indices(basis: Basis, grid: tuple[int, int, int]) -> np.ndarray:
    """PwGridMap.init: (h-min_h) + nx*((k-min_k) + ny*(l-min_l)) (pw_grid_map.zig:18-51).

    Raises ValueError (the reference's error.InvalidGrid) when a G-vector falls outside the grid.
    """
    nx, ny, nz = grid
    hi = basis.h.astype(np.int64) - min_index(nx)
    ki = basis.k.astype(np.int64) - min_index(ny)
    li = basis.l.astype(np.int64) - min_index(nz)
    if (hi < 0).any() or (ki < 0).any() or (li < 0).any() or (hi >= nx).any() or (ki >= ny).any() or (li >= nz).any():
        raise ValueError("InvalidGrid: plane wave outside FFT grid")
    return hi + nx * (ki + ny * li)


def build_fft_index_map(grid: tuple[int, int, int]) -> np.ndarray:
    """map[fft_pos] = centred position (fft_grid.zig:407-432)."""
    nx, ny, nz = grid
    out = np.empty(nx * ny * nz, dtype=np.int64)
    idx = 0
    for z in range(nz):
        fl = index_to_freq(z, nz) - min_index(nz)
        for y in range(ny):
            fk = index_to_freq(y, ny) - min_index(ny)
            for x in range(nx):
                fh = index_to_freq(x, nx) - min_index(nx)
                out[idx] = fh + nx * (fk + ny * fl)
                idx += 1
    return out


def fft_indices(basis: Basis, grid: tuple[int, int, int]) -> np.ndarray:
    """PW -> FFT-order position via the inverse of build_fft_index_map (pw_grid_map.zig:58-79)."""
    fmap = build_fft_index_map(grid)
    inv = np.empty_like(fmap)
    inv[fmap] = np.arange(fmap.shape[0], dtype=np.int64)
    return inv[centered_indices(basis, grid)]


def scatter_fft(idx: np.ndarray, coeffs: np.ndarray, total: int, scale: float) -> np.ndarray:
    """Zero the grid, then grid[idx[i]] = coeffs[i]*scale (pw_grid_map.zig:95-105)."""
    out = np.zeros(total, dtype=np.complex128)
    out[idx] = coeffs * scale
    return out


def gather_fft(idx: np.ndarray, grid_data: np.ndarray, scale: float) -> np.ndarray:
    """out[i] = grid[idx[i]]*scale (pw_grid_map.zig:107-117)."""
    return grid_data[idx] * scale
