"""Caller-side tables for one k-point, vectorised (host, FP64/int).

In dft-zig these are built by the SCF driver before it calls the H.psi
callback, and they are what the HIP descriptor (``dzhip_ham_desc``) carries:
  G-sphere (h, k, l, G, k+G, |k+G|^2)   src/features/plane_wave/basis.zig:26-74
  PW -> FFT-order index                 src/features/scf/pw_grid_map.zig:18-79,
                                        src/features/scf/fft_grid.zig:407-438
  grid sizing                           src/features/scf/pw_grid.zig:54-72, src/lib/fft/sizing.zig:1-15
  phi[(beta,m), G] = 4 pi R_beta(|k+G|) Y_lm(k+G)
                                        src/features/scf/nonlocal_context.zig:437-549,
                                        src/features/pseudopotential/nonlocal.zig:152-241
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

ANGSTROM_TO_BOHR = 1.0 / 0.52917721067


def reciprocal(cell):
    c = np.asarray(cell, dtype=np.float64)
    vol = float(np.dot(c[0], np.cross(c[1], c[2])))
    s = 2.0 * math.pi / vol
    return np.stack([np.cross(c[1], c[2]) * s, np.cross(c[2], c[0]) * s, np.cross(c[0], c[1]) * s])


def next_fft_size(n: int) -> int:
    if n <= 1:
        return 1
    while True:
        m = n
        for f in (2, 3, 5):
            while m % f == 0:
                m //= f
        if m == 1:
            return n
        n += 1


def auto_grid(ecut_ry, recip, grid_scale=1.0):
    gmax = max(2.0, grid_scale if grid_scale > 0 else 1.0) * math.sqrt(ecut_ry)
    return tuple(next_fft_size(max(int(math.ceil(gmax / float(np.linalg.norm(recip[i])))) * 2 + 1, 3))
                 for i in range(3))


@dataclass
class Basis:
    h: np.ndarray
    k: np.ndarray
    l: np.ndarray
    cart: np.ndarray
    kpg: np.ndarray
    kinetic: np.ndarray

    @property
    def n(self) -> int:
        return int(self.h.shape[0])


def generate_basis(recip, ecut_ry, k_cart) -> Basis:
    recip = np.asarray(recip, dtype=np.float64)
    k_cart = np.asarray(k_cart, dtype=np.float64)
    min_b = min(float(np.linalg.norm(recip[i])) for i in range(3))
    max_n = int(math.ceil(math.sqrt(ecut_ry) / min_b)) + 1
    r = np.arange(-max_n, max_n + 1, dtype=np.int64)
    hh, kk, ll = (a.ravel() for a in np.meshgrid(r, r, r, indexing="ij"))
    g = (recip[0] * hh[:, None].astype(np.float64) + recip[1] * kk[:, None].astype(np.float64)) \
        + recip[2] * ll[:, None].astype(np.float64)
    kpg = k_cart + g
    kin = kpg[:, 0] * kpg[:, 0] + kpg[:, 1] * kpg[:, 1] + kpg[:, 2] * kpg[:, 2]
    s = kin <= ecut_ry
    return Basis(hh[s].astype(np.int32), kk[s].astype(np.int32), ll[s].astype(np.int32),
                 g[s].copy(), kpg[s].copy(), kin[s].copy())


def fft_indices(basis: Basis, grid) -> np.ndarray:
    """FFT-order position of each plane wave: negative frequencies wrap to n + f."""
    nx, ny, nz = grid
    for arr, n in ((basis.h, nx), (basis.k, ny), (basis.l, nz)):
        if arr.size and (arr.min() < -(n // 2) or arr.max() > (n - 1) // 2):
            raise ValueError("InvalidGrid: plane wave outside FFT grid")
    x = np.mod(basis.h.astype(np.int64), nx)
    y = np.mod(basis.k.astype(np.int64), ny)
    z = np.mod(basis.l.astype(np.int64), nz)
    return (x + nx * (y + ny * z)).astype(np.int32)


def real_ylm(l: int, v: np.ndarray) -> np.ndarray:
    """[2l+1, n] tesseral harmonics at directions v (rows), m = -l..l; Y_00 at the origin."""
    pi = math.pi
    r2 = np.sum(v * v, axis=1)
    zero = r2 < 1e-30
    inv = np.where(zero, 0.0, 1.0 / np.sqrt(np.where(zero, 1.0, r2)))
    x, y, z = v[:, 0] * inv, v[:, 1] * inv, v[:, 2] * inv
    if l == 0:
        return np.full((1, v.shape[0]), 1.0 / math.sqrt(4.0 * pi))
    if l == 1:
        c = math.sqrt(3.0 / (4.0 * pi))
        out = np.stack([c * y, c * z, c * x])
    elif l == 2:
        c0, c1, c2 = math.sqrt(5 / (16 * pi)), math.sqrt(15 / (4 * pi)), math.sqrt(15 / (16 * pi))
        out = np.stack([c2 * 2 * x * y, c1 * y * z, c0 * (3 * z * z - 1), c1 * x * z, c2 * (x * x - y * y)])
    elif l == 3:
        out = np.stack([
            math.sqrt(35 / (32 * pi)) * (3 * x * x - y * y) * y,
            math.sqrt(105 / (4 * pi)) * x * y * z,
            math.sqrt(21 / (32 * pi)) * y * (5 * z * z - 1),
            math.sqrt(7 / (16 * pi)) * (5 * z ** 3 - 3 * z),
            math.sqrt(21 / (32 * pi)) * x * (5 * z * z - 1),
            math.sqrt(105 / (16 * pi)) * (x * x - y * y) * z,
            math.sqrt(35 / (32 * pi)) * (x * x - 3 * y * y) * x])
    elif l == 4:
        x2, y2, z2 = x * x, y * y, z * z
        out = np.stack([
            math.sqrt(315 / (256 * pi)) * 4 * x * y * (x2 - y2),
            math.sqrt(315 / (32 * pi)) * y * z * (3 * x2 - y2),
            math.sqrt(45 / (64 * pi)) * 2 * x * y * (7 * z2 - 1),
            math.sqrt(45 / (32 * pi)) * y * z * (7 * z2 - 3),
            math.sqrt(9 / (256 * pi)) * (35 * z2 * z2 - 30 * z2 + 3),
            math.sqrt(45 / (32 * pi)) * x * z * (7 * z2 - 3),
            math.sqrt(45 / (64 * pi)) * (x2 - y2) * (7 * z2 - 1),
            math.sqrt(315 / (32 * pi)) * x * z * (x2 - 3 * y2),
            math.sqrt(315 / (256 * pi)) * (x2 * x2 - 6 * x2 * y2 + y2 * y2)])
    else:
        raise ValueError("l > 4 not supported")
    out[:, zero] = 0.0
    return out


def build_phi(kpg: np.ndarray, l_list, radial_fns) -> np.ndarray:
    """[m_total, npw] real projector table, rows (offset_beta + m), m = -l..l."""
    q = np.sqrt(np.sum(kpg * kpg, axis=1))
    rows = []
    for l, fn in zip(l_list, radial_fns):
        rad = fn(q)
        rows.append(4.0 * math.pi * rad[None, :] * real_ylm(l, kpg))
    return np.concatenate(rows, axis=0) if rows else np.zeros((0, kpg.shape[0]))


def gaussian_radial(l: int, rc: float):
    """Vectorised closed-form R_l(q) for the unit-normalised Gaussian beta_l = N r^l e^{-r^2/(2 rc^2)}."""
    norm = 1.0 / math.sqrt(math.gamma(l + 1.5) * rc ** (2 * l + 3) / 2.0)
    c = norm * math.sqrt(math.pi / 2.0) * rc ** (2 * l + 3)
    return lambda q: c * np.power(q, l) * np.exp(-0.5 * q * q * rc * rc)
