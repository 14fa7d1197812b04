"""Kleinman-Bylander projector tables phi[(beta,m), G] = 4*pi*R_beta(|k+G|)*Y_lm(k+G) (FP64).

Restates (paths relative to the dft-zig source tree):
  radial_projector          src/features/pseudopotential/nonlocal.zig:8-21
  angular_factor, legendre  src/features/pseudopotential/nonlocal.zig:24-54
  spherical_bessel (+Miller, small-x series)  nonlocal.zig:59-147
  real_spherical_harmonic   nonlocal.zig:152-241 (l <= 4, m ordered -l..l)
  RadialTable               nonlocal.zig:250-300 (2048 points, linear interpolation, 0 beyond g_max)
  ctrap_weight              src/features/math/radial_integration.zig:7-13
  local_vq_short_range      src/features/pseudopotential/form_factor.zig:225-256
  fill_nonlocal_radial_terms / fill_nonlocal_phi
                            src/features/scf/nonlocal_context.zig:437-477, 524-549
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_CTRAP = np.array([23.75, 95.10, 55.20, 79.30, 70.65]) / 72.0


def ctrap_weight(i: int, n: int) -> float:
    if n < 10:
        return 1.0
    if i < 5:
        return float(_CTRAP[i])
    if i >= n - 5:
        return float(_CTRAP[n - 1 - i])
    return 1.0


def ctrap_weights(n: int) -> np.ndarray:
    return np.array([ctrap_weight(i, n) for i in range(n)])


def legendre_p(l: int, x: float) -> float:
    if l == 0:
        return 1.0
    if l == 1:
        return x
    if l == 2:
        return 0.5 * (3.0 * x * x - 1.0)
    if l == 3:
        return 0.5 * (5.0 * x * x * x - 3.0 * x)
    pm2, pm1 = 1.0, x
    for n in range(1, l):
        pn = ((2.0 * n + 1.0) * x * pm1 - n * pm2) / (n + 1.0)
        pm2, pm1 = pm1, pn
    return pm1


def angular_factor(l: int, cos_gamma: float) -> float:
    """4*pi*(2l+1)*P_l(cos gamma) (nonlocal.zig:24-27)."""
    return 4.0 * math.pi * (2.0 * l + 1.0) * legendre_p(l, cos_gamma)


def _bessel_small(l: int, x: float) -> float:
    x2 = x * x
    if l == 0:
        return 1.0 - x2 / 6.0 + x2 * x2 / 120.0
    if l == 1:
        return x / 3.0 - x * x2 / 30.0
    if l == 2:
        return x2 / 15.0 - x2 * x2 / 210.0
    if l == 3:
        return x * x2 / 105.0
    return 0.0


def _bessel_direct(l: int, x: float) -> float:
    if l > x + 50.0:
        return 0.0
    jm1 = math.sin(x) / x
    j = math.sin(x) / (x * x) - math.cos(x) / x
    for n in range(1, l):
        jp1 = (2.0 * n + 1.0) / x * j - jm1
        jm1, j = j, jp1
        if not math.isfinite(j):
            return 0.0
    return j


def _bessel_miller(l: int, x: float) -> float:
    nm = l + 1
    nlim = nm + int(15.0 * x) + 1 if x < 1.0 else nm + int(1.36 * x) + 15
    if nlim > 256:
        return _bessel_direct(l, x)
    sb = [0.0] * (nlim + 2)
    xi = 1.0 / x
    sb[nlim + 1] = 0.0
    sb[nlim] = 1.0e-18
    for n in range(nlim, 0, -1):
        sb[n - 1] = (2.0 * n + 1.0) * xi * sb[n] - sb[n + 1]
    scale = (math.sin(x) / x) / sb[0]
    return sb[nm - 1] * scale


def spherical_bessel(l: int, x: float) -> float:
    if abs(x) < 1e-6:
        return _bessel_small(l, x)
    if l == 0:
        return math.sin(x) / x
    if l == 1:
        return math.sin(x) / (x * x) - math.cos(x) / x
    return _bessel_miller(l, x)


def radial_projector(beta: np.ndarray, r: np.ndarray, rab: np.ndarray, l: int, g: float) -> float:
    """R(g) = sum_i r_i * (r beta)_i * j_l(g r_i) * rab_i * w_i (nonlocal.zig:8-21)."""
    n = beta.shape[0]
    w = ctrap_weights(n)
    s = 0.0
    for i in range(n):
        s += r[i] * beta[i] * spherical_bessel(l, g * r[i]) * rab[i] * w[i]
    return s


def local_vq_short_range(v_local: np.ndarray, r: np.ndarray, rab: np.ndarray, z_valence: float, q: float) -> float:
    """V_loc(q) in Ry by the short-range split (form_factor.zig:225-256): 4 pi sum_i r_i^2
    (V_loc(r_i) + 2Z/r_i) sinc(q r_i) rab_i w_i (the Coulomb term skipped at r < 1e-10), minus
    8 pi Z / q^2 for q >= 1e-10; at q < 1e-10 the short-range part alone."""
    n = r.shape[0]
    w = ctrap_weights(n)
    s = 0.0
    for i in range(n):
        vc = 0.0 if r[i] < 1e-10 else 2.0 * z_valence / r[i]
        x = q * r[i]
        sinc = 1.0 if abs(x) < 1e-12 else math.sin(x) / x
        s += r[i] * r[i] * (v_local[i] + vc) * sinc * rab[i] * w[i]
    v = 4.0 * math.pi * s
    return v if q < 1e-10 else v - 8.0 * math.pi * z_valence / (q * q)


@dataclass
class RadialTable:
    """Uniform table on [0, g_max], linear interpolation, 0 at/after g_max (nonlocal.zig:250-300)."""

    values: np.ndarray
    g_max: float
    dg: float

    @classmethod
    def build(cls, beta, r, rab, l, g_max, n_points=2048):
        n = max(n_points, 2)
        dg = g_max / (n - 1)
        vals = np.array([radial_projector(beta, r, rab, l, dg * i) for i in range(n)])
        return cls(vals, g_max, dg)

    @classmethod
    def from_function(cls, fn, g_max, n_points=2048):
        n = max(n_points, 2)
        dg = g_max / (n - 1)
        return cls(np.array([fn(dg * i) for i in range(n)]), g_max, dg)

    def eval(self, g: float) -> float:
        if g <= 0.0:
            return float(self.values[0])
        if g >= self.g_max:
            return 0.0
        x = g / self.dg
        i = int(x)
        if i + 1 >= self.values.shape[0]:
            return float(self.values[-1])
        t = x - i
        return float(self.values[i] * (1.0 - t) + self.values[i + 1] * t)


def real_spherical_harmonic(l: int, m: int, x: float, y: float, z: float) -> float:
    """Tesseral harmonics, l <= 4 (nonlocal.zig:152-241)."""
    pi = math.pi
    r2 = x * x + y * y + z * z
    if r2 < 1e-30:
        return 1.0 / math.sqrt(4.0 * pi) if l == 0 else 0.0
    inv_r = 1.0 / math.sqrt(r2)
    nx, ny, nz = x * inv_r, y * inv_r, z * inv_r
    if l == 0:
        return 1.0 / math.sqrt(4.0 * pi) if m == 0 else 0.0
    if l == 1:
        c = math.sqrt(3.0 / (4.0 * pi))
        return {-1: c * ny, 0: c * nz, 1: c * nx}.get(m, 0.0)
    if l == 2:
        c0 = math.sqrt(5.0 / (16.0 * pi))
        c1 = math.sqrt(15.0 / (4.0 * pi))
        c2 = math.sqrt(15.0 / (16.0 * pi))
        return {-2: c2 * 2.0 * nx * ny, -1: c1 * ny * nz, 0: c0 * (3.0 * nz * nz - 1.0),
                1: c1 * nx * nz, 2: c2 * (nx * nx - ny * ny)}.get(m, 0.0)
    if l == 3:
        return {
            -3: math.sqrt(35.0 / (32.0 * pi)) * (3.0 * nx * nx - ny * ny) * ny,
            -2: math.sqrt(105.0 / (4.0 * pi)) * nx * ny * nz,
            -1: math.sqrt(21.0 / (32.0 * pi)) * ny * (5.0 * nz * nz - 1.0),
            0: math.sqrt(7.0 / (16.0 * pi)) * (5.0 * nz * nz * nz - 3.0 * nz),
            1: math.sqrt(21.0 / (32.0 * pi)) * nx * (5.0 * nz * nz - 1.0),
            2: math.sqrt(105.0 / (16.0 * pi)) * (nx * nx - ny * ny) * nz,
            3: math.sqrt(35.0 / (32.0 * pi)) * (nx * nx - 3.0 * ny * ny) * nx,
        }.get(m, 0.0)
    if l == 4:
        nz2, nx2, ny2 = nz * nz, nx * nx, ny * ny
        return {
            -4: math.sqrt(315.0 / (256.0 * pi)) * 4.0 * nx * ny * (nx2 - ny2),
            -3: math.sqrt(315.0 / (32.0 * pi)) * ny * nz * (3.0 * nx2 - ny2),
            -2: math.sqrt(45.0 / (64.0 * pi)) * 2.0 * nx * ny * (7.0 * nz2 - 1.0),
            -1: math.sqrt(45.0 / (32.0 * pi)) * ny * nz * (7.0 * nz2 - 3.0),
            0: math.sqrt(9.0 / (256.0 * pi)) * (35.0 * nz2 * nz2 - 30.0 * nz2 + 3.0),
            1: math.sqrt(45.0 / (32.0 * pi)) * nx * nz * (7.0 * nz2 - 3.0),
            2: math.sqrt(45.0 / (64.0 * pi)) * (nx2 - ny2) * (7.0 * nz2 - 1.0),
            3: math.sqrt(315.0 / (32.0 * pi)) * nx * nz * (nx2 - 3.0 * ny2),
            4: math.sqrt(315.0 / (256.0 * pi)) * (nx2 * nx2 - 6.0 * nx2 * ny2 + ny2 * ny2),
        }.get(m, 0.0)
    return 0.0


def ylm_array(l: int, m: int, v: np.ndarray) -> np.ndarray:
    return np.array([real_spherical_harmonic(l, m, *row) for row in v])


@dataclass
class Species:
    """One KB species: per-beta l and radial function R_beta(q) (Ry units), D_ij (Ry).

    ``radial`` holds callables q -> R(q) (a RadialTable.eval or a closed form).
    ``dij`` is row-major nbeta x nbeta (UPF PP_DIJ); only same-l pairs act
    (apply.zig:940-966).
    """

    l_list: list
    radial: list
    dij: np.ndarray
    name: str = "X"
    m_offsets: list = field(default_factory=list)
    m_counts: list = field(default_factory=list)

    def __post_init__(self):
        off = 0
        self.m_offsets, self.m_counts = [], []
        for l in self.l_list:
            self.m_offsets.append(off)
            self.m_counts.append(2 * l + 1)
            off += 2 * l + 1
        self.dij = np.asarray(self.dij, dtype=np.float64).reshape(len(self.l_list), len(self.l_list))

    @property
    def m_total(self) -> int:
        return sum(self.m_counts)

    def radial_values(self, kpg: np.ndarray) -> np.ndarray:
        """[nbeta, npw] radial terms at |k+G| (nonlocal_context.zig:437-477)."""
        gm = np.linalg.norm(kpg, axis=1)
        return np.array([[fn(float(q)) for q in gm] for fn in self.radial])

    def phi(self, kpg: np.ndarray) -> np.ndarray:
        """[m_total, npw] phi = 4*pi*R*Y_lm(k+G), rows offset+m, m = -l..l (nonlocal_context.zig:524-549)."""
        rad = self.radial_values(kpg)
        out = np.zeros((self.m_total, kpg.shape[0]))
        for b, l in enumerate(self.l_list):
            for mi in range(2 * l + 1):
                y = ylm_array(l, mi - l, kpg)
                out[self.m_offsets[b] + mi] = 4.0 * math.pi * rad[b] * y
        return out

    def dij_m(self) -> np.ndarray:
        """Dense m-resolved [m_total, m_total] equivalent of the per-l radial D (apply.zig:940-966)."""
        mt = self.m_total
        d = np.zeros((mt, mt))
        for b, lb in enumerate(self.l_list):
            for j, lj in enumerate(self.l_list):
                if lj != lb:
                    continue
                for mi in range(2 * lb + 1):
                    d[self.m_offsets[b] + mi, self.m_offsets[j] + mi] = self.dij[b, j]
        return d


def gaussian_beta_radial(l: int, rc: float, norm: float = 1.0):
    """Closed-form R_l(q) for beta_l(r) = norm * r^l * exp(-r^2/(2 rc^2)).

    int_0^inf r^2 beta_l(r) j_l(qr) dr = norm * sqrt(pi/2) * rc^(2l+3) * q^l * exp(-q^2 rc^2/2)
    (Gaussian-Hankel integral).  UPF stores r*beta, and radial_projector
    integrates r*(r beta)*j_l, so the same number comes from the quadrature.
    """
    c = norm * math.sqrt(math.pi / 2.0) * rc ** (2 * l + 3)

    def fn(q: float) -> float:
        return c * q ** l * math.exp(-0.5 * q * q * rc * rc)

    return fn


def gaussian_beta_norm(l: int, rc: float) -> float:
    """Normalisation so that int r^2 beta_l(r)^2 dr = 1 for beta_l = norm r^l e^{-r^2/(2rc^2)}."""
    # int_0^inf r^(2l+2) e^{-r^2/rc^2} dr = Gamma(l+3/2) rc^(2l+3) / 2
    val = math.gamma(l + 1.5) * rc ** (2 * l + 3) / 2.0
    return 1.0 / math.sqrt(val)
