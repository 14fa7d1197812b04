"""UPF pseudopotential files and the radial tables that feed the projector table (host side).

Reads what the reference reads from a UPF file (src/features/pseudopotential/pseudopotential.zig:
126-340): the PP_HEADER attributes element / z_valence / l_max / mesh_size / number_of_proj,
PP_R, PP_RAB, PP_LOCAL (Ry), every PP_BETA.* (r*beta, angular momentum from `l` or
`angular_momentum`), PP_DIJ (Ry), and the optional PP_NLCC and PP_RHOATOM; tags are matched by
name prefix, floats may use Fortran D exponents, and mesh lengths are checked against
mesh_size (error.InvalidUpf there, UpfError here).

Radial tables (nonlocal.zig:8-21, 248-338): R_l(q) = sum_i r_i (r beta)_i j_l(q r_i) rab_i w_i with
the weights w_i of math/radial_integration.zig:7-13, tabulated at 2048 points
on [0, g_max] and interpolated linearly (0 beyond g_max); w_i is the trapezoid rule with
five-point end corrections.  Local form factor by the short-range
split (form_factor.zig:225-254): V(q) = 4 pi sum_i r_i^2 (V_loc(r_i) + 2 Z / r_i) sinc(q r_i)
rab_i w_i - 8 pi Z / q^2 (Ry).

`write_upf` writes a UPF v2 file for GTH-form species (the synthetic files the tests and the SCF
systems use, since no UPF ships with the reference).
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import tables


class UpfError(ValueError):
    pass


@dataclass
class Beta:
    l: int
    values: np.ndarray  # r * beta(r)


@dataclass
class Upf:
    element: str | None
    z_valence: float
    l_max: int | None
    mesh_size: int | None
    r: np.ndarray
    rab: np.ndarray
    v_local: np.ndarray  # Ry
    betas: list
    dij: np.ndarray  # [nbeta, nbeta] Ry
    nlcc: np.ndarray = field(default_factory=lambda: np.zeros(0))
    rho_atom: np.ndarray = field(default_factory=lambda: np.zeros(0))


_TAG = re.compile(r"<(PP_[A-Za-z0-9_.]+)([^>]*)>(.*?)</\1\s*>", re.S)
_ATTR = re.compile(r"""(?:^|[\s<])([A-Za-z_][A-Za-z0-9_]*)\s*=\s*(?:"([^"]*)"|'([^']*)'|(\S+))""")


def _floats(text: str) -> np.ndarray:
    toks = text.replace("D", "E").replace("d", "e").split()
    try:
        return np.array([float(t) for t in toks], dtype=np.float64)
    except ValueError as e:
        raise UpfError(f"bad number in UPF data: {e}") from None


def _attrs(tag_text: str) -> dict:
    out = {}
    for m in _ATTR.finditer(tag_text):
        out[m.group(1)] = next(g for g in m.groups()[1:] if g is not None)
    return out


def parse_upf(text: str) -> Upf:
    head = re.search(r"<PP_HEADER([^>]*)>", text, re.S)
    if not head:
        raise UpfError("no PP_HEADER")
    h = _attrs(head.group(1))
    element = h.get("element", h.get("atomic_symbol"))
    element = element.strip() if element else None
    z_val = float(h["z_valence"].replace("D", "E")) if "z_valence" in h else 0.0
    l_max = int(h.get("l_max", h.get("lmax"))) if ("l_max" in h or "lmax" in h) else None
    mesh = int(h.get("mesh_size", h.get("mesh"))) if ("mesh_size" in h or "mesh" in h) else None
    tags = {}
    betas = []
    for m in _TAG.finditer(text):
        name, attr, body = m.group(1), m.group(2), m.group(3)
        if name.startswith("PP_BETA"):
            a = _attrs(attr)
            lv = a.get("l", a.get("angular_momentum"))
            betas.append(Beta(int(lv) if lv is not None else 0, _floats(body)))
        elif name in ("PP_MESH", "PP_NONLOCAL"):
            for mm in _TAG.finditer(body):  # nested sections
                nm, at, bd = mm.group(1), mm.group(2), mm.group(3)
                if nm.startswith("PP_BETA"):
                    a = _attrs(at)
                    lv = a.get("l", a.get("angular_momentum"))
                    betas.append(Beta(int(lv) if lv is not None else 0, _floats(bd)))
                else:
                    tags.setdefault(nm, bd)
        else:
            tags.setdefault(name, body)
    for need in ("PP_R", "PP_RAB", "PP_LOCAL"):
        if need not in tags:
            raise UpfError(f"missing {need}")
    r, rab, vloc = _floats(tags["PP_R"]), _floats(tags["PP_RAB"]), _floats(tags["PP_LOCAL"])
    nlcc = _floats(tags["PP_NLCC"]) if "PP_NLCC" in tags else np.zeros(0)
    rho = _floats(tags["PP_RHOATOM"]) if "PP_RHOATOM" in tags else np.zeros(0)
    if mesh is not None:
        if len(r) != mesh or len(rab) != mesh or len(vloc) != mesh:
            raise UpfError("mesh arrays do not have mesh_size entries")
        if len(rho) not in (0, mesh) or len(nlcc) not in (0, mesh):
            raise UpfError("PP_RHOATOM / PP_NLCC length differs from mesh_size")
    nb = len(betas)
    dij = _floats(tags["PP_DIJ"]) if "PP_DIJ" in tags else np.zeros(0)
    if nb and dij.size != nb * nb:
        raise UpfError("PP_DIJ is not number_of_proj squared")
    dij = dij.reshape(nb, nb) if nb else np.zeros((0, 0))
    for b in betas:
        if len(b.values) > len(r):
            raise UpfError("PP_BETA longer than the mesh")
        if len(b.values) < len(r):  # short betas are zero beyond their cutoff index
            b.values = np.concatenate([b.values, np.zeros(len(r) - len(b.values))])
    return Upf(element, z_val, l_max, mesh, r, rab, vloc, betas, dij, nlcc, rho)


def read_upf(path) -> Upf:
    return parse_upf(Path(path).read_text())


# ----------------------------------------------------------------- radial integrals
# End corrections of the corrected trapezoid rule the reference integrates with
# (math/radial_integration.zig:7-13): the first and last five weights, in units of 1/72.
_END_WEIGHTS = np.array([23.75, 95.10, 55.20, 79.30, 70.65]) / 72.0


def ctrap_weights(n: int) -> np.ndarray:
    """Weights 1 inside, the five-point end corrections at both ends; all 1 below 10 points."""
    w = np.ones(n)
    if n >= 10:
        w[:5] = _END_WEIGHTS
        w[-5:] = _END_WEIGHTS[::-1]
    return w


def spherical_bessel(l: int, x: np.ndarray) -> np.ndarray:
    """j_l(x) for l <= 4: closed forms, with the small-x series below 1e-3 (relative error < 1e-12)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.empty_like(x)
    small = np.abs(x) < 1e-3
    xs = x[small]
    # j_l(x) ~ x^l / (2l+1)!! (1 - x^2 / (2 (2l+3)) + x^4 / (8 (2l+3)(2l+5)))
    dfac = float(np.prod(np.arange(1, 2 * l + 2, 2)))
    x2 = xs * xs
    out[small] = xs ** l / dfac * (1.0 - x2 / (2 * (2 * l + 3)) + x2 * x2 / (8 * (2 * l + 3) * (2 * l + 5)))
    xb = x[~small]
    s, c = np.sin(xb), np.cos(xb)
    if l == 0:
        v = s / xb
    elif l == 1:
        v = s / xb ** 2 - c / xb
    elif l == 2:
        v = (3.0 / xb ** 3 - 1.0 / xb) * s - 3.0 * c / xb ** 2
    elif l == 3:
        v = (15.0 / xb ** 4 - 6.0 / xb ** 2) * s - (15.0 / xb ** 3 - 1.0 / xb) * c
    elif l == 4:
        v = (105.0 / xb ** 5 - 45.0 / xb ** 3 + 1.0 / xb) * s - (105.0 / xb ** 4 - 10.0 / xb ** 2) * c
    else:
        raise ValueError("l > 4")
    out[~small] = v
    return out


def radial_projector(beta_r: np.ndarray, r: np.ndarray, rab: np.ndarray, l: int, q) -> np.ndarray:
    """R_l(q) = sum_i r_i (r beta)_i j_l(q r_i) rab_i w_i for an array of q."""
    q = np.atleast_1d(np.asarray(q, dtype=np.float64))
    w = r * beta_r * rab * ctrap_weights(len(r))
    return np.array([float(np.dot(w, spherical_bessel(l, qq * r))) for qq in q])


class RadialTable:
    """n points on [0, g_max], linear interpolation, value at 0 below 0 and 0 at or beyond g_max."""

    def __init__(self, beta_r, r, rab, l, g_max, n_points=2048):
        self.n = max(int(n_points), 2)
        self.g_max = float(g_max)
        self.dg = self.g_max / (self.n - 1)
        self.values = radial_projector(beta_r, r, rab, l, self.dg * np.arange(self.n))

    def __call__(self, g):
        g = np.asarray(g, dtype=np.float64)
        x = g / self.dg
        i = np.clip(np.floor(x).astype(np.int64), 0, self.n - 1)
        t = x - i
        i1 = np.minimum(i + 1, self.n - 1)
        v = self.values[i] * (1.0 - t) + self.values[i1] * t
        v = np.where(i + 1 >= self.n, self.values[self.n - 1], v)
        v = np.where(g <= 0.0, self.values[0], v)
        return np.where(g >= self.g_max, 0.0, v)


def local_form_factor(upf: Upf, q) -> np.ndarray:
    """V_loc(q) in Ry by the short-range split; the G = 0 value is the short-range part alone."""
    q = np.atleast_1d(np.asarray(q, dtype=np.float64))
    r, rab, w = upf.r, upf.rab, ctrap_weights(len(upf.r))
    vsr = upf.v_local + np.where(r < 1e-10, 0.0, 2.0 * upf.z_valence / np.where(r < 1e-10, 1.0, r))
    base = r * r * vsr * rab * w
    out = np.empty_like(q)
    for j, qq in enumerate(q):
        x = qq * r
        sinc = np.where(np.abs(x) < 1e-12, 1.0, np.sin(x) / np.where(np.abs(x) < 1e-12, 1.0, x))
        v = 4.0 * math.pi * float(np.dot(base, sinc))
        out[j] = v if qq < 1e-10 else v - 8.0 * math.pi * upf.z_valence / (qq * qq)
    return out


@dataclass
class UpfSpecies:
    """A species for dzhip.Hamiltonian built from a UPF file: l_list, D (Ry) and table-based
    radial functions (g_max from the basis' largest |k+G|, 2048 points as the reference)."""
    upf: Upf
    g_max: float

    @property
    def l_list(self):
        return [b.l for b in self.upf.betas]

    @property
    def dij(self):
        return self.upf.dij

    @property
    def m_total(self):
        return sum(2 * b.l + 1 for b in self.upf.betas)

    def radial_fns(self):
        return [RadialTable(b.values, self.upf.r, self.upf.rab, b.l, self.g_max) for b in self.upf.betas]


# ----------------------------------------------------------------- synthetic files
def log_mesh(n=1200, xmin=-8.0, dx=0.0125, zmesh=1.0):
    """r_i = e^{xmin + i dx} / zmesh, rab_i = r_i dx (the usual UPF logarithmic mesh)."""
    r = np.exp(xmin + dx * np.arange(n)) / zmesh
    return r, r * dx


def gth_local_r(r, z, rloc, c1):
    """GTH local potential in Ry: 2 [-Z erf(r / (sqrt2 rloc)) / r + C1 e^{-r^2 / (2 rloc^2)}]."""
    from scipy.special import erf

    return 2.0 * (-z * erf(r / (math.sqrt(2.0) * rloc)) / r + c1 * np.exp(-0.5 * (r / rloc) ** 2))


def write_upf(path, element, z_valence, rloc, c1, l_list, rc_list, dij_ry, n_mesh=1200):
    """UPF v2 file of a GTH-form local part and unit-normalised Gaussian projectors
    beta_l(r) = N r^l e^{-r^2 / (2 rc^2)} (stored as r beta)."""
    r, rab = log_mesh(n_mesh)
    vloc = gth_local_r(r, z_valence, rloc, c1)

    def block(name, arr, attrs=""):
        body = "\n".join(" ".join(f"{v: .16E}" for v in arr[i:i + 4]) for i in range(0, len(arr), 4))
        return f"    <{name}{attrs}>\n{body}\n    </{name}>\n"

    lines = ['<UPF version="2.0.1">\n',
             f'  <PP_HEADER element="{element}" pseudo_type="NC" z_valence="{z_valence:.6f}" '
             f'l_max="{max(l_list) if l_list else 0}" mesh_size="{n_mesh}" number_of_proj="{len(l_list)}"/>\n',
             "  <PP_MESH>\n", block("PP_R", r), block("PP_RAB", rab), "  </PP_MESH>\n",
             block("PP_LOCAL", vloc), "  <PP_NONLOCAL>\n"]
    for i, (l, rc) in enumerate(zip(l_list, rc_list)):
        norm = 1.0 / math.sqrt(math.gamma(l + 1.5) * rc ** (2 * l + 3) / 2.0)
        beta = norm * r ** l * np.exp(-0.5 * (r / rc) ** 2)
        lines.append(block(f"PP_BETA.{i + 1}", r * beta, f' index="{i + 1}" angular_momentum="{l}"'))
    lines.append(block("PP_DIJ", np.asarray(dij_ry, dtype=np.float64).ravel()))
    lines.append("  </PP_NONLOCAL>\n</UPF>\n")
    Path(path).write_text("".join(lines))
