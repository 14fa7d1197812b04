"""Builds an oracle HamContext from a synthetic workload's raw inputs (duck-typed).

Only the raw inputs are taken (basis vectors, local_r, species l/rc/D, atoms);
every table (radial transforms, Y_lm, phi, index maps) is rebuilt with the
oracle's own restated functions, independently of the product's host tables.
"""
from __future__ import annotations

from . import hamiltonian, projectors, pw


def context_from_workload(w) -> hamiltonian.HamContext:
    sps = []
    for sp in w.species:
        rad = [projectors.gaussian_beta_radial(l, rc, projectors.gaussian_beta_norm(l, rc))
               for l, rc in zip(sp.l_list, sp.rc)]
        sps.append(projectors.Species(list(sp.l_list), rad, sp.dij))
    b = pw.Basis(w.basis.h, w.basis.k, w.basis.l, w.basis.cart, w.basis.kpg, w.basis.kinetic)
    atoms = [hamiltonian.Atom(w.atom_pos[i], int(w.atom_species[i])) for i in range(len(w.atom_species))]
    return hamiltonian.HamContext(tuple(w.grid), b, w.local_r, sps, atoms, w.inv_volume)


def gth_local_q(z_ion: float, rloc: float, c1: float):
    """v(q) (Ry, per atom, without 1/Omega) of the local potential
    V(r) = 2 [ -Z erf(r / (sqrt2 rloc)) / r + C1 exp(-r^2 / (2 rloc^2)) ]  (Hartree -> Ry):
    v(q) = 2 [ -4 pi Z / q^2 + C1 (2 pi)^{3/2} rloc^3 ] exp(-q^2 rloc^2 / 2),  q > 0."""
    import math

    def fn(q: float) -> float:
        e = math.exp(-0.5 * q * q * rloc * rloc)
        return 2.0 * (-4.0 * math.pi * z_ion / (q * q) + c1 * (2.0 * math.pi) ** 1.5 * rloc ** 3) * e

    return fn


# Radial tables of UPF species, built once per (file data, g_max): each is 2048 radial integrals.
_UPF_TABLES = {}


def upf_species(upf, g_max) -> projectors.Species:
    """Oracle species from a UPF file's raw arrays: one RadialTable per beta over [0, g_max]
    (nonlocal.zig:250-300, as fill_nonlocal_phi reads them), D from PP_DIJ."""
    key = (id(upf), float(g_max))
    if key not in _UPF_TABLES:
        tabs = [projectors.RadialTable.build(b.values, upf.r, upf.rab, b.l, g_max) for b in upf.betas]
        _UPF_TABLES[key] = (upf, tabs)
    tabs = _UPF_TABLES[key][1]
    return projectors.Species([b.l for b in upf.betas], [t.eval for t in tabs], upf.dij)


def upf_local_q(upf):
    """v(q) (Ry, per atom, without 1/Omega) of a UPF file's local part by the short-range split."""
    return lambda q: projectors.local_vq_short_range(upf.v_local, upf.r, upf.rab, upf.z_valence, q)


def context_for_k(system, k_cart, local_r) -> hamiltonian.HamContext:
    """Oracle context of one k-point of an SCF system (raw parameters only; species read from
    a UPF file carry `upf` and `g_max`)."""
    basis = pw.generate_basis(system.recip, system.ecut_ry, k_cart)
    sps = []
    for sp in system.species:
        if hasattr(sp, "upf"):
            sps.append(upf_species(sp.upf, sp.g_max))
            continue
        rad = [projectors.gaussian_beta_radial(l, rc, projectors.gaussian_beta_norm(l, rc))
               for l, rc in zip(sp.l_list, sp.rc)]
        sps.append(projectors.Species(list(sp.l_list), rad, sp.dij))
    atoms = [hamiltonian.Atom(system.atom_pos[i], int(system.atom_species[i]))
             for i in range(len(system.atom_species))]
    return hamiltonian.HamContext(tuple(system.grid), basis, local_r, sps, atoms, 1.0 / system.volume)
