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
