"""The oracle's restatement of the reference's short-range local form factor
(oracle/projectors.py::local_vq_short_range, form_factor.zig:225-256) against the GTH closed form
the synthetic file holds, and the product's ionic potential from a UPF file (dzhip.scf.with_upf)
against the oracle's on the Si-8 grid."""
from pathlib import Path

import numpy as np

from dzhip import scf as dscf
from dzhip import upf
from oracle import adapters
from oracle import projectors as oproj
from oracle import scf as oscf

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_oracle_local_form_factor_vs_gth_closed_form():
    u = upf.read_upf(GOLDEN / "si_gth.upf")
    g = dscf.SI_GTH
    vq = adapters.gth_local_q(g["z"], g["rloc"], g["c1"])
    for q in (0.2, 0.7, 1.5, 3.0, 6.0):
        mine = oproj.local_vq_short_range(u.v_local, u.r, u.rab, u.z_valence, q)
        assert abs(mine - vq(q)) < 1e-6 * abs(vq(0.2))
    # q = 0: the short-range part alone (finite).
    assert np.isfinite(oproj.local_vq_short_range(u.v_local, u.r, u.rab, u.z_valence, 0.0))


def test_ionic_potential_from_upf_matches_oracle():
    u = upf.read_upf(GOLDEN / "si_gth.upf")
    s = dscf.with_upf(dscf.si8_small(), [u])
    gv, g2 = dscf.g_vectors(s.grid, s.recip)
    mine = dscf.ionic_potential(s, gv, g2)
    ref = oscf.ionic_potential_r(s.grid, s.recip, s.volume, s.atom_pos, adapters.upf_local_q(u))
    assert np.max(np.abs(mine - ref)) < 1e-12 * np.max(np.abs(ref))
    # and close to the closed-form species' potential (the file's mesh integrals)
    closed = dscf.ionic_potential(dscf.si8_small(), gv, g2)
    assert np.max(np.abs(mine - closed)) < 1e-5 * np.max(np.abs(closed))
    # the radial tables cover every plane wave of the cutoff: g_max = sqrt(2 ecut) 1.5
    assert abs(s.species[0].g_max - np.sqrt(2 * s.ecut_ry) * 1.5) < 1e-15
