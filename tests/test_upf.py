"""UPF reading and the radial tables that feed the projector table (dzhip/upf.py) against the
oracle's radial integrals (oracle/projectors.py) and the closed forms of the Gaussian/GTH forms
the synthetic files hold."""
import math

import numpy as np
import pytest

from dzhip import scf as dscf
from dzhip import tables, upf
from oracle import adapters
from oracle import projectors as oproj


@pytest.fixture(scope="module")
def si_file(tmp_path_factory):
    g = dscf.SI_GTH
    p = tmp_path_factory.mktemp("upf") / "Si.gth.upf"
    upf.write_upf(p, "Si", g["z"], g["rloc"], g["c1"], [0, 1], [g["r_s"], g["r_p"]],
                  np.diag([2 * g["h_s"], 2 * g["h_p"]]))
    return p


def test_read_fields(si_file):
    u = upf.read_upf(si_file)
    assert u.element == "Si" and u.z_valence == 4.0 and u.l_max == 1 and u.mesh_size == 1200
    assert len(u.r) == len(u.rab) == len(u.v_local) == 1200
    assert [b.l for b in u.betas] == [0, 1]
    assert np.allclose(u.dij, np.diag([2 * dscf.SI_GTH["h_s"], 2 * dscf.SI_GTH["h_p"]]))


def test_ctrap_weights_match_oracle():
    for n in (3, 9, 10, 11, 50, 1200):
        assert np.array_equal(upf.ctrap_weights(n), oproj.ctrap_weights(n))


def test_radial_projector_matches_oracle_and_closed_form(si_file):
    u = upf.read_upf(si_file)
    for b, rc in zip(u.betas, [dscf.SI_GTH["r_s"], dscf.SI_GTH["r_p"]]):
        qs = np.array([0.0, 0.3, 1.1, 2.7, 5.0, 8.5])
        mine = upf.radial_projector(b.values, u.r, u.rab, b.l, qs)
        ref = np.array([oproj.radial_projector(b.values, u.r, u.rab, b.l, q) for q in qs])
        assert np.max(np.abs(mine - ref)) < 1e-13 * max(1.0, np.max(np.abs(ref)))
        exact = tables.gaussian_radial(b.l, rc)(qs)
        assert np.max(np.abs(mine - exact)) < 1e-7 * np.max(np.abs(exact))


def test_radial_table_interpolation(si_file):
    u = upf.read_upf(si_file)
    b = u.betas[1]
    t = upf.RadialTable(b.values, u.r, u.rab, b.l, 9.0)
    ot = oproj.RadialTable.build(b.values, u.r, u.rab, b.l, 9.0, 2048)
    qs = np.linspace(0.0, 9.2, 301)
    assert np.max(np.abs(t(qs) - np.array([ot.eval(q) for q in qs]))) < 1e-13
    assert t(np.array([9.0, 9.5])).tolist() == [0.0, 0.0]
    exact = tables.gaussian_radial(1, dscf.SI_GTH["r_p"])(qs[qs < 9.0])
    assert np.max(np.abs(t(qs[qs < 9.0]) - exact)) < 2e-6 * np.max(np.abs(exact))


def test_local_form_factor_vs_gth_closed_form(si_file):
    u = upf.read_upf(si_file)
    g = dscf.SI_GTH
    vq = adapters.gth_local_q(g["z"], g["rloc"], g["c1"])
    qs = np.array([0.2, 0.7, 1.5, 3.0, 6.0])
    mine = upf.local_form_factor(u, qs)
    exact = np.array([vq(q) for q in qs])
    assert np.max(np.abs(mine - exact)) < 1e-6 * np.max(np.abs(exact))


def test_phi_from_upf_equals_closed_form_species(si_file):
    s = dscf.si8_small()
    b = tables.generate_basis(s.recip, s.ecut_ry, s.kpoints[1])
    u = upf.read_upf(si_file)
    gmax = float(np.max(np.linalg.norm(b.kpg, axis=1))) * 1.05
    sp_u = upf.UpfSpecies(u, gmax)
    phi_u = tables.build_phi(b.kpg, sp_u.l_list, sp_u.radial_fns())
    sp_g = s.species[0]
    phi_g = tables.build_phi(b.kpg, sp_g.l_list, sp_g.radial_fns())
    assert np.max(np.abs(phi_u - phi_g)) < 2e-6 * np.max(np.abs(phi_g))


def test_parser_errors_and_fortran_exponents(tmp_path):
    good = """<UPF version="2.0.1"><PP_HEADER element="X" z_valence="1.0D0" mesh_size="3"/>
<PP_MESH><PP_R> 1.0D-1 2.0d-1 3.0E-1 </PP_R><PP_RAB> 1 1 1 </PP_RAB></PP_MESH>
<PP_LOCAL> -1 -2 -3 </PP_LOCAL></UPF>"""
    u = upf.parse_upf(good)
    assert np.allclose(u.r, [0.1, 0.2, 0.3]) and u.z_valence == 1.0 and u.betas == []
    with pytest.raises(upf.UpfError):
        upf.parse_upf(good.replace("<PP_LOCAL> -1 -2 -3 </PP_LOCAL>", ""))
    with pytest.raises(upf.UpfError):
        upf.parse_upf(good.replace('mesh_size="3"', 'mesh_size="4"'))
    with pytest.raises(upf.UpfError):
        upf.parse_upf("<UPF></UPF>")
