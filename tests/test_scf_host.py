"""Host side of the SCF layer: the oracle's potentials and mixing pinned by closed forms, and the
GPU driver's host functions (dzhip.scf) against the oracle.

Closed forms: Hartree potential of a single cosine density (V = 8 pi A / G^2 cos, E_H =
Omega 4 pi A^2 / G^2 ... in Ry); the GTH local form factor against a numerical radial Fourier
transform of the real-space GTH potential; the PZ81 potential against the numerical derivative
of its energy density; Pulay on a linear fixed-point map returns the exact fixed point once the
history spans the space.  The oracle SCF reproduces its golden file and, at self-consistency,
the reference's energy assembly band - E_H - int v_xc rho + E_xc (energy.zig:372-396).
"""
import math
from pathlib import Path

import numpy as np
import pytest
from scipy import integrate, special

from dzhip import scf as dscf
from oracle import adapters
from oracle import density as odens
from oracle import scf as oscf

GOLD = Path(__file__).resolve().parent / "golden" / "scf_si8_small.npz"


def test_lda_potential_is_energy_derivative():
    n = np.logspace(-6, 1, 400)
    f, v = oscf.lda_pz(n)
    h = 1e-6 * n
    fp, _ = oscf.lda_pz(n + h)
    fm, _ = oscf.lda_pz(n - h)
    num = (fp - fm) / (2 * h)
    far = np.abs(np.cbrt(3.0 / (4 * math.pi * n)) - 1.0) > 1e-3  # away from the rs = 1 seam
    assert np.max(np.abs(num - v)[far] / np.abs(v[far])) < 1e-7
    # Exchange-only limit and the low-density correlation branch at rs = 2 (Ha): -0.1423/(1+1.0529 sqrt2+0.6668).
    rs = 2.0
    nn = 3.0 / (4 * math.pi * rs ** 3)
    f2, _ = oscf.lda_pz(np.array([nn]))
    ex = -0.75 * (3.0 / math.pi) ** (1.0 / 3.0) * nn ** (4.0 / 3.0)
    ec = -0.1423 / (1 + 1.0529 * math.sqrt(2.0) + 0.3334 * 2.0)
    assert abs(f2[0] / 2.0 - (ex + nn * ec)) < 1e-15
    # PZ81 is continuous at rs = 1.
    n1 = 3.0 / (4 * math.pi)
    fa, va = oscf.lda_pz(np.array([n1 * (1 + 1e-9), n1 * (1 - 1e-9)]))
    assert abs(fa[0] - fa[1]) < 1e-4 * abs(fa[0]) and abs(va[0] - va[1]) < 1e-3 * abs(va[0])
    assert np.all(oscf.lda_pz(np.array([0.0, 1e-13]))[0] == 0.0)


def test_hartree_single_cosine():
    a = 7.3
    cell = np.diag([a, a * 1.1, a * 0.9])
    recip = 2 * math.pi * np.linalg.inv(cell).T
    grid = (12, 10, 9)
    vol = abs(np.linalg.det(cell))
    nx, ny, nz = grid
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    r = np.stack([x.ravel(), y.ravel(), z.ravel()], 1) @ cell
    gv = 2 * recip[0] - 1 * recip[2]
    amp = 0.013
    rho = 0.5 + amp * np.cos(r @ gv)
    vh, eh = oscf.hartree(rho, grid, recip, vol)
    g2 = float(gv @ gv)
    assert np.max(np.abs(vh - 8 * math.pi * amp / g2 * np.cos(r @ gv))) < 1e-12
    assert abs(eh - 0.5 * vol * 8 * math.pi * amp ** 2 / (2 * g2)) < 1e-12
    g, g2g = dscf.g_vectors(grid, recip)
    vh2, eh2 = dscf.hartree(rho, grid, g2g, vol)
    assert np.max(np.abs(vh2 - vh)) < 1e-12 and abs(eh2 - eh) < 1e-12


@pytest.mark.parametrize("q", [0.3, 1.0, 2.5, 4.0])
def test_gth_local_form_factor_numerical(q):
    z, rloc, c1 = 4.0, 0.44, -7.336103
    vq = adapters.gth_local_q(z, rloc, c1)(q)
    # Short-range part of V(r) (Ha): Z erfc(r/(sqrt2 rloc))/r + C1 exp(-r^2/(2 rloc^2)); -Z/r gives -4 pi Z/q^2.
    def integrand(r):
        vs = z * special.erfc(r / (math.sqrt(2) * rloc)) / r + c1 * math.exp(-r * r / (2 * rloc * rloc))
        return 4 * math.pi * r * r * vs * math.sin(q * r) / (q * r)
    num, _ = integrate.quad(integrand, 1e-12, 12.0, limit=400, epsabs=1e-13)
    ha = num - 4 * math.pi * z / q ** 2
    assert abs(vq - 2.0 * ha) < 1e-9 * max(1.0, abs(vq))


def test_ionic_potential_product_matches_oracle():
    s = dscf.si8_small()
    g, g2 = dscf.g_vectors(s.grid, s.recip)
    v_prod = dscf.ionic_potential(s, g, g2)
    v_orc = oscf.ionic_potential_r(s.grid, s.recip, s.volume, s.atom_pos, adapters.gth_local_q(s.z_ion, s.rloc, s.c1))
    assert np.max(np.abs(v_prod - v_orc)) < 1e-10 * np.max(np.abs(v_orc))


def test_band_weights_match_oracle():
    for nelec, nb, wk in ((32.0, 20, 0.5), (31.0, 20, 0.125), (8.0, 4, 1.0)):
        assert np.array_equal(dscf.band_weights(nelec, nb, wk), odens.band_weights(nelec, nb, wk))


def _linear_map(d, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    a = q @ np.diag(rng.uniform(-0.6, 0.6, d)) @ q.T
    b = rng.standard_normal(d)
    return a, b, np.linalg.solve(np.eye(d) - a, b)


@pytest.mark.parametrize("make", [lambda: oscf.PulayMixer(8), lambda: dscf.Pulay(8)])
def test_pulay_solves_linear_fixed_point(make):
    d = 5
    a, b, fixed = _linear_map(d, 3)
    mixer = make()
    v = np.zeros(d)
    for _ in range(d + 1):
        resid = (a @ v + b) - v
        v = (mixer.mix if hasattr(mixer, "mix") else mixer.step)(v, resid, 0.3)
    assert np.max(np.abs(v - fixed)) < 1e-9


def test_pulay_product_equals_oracle_and_reference_fallback():
    rng = np.random.default_rng(11)
    om, pm = oscf.PulayMixer(4), dscf.Pulay(4)
    v = rng.standard_normal(50)
    for it in range(9):
        r = rng.standard_normal(50) * 10.0 ** (-it / 2)
        vo = om.mix(v, r, 0.3)
        vp = pm.step(v, r, 0.3)
        assert np.max(np.abs(vo - vp)) < 1e-12
        v = vo
    # The reference's absolute pivot threshold: tiny residuals take the equal-weights branch.
    ref = oscf.PulayMixer(3, relative=False)
    v0 = np.ones(8)
    for k in range(3):
        out = ref.mix(v0 * (k + 1), np.full(8, 1e-8 * (k + 1)), 0.3)
    expect = np.mean([v0 * (k + 1) + 0.3 * np.full(8, 1e-8 * (k + 1)) for k in range(3)], axis=0)
    assert np.max(np.abs(out - expect)) < 1e-14


def test_oracle_scf_golden_and_energy_identity():
    s = dscf.si8_small()
    gold = np.load(GOLD)
    res = oscf.run_scf(s, n_iter=40, beta=0.3, conv=1e-10, pulay_history=8, pulay_start=4)
    assert res.converged and len(res.energies) == len(gold["energies"])
    assert np.max(np.abs(np.array(res.energies) - gold["energies"])) < 1e-10
    rho = res.rho
    n = rho.size
    assert abs(np.sum(rho) * s.volume / n - s.nelec) < 1e-9
    # energy.zig:372-396 at self-consistency: band - E_H - int vxc rho + E_xc.
    band = sum(float(np.dot(odens.band_weights(s.nelec, s.nbands, wk), ev))
               for wk, ev in zip(s.kweights, res.eigenvalues))
    _, e_h = oscf.hartree(rho, s.grid, s.recip, s.volume)
    f, vxc = oscf.lda_pz(rho)
    e_ref_form = band - e_h - float(np.dot(vxc, rho)) * s.volume / n + float(np.sum(f)) * s.volume / n
    assert abs(e_ref_form - res.energies[-1]) < 1e-8


def test_fermi_level_and_entropy():
    rng = np.random.default_rng(5)
    ev = [np.sort(rng.normal(0.0, 0.4, 10)) for _ in range(6)]
    w = list(rng.dirichlet(np.ones(6)))
    for sigma in (0.005, 0.01, 0.05):
        mu_o = oscf.find_fermi_level(9.0, sigma, ev, w)
        mu_p = dscf.fermi_level(9.0, sigma, ev, w)
        assert abs(mu_o - mu_p) < 1e-12
        n = sum(2 * wk * float(np.sum(oscf.smearing_occ(v, mu_o, sigma))) for v, wk in zip(ev, w))
        assert abs(n - 9.0) < 1e-8
    # Entropy closed forms: S(1/2) = ln 2, S(0) = S(1) = 0, symmetric.
    assert abs(oscf.smearing_entropy(np.array([0.5]))[0] - math.log(2.0)) < 1e-15
    assert oscf.smearing_entropy(np.array([0.0, 1.0])).tolist() == [0.0, 0.0]
    f = rng.uniform(0, 1, 50)
    assert np.max(np.abs(oscf.smearing_entropy(f) - oscf.smearing_entropy(1 - f))) < 1e-14
    assert np.max(np.abs(dscf.fd_entropy(f) - oscf.smearing_entropy(f))) < 1e-14
    # Fermi-Dirac at e = mu is 1/2; sigma -> 0 is a step.
    assert oscf.smearing_occ(np.array([0.3]), 0.3, 0.01)[0] == 0.5
    assert oscf.smearing_occ(np.array([0.2, 0.4]), 0.3, 0.0).tolist() == [1.0, 0.0]


def test_time_reversal_mesh():
    s = dscf.aluminium(ecut_ry=8.0, kmesh=(4, 4, 4))
    assert len(s.kpoints) == 36 and abs(sum(s.kweights) - 1.0) < 1e-15
    frac = [np.linalg.solve(s.recip.T, k) for k in s.kpoints]
    keys = {tuple(np.round(np.mod(f * 4, 4)).astype(int)) for f in frac}
    for kk in keys:  # no point together with its time-reversed partner unless self-inverse
        neg = tuple(int(v) for v in np.mod(-np.array(kk), 4))
        assert neg == kk or neg not in keys


def test_oracle_smearing_scf_golden():
    gold = np.load(GOLD.parent / "scf_al4_small.npz")
    s = dscf.al4_small()
    res = oscf.run_scf(s, n_iter=60, beta=0.3, conv=1e-12, pulay_history=8, pulay_start=4)
    assert res.converged and len(res.energies) == len(gold["energies"])
    assert np.max(np.abs(np.array(res.energies) - gold["energies"])) < 1e-10
    assert abs(np.sum(res.rho) * s.volume / res.rho.size - s.nelec) < 1e-8


def test_model_dielectric_single_mode():
    """P(G) on a single cosine: scaled by (l^2 G^2 + 1/eps) / (l^2 G^2 + 1); a constant by 1/eps."""
    s = dscf.si8_small()
    nx, ny, nz = s.grid
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    r = np.stack([x.ravel(), y.ravel(), z.ravel()], 1) @ s.cell
    gv = 1 * s.recip[0] + 2 * s.recip[1]
    res = 0.3 + np.cos(r @ gv)
    g2 = float(gv @ gv)
    for eps, l in ((12.0, 1.0), (4.0, 2.5)):
        want = 0.3 / eps + (l * l * g2 + 1 / eps) / (l * l * g2 + 1) * np.cos(r @ gv)
        got = oscf.model_dielectric(res, s.grid, s.recip, eps, l)
        assert np.max(np.abs(got - want)) < 1e-12
        _, g2g = dscf.g_vectors(s.grid, s.recip)
        assert np.max(np.abs(dscf.model_dielectric(res, s.grid, g2g, eps, l) - want)) < 1e-12


def test_atomic_start_density():
    for s in (dscf.si8_small(), dscf.si64()):
        g, g2 = dscf.g_vectors(s.grid, s.recip)
        rho = dscf.atomic_density(s, g, g2)
        assert abs(np.sum(rho) * s.volume / rho.size - s.nelec) < 1e-9 * s.nelec
        assert rho.min() > -1e-6


def test_oracle_pbe_scf_golden():
    gold = np.load(GOLD.parent / "scf_al4_pbe_small.npz")
    s = dscf.al4_small(xc="pbe")
    res = oscf.run_scf(s, n_iter=60, beta=0.3, conv=1e-12, pulay_history=8, pulay_start=4)
    assert res.converged and len(res.energies) == len(gold["energies"])
    assert np.max(np.abs(np.array(res.energies) - gold["energies"])) < 1e-10
    # PBE differs from LDA-PZ on the same cell (the golden files are not the same run).
    lda = np.load(GOLD.parent / "scf_al4_small.npz")
    assert abs(gold["energies"][-1] - lda["energies"][-1]) > 1e-4


def test_pbe_needs_the_device_grid():
    with pytest.raises(ValueError):
        dscf.ScfDriver(dscf.al4_small(xc="pbe"), device_grid=False)


def test_ibz_kmesh_reference_counts():
    """Symmetry reduction of unshifted meshes (reduction.reduce_kmesh,
    src/features/kpoints/reduction.zig:55-120): the reference's own test, fcc Al primitive cell
    6x6x6 -> 16 points with weights summing to 1 (src/features/kpoints/kpoints.zig:94-125); the
    cubic 4-atom Al cell 8x8x8 -> 35 points whose plane-wave counts at 30 Ry span 1,208-1,260;
    2x2x2 -> 4 (Gamma, X, M, R with weights 1, 3, 3, 1 over 8); every mesh point lies in exactly one
    orbit."""
    from dzhip import scf as dscf
    from dzhip import tables as dtab

    prim = np.array([[0, 2.025, 2.025], [2.025, 0, 2.025], [2.025, 2.025, 0]]) * dtab.ANGSTROM_TO_BOHR
    pts, wts, _ = dscf.ibz_kmesh(prim, (6, 6, 6), np.zeros((1, 3)), np.zeros(1, int))
    assert len(pts) == 16 and abs(sum(wts) - 1.0) < 1e-12
    a = 4.05 * dtab.ANGSTROM_TO_BOHR
    cell = np.diag([a, a, a])
    frac = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    assert len(dscf.lattice_symmetry_ops(cell, frac, np.zeros(4, int))) == 48
    pts, wts, fr = dscf.ibz_kmesh(cell, (2, 2, 2), frac, np.zeros(4, int))
    assert sorted(round(w * 8) for w in wts) == [1, 1, 3, 3]
    al = dscf.aluminium(symmetry=True, kmesh=(8, 8, 8))
    assert len(al.kpoints) == 35 and abs(sum(al.kweights) - 1.0) < 1e-12
    npw = [dtab.generate_basis(al.recip, al.ecut_ry, k).kinetic.shape[0] for k in al.kpoints]
    assert min(npw) == 1208 and max(npw) == 1260 and al.grid == (30, 30, 30)
    # Orbits partition the mesh: the weights count 512 distinct points.
    assert sum(round(w * 512) for w in al.kweights) == 512
    # A structure with only the identity (one displaced atom) keeps k and -k together only.
    pts, wts, _ = dscf.ibz_kmesh(cell, (4, 4, 4), frac + np.array([[0, 0, 0], [0.013, 0.007, 0.002], [0, 0, 0],
                                                                      [0, 0, 0]]), np.zeros(4, int))
    assert len(pts) == 36
