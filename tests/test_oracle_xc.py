"""The oracle's PBE functional and GGA fields (oracle/scf.py: pw92, pbe, gradient, divergence,
xc_fields), pinned without the reference: PW92 against its published correlation energies and
against PZ81 (an independent fit of the same Ceperley-Alder data), the analytic derivatives
against central differences, the large-gradient limits fixed by the PBE construction, and the
GGA potential against the functional derivative of E_xc on the grid."""
import math

import numpy as np

from dzhip import scf as dscf
from oracle import scf as oscf

C_X = -0.75 * (3.0 / math.pi) ** (1.0 / 3.0)


def test_pw92_published_values_and_pz81():
    # Perdew-Wang 1992 unpolarised correlation energy per electron (Ha) at rs = 1, 2, 5, 10.
    ref = {1.0: -0.0598, 2.0: -0.0448, 5.0: -0.0282, 10.0: -0.0186}
    for rs, e in ref.items():
        assert abs(float(oscf.pw92(np.array([rs]))[0][0]) - e) < 1e-4
    rs = np.linspace(0.3, 10.0, 200)
    n = 3.0 / (4.0 * math.pi * rs ** 3)
    f, _ = oscf.lda_pz(n)
    ec_pz = f / (2.0 * n) - C_X * np.cbrt(n)
    assert np.max(np.abs(oscf.pw92(rs)[0] - ec_pz)) < 1e-3
    # v_c = d(n eps_c)/dn by central differences in n.
    h = 1e-6 * n
    e_p = oscf.pw92(np.cbrt(3.0 / (4.0 * math.pi * (n + h))))[0] * (n + h)
    e_m = oscf.pw92(np.cbrt(3.0 / (4.0 * math.pi * (n - h))))[0] * (n - h)
    np.testing.assert_allclose(oscf.pw92(rs)[2], (e_p - e_m) / (2.0 * h), rtol=1e-7)


def test_pbe_without_gradient_is_slater_exchange_plus_pw92():
    n = np.geomspace(1e-4, 5.0, 50)
    f, fn, fg = oscf.pbe(n, np.zeros_like(n))
    eps, _, vc = oscf.pw92(np.cbrt(3.0 / (4.0 * math.pi * n)))
    np.testing.assert_allclose(f, 2.0 * n * (C_X * np.cbrt(n) + eps), rtol=1e-13)
    np.testing.assert_allclose(fn, 2.0 * (4.0 / 3.0 * C_X * np.cbrt(n) + vc), rtol=1e-13)
    # At zero gradient the exchange and correlation gradient terms cancel exactly, because PBE
    # fixes mu = beta pi^2 / 3: -3 mu / (16 pi k_F n) + beta pi / (16 k_F n) = 0.
    assert abs(0.2195149727645171 - 0.06672455060314922 * math.pi ** 2 / 3.0) < 1e-15
    kf = np.cbrt(3.0 * math.pi ** 2 * n)
    scale = np.abs(2.0 * n * C_X * np.cbrt(n) * 0.2195149727645171 / (4.0 * kf * kf * n * n))
    assert np.all(np.abs(fg) < 1e-12 * scale)
    f0, fn0, fg0 = oscf.pbe(np.array([0.0, 1e-13]), np.array([1.0, 1.0]))
    assert np.all(f0 == 0.0) and np.all(fn0 == 0.0) and np.all(fg0 == 0.0)


def test_pbe_derivatives_central_differences():
    rng = np.random.Generator(np.random.PCG64(3))
    n = np.exp(rng.uniform(np.log(1e-3), np.log(2.0), 400))
    # |grad n|^2 spanning s^2 from 1e-3 to 10: s^2 = g2 / (4 kf^2 n^2).
    kf = np.cbrt(3.0 * math.pi ** 2 * n)
    g2 = np.exp(rng.uniform(np.log(1e-3), np.log(10.0), 400)) * 4.0 * kf * kf * n * n
    f, fn, fg = oscf.pbe(n, g2)
    hn, hg = 1e-6 * n, 1e-6 * g2
    num_n = (oscf.pbe(n + hn, g2)[0] - oscf.pbe(n - hn, g2)[0]) / (2.0 * hn)
    num_g = (oscf.pbe(n, g2 + hg)[0] - oscf.pbe(n, g2 - hg)[0]) / (2.0 * hg)
    np.testing.assert_allclose(fn, num_n, rtol=1e-6, atol=1e-9 * np.max(np.abs(fn)))
    # df/dg2 is a difference of the exchange and correlation terms, which cancel as s -> 0, so
    # its error is measured against the size of the exchange term.
    scale = np.abs(2.0 * n * C_X * np.cbrt(n) * 0.2195149727645171 / (4.0 * kf * kf * n * n))
    assert np.all(np.abs(fg - num_g) < 1e-6 * scale)


def test_pbe_large_gradient_limits():
    """s, t -> infinity: F_x -> 1 + kappa and H -> -eps_c (correlation vanishes), so
    f -> 2 n eps_x^LDA (1 + kappa)."""
    n = np.array([0.01, 0.1, 1.0])
    ex0 = n * C_X * np.cbrt(n)
    kf = np.cbrt(3.0 * math.pi ** 2 * n)
    f, _, _ = oscf.pbe(n, 1e14 * 4.0 * kf * kf * n * n)
    np.testing.assert_allclose(f / (2.0 * ex0), 1.804, rtol=1e-5)


def test_gradient_of_single_cosine():
    s = dscf.al4_small()
    nx, ny, nz = s.grid
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    r = np.stack([x.ravel(), y.ravel(), z.ravel()], 1) @ s.cell
    gv = 2 * s.recip[0] - 1 * s.recip[1] + 3 * s.recip[2]
    g = oscf.gradient(0.4 + np.cos(r @ gv), s.grid, s.recip)
    for c in range(3):
        np.testing.assert_allclose(g[c], -gv[c] * np.sin(r @ gv), atol=1e-12)
    div = oscf.divergence(*[np.cos(r @ gv) * gv[c] for c in range(3)], s.grid, s.recip)
    np.testing.assert_allclose(div, -(gv @ gv) * np.sin(r @ gv), atol=1e-11)


def test_gga_potential_is_the_functional_derivative():
    """E_xc[rho] = sum f dV on an odd grid (no Nyquist plane, so the spectral gradient's
    adjoint is minus the spectral divergence): dE/d eps along a smooth direction equals
    sum v_xc d rho dV, which fixes the sign and the factor 2 of the divergence term."""
    s = dscf.al4_small()
    nx, ny, nz = s.grid
    assert nx % 2 and ny % 2 and nz % 2
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    r = np.stack([x.ravel(), y.ravel(), z.ravel()], 1) @ s.cell
    g1, g2 = s.recip[0] + s.recip[1], s.recip[2] - 2 * s.recip[0]
    rho = 0.03 + 0.012 * np.cos(r @ g1) + 0.008 * np.sin(r @ g2 + 0.3)
    drho = np.cos(r @ (g1 + g2) + 0.7) + 0.5 * np.sin(r @ s.recip[1])
    dv = s.volume / rho.size
    f, v = oscf.xc_fields(rho, s.grid, s.recip, "pbe")
    eps = 1e-5

    def energy(e):
        return float(np.sum(oscf.xc_fields(rho + e * drho, s.grid, s.recip, "pbe")[0])) * dv

    def central(h):
        return (energy(h) - energy(-h)) / (2.0 * h)

    num = (4.0 * central(eps) - central(2.0 * eps)) / 3.0  # Richardson: O(eps^4)
    ana = float(np.sum(v * drho)) * dv
    assert abs(num - ana) < 1e-9 * abs(ana)
    # The gradient term matters here: dropping it moves the result well outside that bound.
    f_lda, vn, _ = oscf.pbe(rho, np.zeros_like(rho))
    assert abs(float(np.sum(vn * drho)) * dv - ana) > 1e-4 * abs(ana)
