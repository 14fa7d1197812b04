"""Pins the oracle's projector tables, H.psi, eigensolver and density against closed forms,
brute force and the reference's own exact-answer tests (CPU)."""
import math

import numpy as np
import pytest

from oracle import adapters, density, eigensolver, hamiltonian, projectors, pw
import dzhip.tables as tables
from dzhip import workloads


# ---------------------------------------------------------------- projectors
def test_spherical_bessel_closed_forms():
    """src/tests.zig:1247+: j0, j1, j2 vs closed forms."""
    for x in (0.1, 1.0, 2.0, 5.0, 10.0, 20.0):
        assert abs(projectors.spherical_bessel(0, x) - math.sin(x) / x) < 1e-12
        assert abs(projectors.spherical_bessel(1, x) - (math.sin(x) / x ** 2 - math.cos(x) / x)) < 1e-12
        j2 = (3 / x ** 2 - 1) * math.sin(x) / x - 3 * math.cos(x) / x ** 2
        assert abs(projectors.spherical_bessel(2, x) - j2) < 1e-9
    assert projectors.spherical_bessel(0, 0.0) == 1.0
    assert projectors.spherical_bessel(2, 1e-8) == pytest.approx(1e-16 / 15, rel=1e-6)


def test_angular_factor_values():
    """src/tests.zig:1218-1245."""
    assert abs(projectors.angular_factor(0, 1.0) - 4 * math.pi) < 1e-10
    assert abs(projectors.angular_factor(1, 1.0) - 12 * math.pi) < 1e-10
    assert abs(projectors.legendre_p(5, 0.3) - (63 * 0.3 ** 5 - 70 * 0.3 ** 3 + 15 * 0.3) / 8) < 1e-14


@pytest.mark.parametrize("l", [0, 1, 2, 3, 4])
def test_ylm_addition_theorem(l):
    """sum_m Y_lm(a) Y_lm(b) = (2l+1)/(4 pi) P_l(cos gamma) (nonlocal.zig:151); catches a wrong
    coefficient or a swapped m in any of the tesseral harmonics."""
    rng = np.random.default_rng(l)
    for _ in range(20):
        a, b = rng.standard_normal(3), rng.standard_normal(3)
        s = sum(projectors.real_spherical_harmonic(l, m, *a) * projectors.real_spherical_harmonic(l, m, *b)
                for m in range(-l, l + 1))
        c = float(a @ b / np.linalg.norm(a) / np.linalg.norm(b))
        assert abs(s - (2 * l + 1) / (4 * math.pi) * projectors.legendre_p(l, c)) < 1e-12
    # Vectorised host tables agree with the scalar restatement.
    v = rng.standard_normal((7, 3))
    yv = tables.real_ylm(l, v)
    for mi in range(2 * l + 1):
        for i in range(7):
            assert abs(yv[mi, i] - projectors.real_spherical_harmonic(l, mi - l, *v[i])) < 1e-13


@pytest.mark.parametrize("l", [0, 1, 2])
def test_radial_projector_gaussian_closed_form(l):
    """Quadrature of radial_projector (ctrap weights) vs the Gaussian-Hankel closed form."""
    rc = 1.0
    n = 2001
    r = np.linspace(0.0, 12.0, n)
    rab = np.full(n, r[1] - r[0])
    norm = projectors.gaussian_beta_norm(l, rc)
    beta_rb = r * norm * r ** l * np.exp(-r ** 2 / (2 * rc ** 2))  # UPF stores r*beta
    exact = projectors.gaussian_beta_radial(l, rc, norm)
    for q in (0.0, 0.7, 1.9, 3.5):
        assert abs(projectors.radial_projector(beta_rb, r, rab, l, q) - exact(q)) < 1e-8
    # Host vectorised closed form == oracle closed form.
    qs = np.array([0.0, 0.5, 2.0])
    np.testing.assert_allclose(tables.gaussian_radial(l, rc)(qs), [exact(q) for q in qs], rtol=1e-14)


def test_radial_table_interpolation():
    t = projectors.RadialTable.from_function(lambda q: 2.0 * q + 1.0, 3.0, 7)
    assert t.eval(-1.0) == 1.0 and t.eval(3.0) == 0.0 and t.eval(5.0) == 0.0
    assert abs(t.eval(1.25) - 3.5) < 1e-14


# ---------------------------------------------------------------- H.psi
def oracle_ctx(w):
    return adapters.context_from_workload(w)


@pytest.fixture(scope="module")
def tiny():
    w = workloads.tiny()
    return w, oracle_ctx(w)


def test_phi_tables_match_oracle(tiny):
    w, ctx = tiny
    for s in range(len(w.species)):
        np.testing.assert_allclose(w.phi(s), ctx.phi(s), rtol=1e-12, atol=1e-14)


def test_apply_matches_dense_bruteforce(tiny):
    w, ctx = tiny
    h = hamiltonian.dense_hamiltonian(ctx)
    assert np.max(np.abs(h - h.conj().T)) < 1e-12  # Hermitian
    for b in range(w.nbands):
        y = hamiltonian.apply_hamiltonian(ctx, w.psi[b])
        ref = h @ w.psi[b]
        assert np.max(np.abs(y - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_nonlocal_apply_vs_legendre_dense(tiny):
    w, ctx = tiny
    hn = hamiltonian.dense_nonlocal(ctx)
    x = w.psi[0]
    y = hamiltonian.apply_nonlocal(ctx, x)
    assert np.max(np.abs(y - hn @ x)) < 1e-11 * np.max(np.abs(hn @ x))
    assert np.max(np.abs(y)) > 1e-3  # the nonlocal term is not trivially zero


def test_batched_equals_single(tiny):
    w, ctx = tiny
    yb = hamiltonian.apply_hamiltonian_batched(ctx, w.psi)
    for b in range(w.nbands):
        np.testing.assert_allclose(yb[b], hamiltonian.apply_hamiltonian(ctx, w.psi[b]), atol=1e-12)
    with pytest.raises(ValueError):
        hamiltonian.apply_hamiltonian(ctx, w.psi[0][:-1])


def test_planted_mistakes_are_caught(tiny):
    """A flipped phase sign or dropped 1/Omega in the oracle would disagree with the dense matrix."""
    w, ctx = tiny
    h = hamiltonian.dense_nonlocal(ctx)
    x = w.psi[1]
    good = hamiltonian.apply_nonlocal(ctx, x)
    saved = ctx.inv_volume
    ctx.inv_volume = 1.0
    bad = hamiltonian.apply_nonlocal(ctx, x)
    ctx.inv_volume = saved
    assert np.max(np.abs(bad - h @ x)) > 1e-3
    pos = [a.position.copy() for a in ctx.atoms]
    for a in ctx.atoms:
        a.position = -a.position
    bad2 = hamiltonian.apply_nonlocal(ctx, x)
    for a, p in zip(ctx.atoms, pos):
        a.position = p
    assert np.max(np.abs(bad2 - good)) > 1e-3


# ---------------------------------------------------------------- eigensolver
def _diag_op(d):
    d = np.asarray(d, float)
    return lambda xb: xb * d[None, :]


def test_davidson_diagonal_operators():
    """linalg/iterative.zig:80-203: eigenvalues {1,2} of diag(1,2,3) and {1,3} of diag(1,3,5,7)."""
    r = eigensolver.davidson(_diag_op([1, 2, 3]), 3, np.array([1.0, 2.0, 3.0]), 2,
                             eigensolver.Options(max_iter=20, tol=1e-8))
    np.testing.assert_allclose(r.values, [1.0, 2.0], atol=1e-6)
    r = eigensolver.davidson(_diag_op([1, 3, 5, 7]), 4, np.array([1.0, 3, 5, 7]), 2,
                             eigensolver.Options(max_iter=20, tol=1e-8))
    np.testing.assert_allclose(r.values, [1.0, 3.0], atol=1e-6)


def test_davidson_tridiagonal_known_spectrum():
    """linalg/cg/cg.zig:543-593: 5x5 tridiag(2 on diagonal, -1 off) lowest eigenvalues
    0.2679, 1.0, 2.0 (closed form 2 - 2 cos(k pi / 6))."""
    n = 5
    a = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    r = eigensolver.davidson(lambda xb: xb @ a.T, n, np.diag(a).copy(), 3,
                             eigensolver.Options(max_iter=50, tol=1e-10))
    np.testing.assert_allclose(r.values, [2 - 2 * math.cos(k * math.pi / 6) for k in (1, 2, 3)], atol=1e-9)


def test_davidson_matches_eigh_on_tiny(tiny):
    w, ctx = tiny
    h = hamiltonian.dense_hamiltonian(ctx)
    exact = np.linalg.eigvalsh(h)[:6]
    r = eigensolver.davidson(lambda xb: hamiltonian.apply_hamiltonian_batched(ctx, xb), ctx.npw,
                             ctx.basis.kinetic.copy(), 6, eigensolver.Options(max_iter=200, tol=1e-9))
    assert r.converged
    np.testing.assert_allclose(r.values, exact, atol=1e-9)


def test_lcg_sequence():
    g = eigensolver.Lcg()
    s = 0x6F3D3CBE
    for _ in range(5):
        s = (s * 6364136223846793005 + 1) % 2 ** 64
        assert g.next01() == (s >> 11) / 2.0 ** 53


# ---------------------------------------------------------------- density
def test_density_bruteforce_and_charge(tiny):
    w, ctx = tiny
    nx, ny, nz = w.grid
    q, _ = np.linalg.qr(w.psi.T)  # orthonormal columns
    psi = q.T
    wts = density.band_weights(5.0, psi.shape[0], 1.0)  # nelec 5 -> occ 1,1,0.5
    np.testing.assert_allclose(wts, [2.0, 2.0, 1.0, 0.0])
    rho = density.accumulate_density(ctx.fft_idx, w.grid, psi, wts, w.inv_volume)
    # Integral of rho = sum of weights for normalised bands.
    assert abs(rho.sum() * w.volume / rho.size - wts.sum()) < 1e-10
    # Brute force |sum_G c_G e^{iG.r}|^2 at a few grid points (fractional r = (x/nx, y/ny, z/nz)).
    for idx in (0, 17, 300):
        x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
        val = 0.0
        for b in range(3):
            ph = np.exp(2j * np.pi * (w.basis.h * x / nx + w.basis.k * y / ny + w.basis.l * z / nz))
            val += wts[b] * abs(np.sum(psi[b] * ph)) ** 2 * w.inv_volume
        assert abs(rho[idx] - val) < 1e-12 * max(1.0, val)
