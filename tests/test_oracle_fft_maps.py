"""Pins the oracle's FFT conventions, grid sizing, basis and index maps (CPU)."""
import math

import numpy as np
import pytest

from oracle import fftconv, pw
import dzhip.tables as tables


@pytest.mark.parametrize("grid", [(4, 4, 4), (6, 5, 4), (3, 8, 5)])
def test_fft3d_matches_bruteforce_dft(grid):
    rng = np.random.default_rng(1)
    n = int(np.prod(grid))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(fftconv.forward3d(x, grid), fftconv.dft_direct_3d(x, grid, False), atol=1e-11)
    np.testing.assert_allclose(fftconv.inverse3d(x, grid), fftconv.dft_direct_3d(x, grid, True), atol=1e-12)
    np.testing.assert_allclose(fftconv.inverse3d(fftconv.forward3d(x, grid), grid), x, atol=1e-12)


def test_fft_sign_on_single_mode():
    # x_j = e^{+2 pi i j/n} has forward transform n at k = 1 (sign e^{-i} in the forward).
    n = 6
    j = np.arange(n)
    x = np.exp(2j * np.pi * j / n)
    fx = fftconv.dft_direct_1d(x, False)
    expect = np.zeros(n, complex)
    expect[1] = n
    np.testing.assert_allclose(fx, expect, atol=1e-12)


def test_apply_local_delta_psi_closed_form():
    """src/tests.zig:872-973: psi = delta(G=0); FFT path reproduces V_G to 1e-10."""
    grid = (4, 4, 4)
    nx, ny, nz = grid
    total = nx * ny * nz
    v_g = np.zeros(total, complex)
    v_g[0], v_g[1], v_g[nx - 1] = 0.5, 0.1, 0.1
    local_r = fftconv.inverse3d(v_g * total, grid).real
    assert abs(local_r[0] - 0.7) < 1e-12
    psi_g = np.zeros(total, complex)
    psi_g[0] = 1.0
    psi_r = fftconv.inverse3d(psi_g * total, grid)
    out = fftconv.forward3d(psi_r * local_r, grid) / total
    assert np.max(np.abs(out - v_g)) < 1e-10


def test_local_fft_vs_direct_convolution():
    """src/tests.zig:975-1106: N * FFT(IFFT(V) IFFT(psi)) equals the direct convolution."""
    grid = (4, 4, 4)
    nx, ny, nz = grid
    total = nx * ny * nz
    rng = np.random.default_rng(3)
    v_g = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    psi_g = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    direct = np.zeros(total, complex)
    for gz in range(nz):
        for gy in range(ny):
            for gx in range(nx):
                s = 0j
                for pz in range(nz):
                    for py in range(ny):
                        for px in range(nx):
                            dq = (gx - px) % nx + nx * (((gy - py) % ny) + ny * ((gz - pz) % nz))
                            s += v_g[dq] * psi_g[px + nx * (py + ny * pz)]
                direct[gx + nx * (gy + ny * gz)] = s
    conv = fftconv.forward3d(fftconv.inverse3d(v_g, grid) * fftconv.inverse3d(psi_g, grid), grid) * total
    assert np.max(np.abs(conv - direct)) < 1e-10


def test_next_fft_size_and_auto_grid():
    assert [pw.next_fft_size(n) for n in (1, 7, 11, 13, 31, 33, 71, 73, 77, 91)] == [1, 8, 12, 15, 32, 36, 72, 75, 80, 96]
    a_si = 5.431 * pw.ANGSTROM_TO_BOHR
    # Appendix sizes derived from auto_grid (pw_grid.zig:54-68): Si-8 at 20 Ry -> 32^3 (not 24^3),
    # Si-64 (2a) at 30 Ry -> 75^3, Al-4 (a = 4.05 A) at 30 Ry -> 30^3, GaAs-64 at 35 Ry -> 90^3.
    assert pw.auto_grid(20.0, pw.reciprocal(np.eye(3) * a_si)) == (32, 32, 32)
    assert pw.auto_grid(30.0, pw.reciprocal(np.eye(3) * 2 * a_si)) == (75, 75, 75)
    assert pw.auto_grid(30.0, pw.reciprocal(np.eye(3) * 4.05 * pw.ANGSTROM_TO_BOHR)) == (30, 30, 30)
    assert pw.auto_grid(35.0, pw.reciprocal(np.eye(3) * 11.306 * pw.ANGSTROM_TO_BOHR)) == (90, 90, 90)
    for e, g in ((20.0, np.eye(3) * a_si), (30.0, np.eye(3) * 3.3)):
        assert tables.auto_grid(e, tables.reciprocal(g)) == pw.auto_grid(e, pw.reciprocal(g))


def test_reciprocal_duality():
    cell = np.array([[7.0, 0.0, 0.0], [0.6, 6.5, 0.0], [0.3, -0.4, 7.4]])
    b = pw.reciprocal(cell)
    np.testing.assert_allclose(cell @ b.T, 2 * math.pi * np.eye(3), atol=1e-12)


def test_basis_counts_and_order():
    # Bench box: 23,847 plane waves at 15 Ha in a 20.5 bohr cube.
    b = pw.generate_basis(pw.reciprocal(np.eye(3) * 20.5), 30.0)
    assert b.n == 23847
    # Si-8 at Gamma, 20 Ry: 1,647 plane waves.
    a_si = 5.431 * pw.ANGSTROM_TO_BOHR
    assert pw.generate_basis(pw.reciprocal(np.eye(3) * a_si), 20.0).n == 1647
    # Loop order h, k, l ascending (basis.zig:43-70): lexicographic.
    key = b.h.astype(np.int64) * 10 ** 6 + b.k.astype(np.int64) * 10 ** 3 + b.l
    assert np.all(np.diff(key) > 0)
    g0 = np.where((b.h == 0) & (b.k == 0) & (b.l == 0))[0]
    assert g0.size == 1 and b.kinetic[g0[0]] == 0.0


def test_basis_vectorised_matches_oracle():
    cell = np.array([[7.0, 0.0, 0.0], [0.6, 6.5, 0.0], [0.3, -0.4, 7.4]])
    rec = pw.reciprocal(cell)
    k = np.array([0.1, 0.2, -0.15]) @ rec
    bo = pw.generate_basis(rec, 6.0, k)
    bt = tables.generate_basis(tables.reciprocal(cell), 6.0, k)
    assert bo.n == bt.n
    for f in ("h", "k", "l", "cart", "kpg", "kinetic"):
        np.testing.assert_array_equal(getattr(bo, f), getattr(bt, f))


@pytest.mark.parametrize("grid", [(8, 8, 8), (7, 6, 9), (12, 10, 8)])
def test_fft_index_maps(grid):
    cell = np.array([[7.0, 0.0, 0.0], [0.6, 6.5, 0.0], [0.3, -0.4, 7.4]])
    rec = pw.reciprocal(cell)
    b = pw.generate_basis(rec, 3.0, np.array([0.1, 0.2, -0.15]) @ rec)
    fo = pw.fft_indices(b, grid)
    ft = tables.fft_indices(tables.Basis(b.h, b.k, b.l, b.cart, b.kpg, b.kinetic), grid)
    np.testing.assert_array_equal(fo, ft)
    assert len(set(fo.tolist())) == b.n  # injective
    # Independent restatement: FFT position holds frequency index_to_freq(x) == h.
    nx, ny, nz = grid
    x, y, z = fo % nx, (fo // nx) % ny, fo // (nx * ny)
    assert all(pw.index_to_freq(int(xi), nx) == hi for xi, hi in zip(x, b.h))
    assert all(pw.index_to_freq(int(yi), ny) == ki for yi, ki in zip(y, b.k))
    assert all(pw.index_to_freq(int(zi), nz) == li for zi, li in zip(z, b.l))
    rng = np.random.default_rng(0)
    c = rng.standard_normal(b.n) + 1j * rng.standard_normal(b.n)
    total = nx * ny * nz
    np.testing.assert_array_equal(pw.gather_fft(fo, pw.scatter_fft(fo, c, total, 1.0), 1.0), c)


def test_plane_wave_outside_grid_is_rejected():
    rec = pw.reciprocal(np.eye(3) * 10.0)
    b = pw.generate_basis(rec, 10.0)
    with pytest.raises(ValueError):
        pw.fft_indices(b, (4, 4, 4))
    with pytest.raises(ValueError):
        tables.fft_indices(tables.Basis(b.h, b.k, b.l, b.cart, b.kpg, b.kinetic), (4, 4, 4))
