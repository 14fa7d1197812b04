"""GPU parity of the device eigensolver (dzhip_eig_solve) against exact spectra.

Reference semantics: src/features/linalg/lobpcg/serial.zig:55-307 (converged when the largest
residual 2-norm is below tol, in Ry).  Exact answers: numpy.linalg.eigh of the dense FP64
Hamiltonian built by the oracle, and the reference's diagonal-operator tests
(src/features/linalg/iterative.zig:80-203) realised as a Hamiltonian with V = 0 and no
projectors, whose eigenvalues are the sorted |k+G|^2.
Tolerance on eigenvalues: 1e-8 Ha = 2e-8 Ry, with solver tol 1e-9 Ry.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from dzhip import workloads
import dzhip.tables as tables
from oracle import hamiltonian

sys.path.insert(0, str(Path(__file__).resolve().parent))
from test_oracle_physics import oracle_ctx  # noqa: E402

pytestmark = pytest.mark.gpu
EV_TOL = 2e-8  # Ry


@pytest.fixture(scope="module")
def si8_exact():
    w = workloads.si8(nbands=1)
    h = hamiltonian.dense_hamiltonian(oracle_ctx(w))
    return w, np.linalg.eigvalsh(h)


def test_eig_tiny_matches_dense(dz, ctx):
    w = workloads.tiny()
    exact = np.linalg.eigvalsh(hamiltonian.dense_hamiltonian(oracle_ctx(w)))
    ham = dz.Hamiltonian.from_workload(ctx, w)
    vals, vecs, st = ham.eig_solve(6, tol=1e-10, max_iter=200)
    print("tiny", st, vals - exact[:6])
    assert st["converged"] == 1, st
    np.testing.assert_allclose(vals, exact[:6], atol=EV_TOL)
    # Eigenvectors: orthonormal and H x = lambda x.
    np.testing.assert_allclose(vecs.conj() @ vecs.T, np.eye(6), atol=1e-10)
    hx = ham.apply(vecs)
    assert np.max(np.linalg.norm(hx - vals[:, None] * vecs, axis=1)) < 1e-9


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("precond", [0, 1])
def test_eig_si8_matches_dense(dz, ctx, si8_exact, prec, precond):
    w, exact = si8_exact
    ham = dz.Hamiltonian.from_workload(ctx, w)
    vals, vecs, st = ham.eig_solve(20, tol=1e-9, max_iter=300, precision=prec, preconditioner=precond)
    print("si8 prec", prec, "precond", precond, st, np.max(np.abs(vals - exact[:20])))
    assert st["converged"] == 1, st
    np.testing.assert_allclose(vals, exact[:20], atol=EV_TOL)
    if prec == 1:
        assert st["fp32_applies"] > 0 and st["precision_switch_iter"] > 0
    print("si8 prec", prec, "precond", precond, st)


def test_eig_diagonal_operator(dz, ctx):
    """V = 0, no projectors: H = diag(|k+G|^2); lowest values are the sorted kinetic energies."""
    w = workloads.si8(nbands=1)
    ham = dz.Hamiltonian(ctx, w.grid, tables.fft_indices(w.basis, w.grid), w.basis.kinetic, w.basis.cart,
                         np.zeros(int(np.prod(w.grid))), w.inv_volume)
    vals, _, st = ham.eig_solve(10, tol=1e-9, max_iter=100, init_diagonal=True)
    np.testing.assert_allclose(vals, np.sort(w.basis.kinetic)[:10], atol=1e-10)
    assert st["converged"] == 1


def test_eig_warm_start_and_errors(dz, ctx, si8_exact):
    w, exact = si8_exact
    ham = dz.Hamiltonian.from_workload(ctx, w)
    vals, vecs, st = ham.eig_solve(12, tol=1e-9, max_iter=300)
    vals2, _, st2 = ham.eig_solve(12, tol=1e-9, max_iter=300, init=vecs)
    np.testing.assert_allclose(vals2, exact[:12], atol=EV_TOL)
    assert st2["apply_calls"] < st["apply_calls"] / 2, (st, st2)
    with pytest.raises(dz.InvalidMatrixSize):
        ham.eig_solve(4, diag=w.basis.kinetic[:-1])
    with pytest.raises(dz.BadArgument):
        ham.eig_solve(w.npw + 1)


@pytest.mark.parametrize("name", ["tiny", "si8"])
def test_density_matches_oracle(dz, ctx, name):
    """rho(r) = (1/Omega) sum_b w_b |psi_b(r)|^2 vs the oracle (kpoint/density.zig:444-474), FP64 1e-12."""
    from oracle import density as odens

    w = workloads.BUILDERS[name]()
    o = oracle_ctx(w)
    nb = w.nbands
    wts = odens.band_weights(2.0 * nb - 1.0, nb, 0.5)  # last band half occupied
    ref = odens.accumulate_density(o.fft_idx, w.grid, w.psi, wts, w.inv_volume)
    ham = dz.Hamiltonian.from_workload(ctx, w)
    rho = ham.density(w.psi, wts)
    assert np.max(np.abs(rho - ref)) < 1e-12 * np.max(np.abs(ref))
    rho2 = ham.density(w.psi[:2], wts[:2], rho=rho)  # accumulate mode
    ref2 = odens.accumulate_density(o.fft_idx, w.grid, w.psi[:2], wts[:2], w.inv_volume, ref.copy())
    assert np.max(np.abs(rho2 - ref2)) < 1e-12 * np.max(np.abs(ref2))
    # zero weights skip bands
    rz = ham.density(w.psi, np.zeros(nb))
    assert np.all(rz == 0.0)


def test_eig_nconv_converges_lowest_bands(dz, ctx, si8_exact):
    """nconv = 8 of 12 bands: converged once the lowest 8 residuals are below tol; those match
    the dense spectrum, and the solve needs no more applies than converging all 12."""
    w, exact = si8_exact
    ham = dz.Hamiltonian.from_workload(ctx, w)
    v_all, _, st_all = ham.eig_solve(12, tol=1e-9, max_iter=300)
    v8, _, st8 = ham.eig_solve(12, tol=1e-9, max_iter=300, nconv=8)
    print(st_all["apply_calls"], st8["apply_calls"])
    assert st8["converged"] == 1
    np.testing.assert_allclose(v8[:8], exact[:8], atol=EV_TOL)
    assert st8["apply_calls"] <= st_all["apply_calls"]


def test_density_solution_guard(dz, ctx):
    """Regression: the density of the last solution reads that solution's columns on the device;
    asking for it before any solve, or for more bands than the solve returned, is DZHIP_ERR_STATE
    (BadState), not an access past the solution buffer."""
    w = workloads.si8(nbands=4)
    ham = dz.Hamiltonian.from_workload(ctx, w)
    try:
        with pytest.raises(dz.BadState):
            ham.density_solution(np.ones(2))
        ham.eig_solve(4, tol=1e-6, max_iter=60, init_diagonal=True, want_vectors=False)
        with pytest.raises(dz.BadState):
            ham.density_solution(np.ones(6))
        rho = ham.density_solution(np.full(4, 0.5))
        assert rho.shape == (int(np.prod(w.grid)),) and np.all(np.isfinite(rho))
        # Integrated density = sum of weights (each solution vector has unit norm).
        assert abs(rho.sum() * (1.0 / w.inv_volume) / rho.size - 2.0) < 1e-10
    finally:
        ham.close()
