"""GPU band-by-band CG (dzhip_cg_solve) against the CG oracle and exact spectra.

Cases follow the reference's own CG tests (src/features/linalg/cg/cg.zig:500-593): a diagonal
operator (V = 0, no projectors: H = diag(|k+G|^2)) and non-diagonal Hamiltonians with identity
preconditioning (diag = 0), plus the same start vectors and iteration budget as the oracle with
the reference's kinetic-shift preconditioner, where both follow the same trajectory.
"""
import numpy as np
import pytest

from dzhip import workloads
import dzhip.tables as tables
from oracle import cg as ocg
from oracle import hamiltonian
from test_oracle_physics import oracle_ctx

pytestmark = pytest.mark.gpu


def test_cg_diagonal_operator(dz, ctx):
    w = workloads.si8(nbands=1)
    ham = dz.Hamiltonian(ctx, w.grid, tables.fft_indices(w.basis, w.grid), w.basis.kinetic, w.basis.cart,
                         np.zeros(int(np.prod(w.grid))), w.inv_volume)
    try:
        vals, _, st = ham.cg_solve(6, diag=np.zeros(w.npw), tol=1e-10, max_iter=200)
        print(st)
        np.testing.assert_allclose(vals, np.sort(w.basis.kinetic)[:6], atol=1e-8)
    finally:
        ham.close()


@pytest.mark.parametrize("name,nb", [("si8", 8)])
def test_cg_identity_preconditioner_vs_dense(dz, ctx, name, nb):
    w = workloads.BUILDERS[name]()
    exact = np.linalg.eigvalsh(hamiltonian.dense_hamiltonian(oracle_ctx(w)))
    ham = dz.Hamiltonian.from_workload(ctx, w)
    try:
        vals, vecs, st = ham.cg_solve(nb, diag=np.zeros(w.npw), tol=1e-12, max_iter=2000)
        print(name, st["iterations"], np.max(np.abs(vals - exact[:nb])))
        assert np.max(np.abs(vals - exact[:nb])) < 2e-8
        assert np.max(np.abs(vecs.conj() @ vecs.T - np.eye(nb))) < 1e-10
    finally:
        ham.close()


@pytest.mark.parametrize("max_iter", [1, 3, 6])
def test_cg_trajectory_matches_oracle(dz, ctx, max_iter):
    """Same LCG start, kinetic-shift preconditioner, fixed budget: device and oracle agree."""
    w = workloads.tiny()
    octx = oracle_ctx(w)
    h = hamiltonian.dense_hamiltonian(octx)
    ref_vals, ref_vecs = ocg.solve(lambda x: h @ x, w.basis.kinetic, 4, max_iter=max_iter, tol=1e-14)
    ham = dz.Hamiltonian.from_workload(ctx, w)
    try:
        vals, vecs, st = ham.cg_solve(4, tol=1e-14, max_iter=max_iter)
        print(max_iter, vals - ref_vals)
        assert np.max(np.abs(vals - ref_vals)) < 1e-9
    finally:
        ham.close()
