"""GPU parity of the SCF driver (device eigensolver + device density, host potentials and
Pulay mixing) against the FP64 oracle SCF (oracle/scf.py, dense eigh per k).

Case: the 8-atom Si cell at 8 Ry, k = Gamma and (1/4,1/4,1/4), 20 bands, GTH-form synthetic
pseudopotential (dzhip.scf.si8_small); oracle results in tests/golden/scf_si8_small.npz.
Bars: eigenvalues within 1e-8 Ha (2e-8 Ry), total energy within 1e-6 Ha (2e-6 Ry), density
L1 error int |rho - rho_ref| dV within 1e-7 electrons.
"""
from pathlib import Path

import numpy as np
import pytest

from dzhip import FP32, FP64, PRECOND_TPA
from dzhip import scf as dscf

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden" / "scf_si8_small.npz"
EV_TOL = 2e-8  # Ry
E_TOL = 2e-6  # Ry
RHO_L1 = 1e-7


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.fixture(scope="module")
def system():
    return dscf.si8_small()


def _l1(system, rho, rho_ref):
    return float(np.sum(np.abs(rho - rho_ref))) * system.volume / rho.shape[0]


def test_scf_trajectory_matches_oracle(dz, ctx, gold, system):
    """Tight eigen-solves every iteration: the same SCF trajectory as the dense oracle."""
    drv = dscf.ScfDriver(system)
    try:
        run = drv.run(max_iter=40, conv=1e-10, eig_tol=1e-10, eig_max_iter=400, warmup_steps=0,
                      final_tol=1e-10)
    finally:
        drv.close()
    e = np.array(run.energies)
    print("iterations", len(e), "oracle", len(gold["energies"]))
    print("dE per iteration", e[:len(gold["energies"])] - gold["energies"][:len(e)])
    assert run.converged
    assert len(e) == len(gold["energies"])
    assert np.max(np.abs(e - gold["energies"])) < 1e-8
    ev = np.array(run.eigenvalues)
    print("eigenvalue error", np.max(np.abs(ev - gold["eigenvalues"])))
    assert np.max(np.abs(ev - gold["eigenvalues"])) < EV_TOL
    l1 = _l1(system, run.rho, gold["rho"])
    print("density L1", l1)
    assert l1 < RHO_L1


def test_scf_reference_fixed_tolerance(dz, ctx, gold, system):
    """The reference's eigensolver schedule as written (warm-up tol 1e-3, then a fixed 1e-4 Ry with
    reused vectors): the energy reaches the oracle's within 1e-6 Ha, but once the start residual
    is below 1e-4 the solver returns at once, so the density is built from vectors with 1e-4
    residuals and the fixed point (eigenvalues) is only good to about 1e-5 Ry."""
    drv = dscf.ScfDriver(system)
    try:
        run = drv.run(max_iter=60, conv=1e-9, precision=FP64, final_tol=1e-10)
    finally:
        drv.close()
    e_ref = gold["energies"][-1]
    ev_err = np.max(np.abs(np.array(run.eigenvalues) - gold["eigenvalues"]))
    print("iterations", len(run.records), "E - E_ref", run.energies[-1] - e_ref, "eigenvalue error", ev_err,
          [r.eig_iterations for r in run.records])
    assert run.converged
    assert abs(run.energies[-1] - e_ref) < E_TOL
    assert run.time_to(e_ref) is not None
    assert ev_err < 1e-4


@pytest.mark.parametrize("precision,concurrent,precond", [(FP64, False, 0), (FP32, True, 0), (FP32, True, PRECOND_TPA)])
def test_scf_adaptive_tolerance(dz, ctx, gold, system, precision, concurrent, precond):
    """Eigensolver tolerance tied to the density change (min(1e-4, max(1e-9, drho))): FP64 or
    mixed FP32-apply (FP64 below 1e-5), k-points serial or on their own streams, kinetic or TPA
    preconditioner: the oracle's fixed point within the eigenvalue, energy and density bars."""
    drv = dscf.ScfDriver(system, concurrent_k=concurrent)
    try:
        run = drv.run(max_iter=60, conv=1e-9, precision=precision, preconditioner=precond, final_tol=1e-10,
                      eig_tol_ratio=1.0, eig_tol_min=1e-9)
    finally:
        drv.close()
    e_ref = gold["energies"][-1]
    ev = np.array(run.eigenvalues)
    l1 = _l1(system, run.rho, gold["rho"])
    print("iterations", len(run.records), "E - E_ref", run.energies[-1] - e_ref, "t(1e-6 Ha)", run.time_to(e_ref),
          "eigenvalue error", np.max(np.abs(ev - gold["eigenvalues"])), "density L1", l1,
          [(r.eig_iterations, r.fp32_applies) for r in run.records])
    assert run.converged
    assert abs(run.energies[-1] - e_ref) < E_TOL
    assert run.time_to(e_ref) is not None
    assert np.max(np.abs(ev - gold["eigenvalues"])) < EV_TOL
    assert l1 < RHO_L1
    if precision == FP32:
        assert sum(r.fp32_applies for r in run.records) > 0


def test_warm_start_and_device_density(dz, ctx, system):
    """reuse=True restarts from the device Ritz vectors; the density of the device solution
    equals the density of the returned vectors."""
    drv = dscf.ScfDriver(system)
    try:
        ham = drv.hams[1]
        ham.set_local(drv.v_ion)
        v1, x1, st1 = ham.eig_solve(20, tol=1e-9, max_iter=300)
        v2, x2, st2 = ham.eig_solve(20, tol=1e-9, max_iter=300, reuse=True)
        print(st1["iterations"], st2["iterations"])
        assert st1["converged"] and st2["converged"]
        assert st2["iterations"] <= 2
        assert np.max(np.abs(v1 - v2)) < 1e-10
        w = dscf.band_weights(system.nelec, 20, 0.5)
        r_dev = ham.density_solution(w)
        r_host = ham.density(x2, w)
        assert np.max(np.abs(r_dev - r_host)) < 1e-13 * max(1.0, np.max(np.abs(r_host)))
        assert abs(np.sum(r_dev) * system.volume / r_dev.shape[0] - 16.0) < 1e-9
    finally:
        drv.close()
