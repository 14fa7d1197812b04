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
from oracle import scf as oscf

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
    """Eigensolver tolerance tied to the density change (min(1e-4, max(1e-10, 0.3 drho))): FP64 or
    mixed FP32-apply (FP64 below 1e-5), k-points serial or on their own streams, kinetic or TPA
    preconditioner: the oracle's fixed point within the eigenvalue, energy and density bars."""
    drv = dscf.ScfDriver(system, concurrent_k=concurrent)
    try:
        run = drv.run(max_iter=60, conv=1e-11, precision=precision, preconditioner=precond, final_tol=1e-10,
                      eig_tol_ratio=0.3, eig_tol_min=1e-10)
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


@pytest.mark.parametrize("precision,concurrent", [(FP64, False), (FP32, True)])
def test_scf_device_grid_matches_host(dz, ctx, gold, system, precision, concurrent):
    """rho, V and Pulay on the device (dzhip_scf_grid): the oracle's trajectory with tight solves,
    and the same fixed point as the host path."""
    drv = dscf.ScfDriver(system, concurrent_k=concurrent, device_grid=True)
    try:
        # FP32 applies: the x-line kernel the autotuner times fastest may differ from run to run,
        # and with it the rounding of the early iterations; the way down to conv = 1e-11 then
        # takes a varying number of steps (up to about 60 seen), hence the headroom.
        run = drv.run(max_iter=60 if precision == FP64 else 90, conv=1e-11, precision=precision, final_tol=1e-10,
                      eig_tol=1e-10 if precision == FP64 else 1e-4,
                      eig_max_iter=400, warmup_steps=0 if precision == FP64 else 2, eig_tol_ratio=None if precision == FP64 else 0.3,
                      eig_tol_min=1e-10)
        v_dev = drv.grid.read(dz.GRID_VIN)
    finally:
        drv.close()
    e = np.array(run.energies)
    ev = np.array(run.eigenvalues)
    l1 = _l1(system, run.rho, gold["rho"])
    print("iterations", len(e), "E - E_ref", e[-1] - gold["energies"][-1], "eigenvalue error",
          np.max(np.abs(ev - gold["eigenvalues"])), "density L1", l1)
    assert run.converged
    if precision == FP64:
        # Same trajectory; near 1e-11 the stopping iteration depends on summation order.
        m = min(len(e), len(gold["energies"]))
        assert abs(len(e) - len(gold["energies"])) <= 3
        assert np.max(np.abs(e[:m] - gold["energies"][:m])) < 1e-8
    assert abs(e[-1] - gold["energies"][-1]) < E_TOL
    assert np.max(np.abs(ev - gold["eigenvalues"])) < EV_TOL
    assert l1 < RHO_L1
    assert np.all(np.isfinite(v_dev))


def test_scf_grid_potential_matches_host(dz, ctx, system):
    """One hxc evaluation on the device against the host formulas (FP64, 1e-12 relative)."""
    drv = dscf.ScfDriver(system, device_grid=True)
    try:
        g = drv.grid
        g.init(system.nelec)
        rho0 = g.read(dz.GRID_RHO)
        n = rho0.shape[0]
        assert np.allclose(rho0, system.nelec / system.volume, rtol=0, atol=1e-15)
        vh, eh = dscf.hartree(rho0, system.grid, drv.g2, system.volume)
        v_host = vh + dscf.lda_pz(rho0)[1]
        np.testing.assert_allclose(g.read(dz.GRID_VIN), v_host, rtol=0, atol=1e-12 * np.max(np.abs(v_host)))
        np.testing.assert_allclose(g.read(dz.GRID_LOCAL), drv.v_ion + v_host, rtol=0,
                                   atol=1e-12 * np.max(np.abs(drv.v_ion + v_host)))
        # One solve, density sum, hxc: energies against the host formulas.
        ham = drv.hams[0]
        g.set_local(ham)
        vals, _, _ = ham.eig_solve(system.nbands, tol=1e-9, max_iter=300, want_vectors=False)
        w = dscf.band_weights(system.nelec, system.nbands, 1.0)
        g.add_density(ham, w)
        rho_h = ham.density_solution(w)
        e_h, e_xc, vin_rho, drho = g.hxc()
        vh2, eh2 = dscf.hartree(rho_h, system.grid, drv.g2, system.volume)
        f2, vx2 = dscf.lda_pz(rho_h)
        assert abs(e_h - eh2) < 1e-11 * abs(eh2)
        assert abs(e_xc - float(np.sum(f2)) * system.volume / n) < 1e-11 * abs(e_xc)
        assert abs(vin_rho - float(np.dot(v_host, rho_h)) * system.volume / n) < 1e-11 * abs(vin_rho)
        assert abs(drho - np.sqrt(np.mean((rho_h - rho0) ** 2))) < 1e-12
        np.testing.assert_allclose(g.read(dz.GRID_VOUT), vh2 + vx2, rtol=0, atol=1e-12 * np.max(np.abs(vh2 + vx2)))
        np.testing.assert_allclose(g.read(dz.GRID_RHO), rho_h, rtol=0, atol=1e-14)
    finally:
        drv.close()


@pytest.mark.parametrize("concurrent,device_grid", [(False, False), (True, True)])
def test_scf_smearing_al(dz, ctx, concurrent, device_grid):
    """Fermi-Dirac smearing (fermi_level.zig, kpoint/density.zig:398-404) on fcc Al, 8 k-points
    with different plane-wave counts (one atom displaced, so no partially occupied level is
    degenerate): the oracle's trajectory and fixed point."""
    gold = np.load(GOLD.parent / "scf_al4_small.npz")
    s = dscf.al4_small()
    drv = dscf.ScfDriver(s, concurrent_k=concurrent, device_grid=device_grid)
    try:
        run = drv.run(max_iter=60, conv=1e-12, eig_tol=1e-11, eig_max_iter=600, warmup_steps=0, final_tol=1e-11)
    finally:
        drv.close()
    e = np.array(run.energies)
    m = min(len(e), len(gold["energies"]))
    print("iterations", len(e), len(gold["energies"]), "dE", np.max(np.abs(e[:m] - gold["energies"][:m])),
          "eig", np.max(np.abs(np.array(run.eigenvalues) - gold["eigenvalues"])))
    assert run.converged
    assert abs(len(e) - len(gold["energies"])) <= 3
    assert np.max(np.abs(e[:m] - gold["energies"][:m])) < 1e-8
    assert np.max(np.abs(np.array(run.eigenvalues) - gold["eigenvalues"])) < EV_TOL
    assert float(np.sum(np.abs(run.rho - gold["rho"]))) * s.volume / run.rho.size < RHO_L1


def test_scf_dielectric_pulay_host_and_device_vs_oracle(dz, ctx, system):
    """diemac = 12 model-dielectric Pulay (mixing.zig:309-338): host and device drivers follow
    the oracle's trajectory with tight eigen-solves."""
    from oracle import scf as oscf

    ref = oscf.run_scf(system, n_iter=40, beta=0.3, conv=1e-10, pulay_history=8, pulay_start=4, diemac=12.0)
    assert ref.converged
    for device_grid in (False, True):
        drv = dscf.ScfDriver(system, device_grid=device_grid)
        try:
            run = drv.run(max_iter=40, conv=1e-10, eig_tol=1e-10, eig_max_iter=400, warmup_steps=0, diemac=12.0)
        finally:
            drv.close()
        e = np.array(run.energies)
        m = min(len(e), len(ref.energies))
        print(device_grid, len(e), len(ref.energies), np.max(np.abs(e[:m] - np.array(ref.energies[:m]))))
        assert run.converged and abs(len(e) - len(ref.energies)) <= 3
        assert np.max(np.abs(e[:m] - np.array(ref.energies[:m]))) < 1e-8


def test_scf_grid_pbe_matches_oracle(dz, ctx):
    """PBE on the device grid (gradient, point-wise PBE, divergence term) against the oracle's
    xc_fields on an atomic-superposition density and on a solved density (1e-11 relative)."""
    s = dscf.al4_small(xc="pbe")
    drv = dscf.ScfDriver(s, device_grid=True)
    try:
        g = drv.grid
        rho0 = dscf.atomic_density(s, drv.g, drv.g2)
        g.init_rho(rho0)
        n = rho0.shape[0]
        vh, _ = oscf.hartree(rho0, s.grid, s.recip, s.volume)
        _, vx = oscf.xc_fields(rho0, s.grid, s.recip, "pbe")
        _, vx_lda = oscf.xc_fields(rho0, s.grid, s.recip, "lda_pz")
        assert np.max(np.abs(vx - vx_lda)) > 1e-3  # the gradient terms are not negligible here
        np.testing.assert_allclose(g.read(dz.GRID_VIN), vh + vx, rtol=0, atol=1e-11 * np.max(np.abs(vh + vx)))
        ham = drv.hams[0]
        g.set_local(ham)
        ham.eig_solve(s.nbands, tol=1e-9, max_iter=300, want_vectors=False)
        w = dscf.band_weights(s.nelec, s.nbands, 1.0)
        g.add_density(ham, w)
        rho_h = ham.density_solution(w)
        e_h, e_xc, _, _ = g.hxc()
        vh2, eh2 = oscf.hartree(rho_h, s.grid, s.recip, s.volume)
        f2, vx2 = oscf.xc_fields(rho_h, s.grid, s.recip, "pbe")
        print("E_xc", e_xc, float(np.sum(f2)) * s.volume / n, "max |dV|", np.max(np.abs(g.read(dz.GRID_VOUT) - vh2 - vx2)))
        assert abs(e_h - eh2) < 1e-11 * abs(eh2)
        assert abs(e_xc - float(np.sum(f2)) * s.volume / n) < 1e-11 * abs(e_xc)
        np.testing.assert_allclose(g.read(dz.GRID_VOUT), vh2 + vx2, rtol=0, atol=1e-11 * np.max(np.abs(vh2 + vx2)))
    finally:
        drv.close()


@pytest.mark.parametrize("concurrent", [False, True])
def test_scf_pbe_al(dz, ctx, concurrent):
    """PBE + Fermi-Dirac SCF on the device grid (xc_fields.zig:104-140): the oracle's trajectory
    and fixed point on the small displaced Al cell."""
    gold = np.load(GOLD.parent / "scf_al4_pbe_small.npz")
    s = dscf.al4_small(xc="pbe")
    drv = dscf.ScfDriver(s, concurrent_k=concurrent, device_grid=True)
    try:
        run = drv.run(max_iter=60, conv=1e-12, eig_tol=1e-11, eig_max_iter=600, warmup_steps=0, final_tol=1e-11)
    finally:
        drv.close()
    e = np.array(run.energies)
    m = min(len(e), len(gold["energies"]))
    print("iterations", len(e), len(gold["energies"]), "dE", np.max(np.abs(e[:m] - gold["energies"][:m])),
          "eig", np.max(np.abs(np.array(run.eigenvalues) - gold["eigenvalues"])),
          "L1", float(np.sum(np.abs(run.rho - gold["rho"]))) * s.volume / run.rho.size)
    assert run.converged
    assert abs(len(e) - len(gold["energies"])) <= 3
    assert np.max(np.abs(e[:m] - gold["energies"][:m])) < 1e-8
    assert np.max(np.abs(np.array(run.eigenvalues) - gold["eigenvalues"])) < EV_TOL
    assert float(np.sum(np.abs(run.rho - gold["rho"]))) * s.volume / run.rho.size < RHO_L1
