"""The device path fed from a UPF file (dzhip.upf): the reference reads its pseudopotentials with
pseudopotential.zig:126-257, tabulates each projector's radial transform (RadialTable, 2048
points over [0, sqrt(2 ecut) 1.5], nonlocal.zig:248-338, common.zig:124-147) and fills phi from the
tables (nonlocal_context.zig:387-435, 524-549); V_loc(q) is the file's short-range local form
factor (form_factor.zig:225-256).  Fixture: tests/golden/si_gth.upf, the GTH-form Si set of the
synthetic Si cells written as a UPF v2 file on a 1200-point log mesh.

* H.psi at a Si-8 k-point with phi from dzhip.upf's tables on the device against the oracle apply
  whose species are rebuilt from the same file's raw arrays by the oracle's own RadialTable and
  radial integrals (FP64 1e-12, FP32 2e-5).
* The device SCF of the 8-atom cell with the UPF species against the oracle SCF on the same file
  (tests/golden/scf_si8_upf.npz): every iteration's energy within 1e-8 Ry, eigenvalues within
  2e-8 Ry, density L1 within 1e-7 electrons.
"""
import hashlib
from pathlib import Path

import numpy as np
import pytest

from dzhip import scf as dscf
from dzhip import upf
from oracle import adapters, hamiltonian

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def upf_si():
    path = GOLDEN / "si_gth.upf"
    return upf.read_upf(path), hashlib.sha256(path.read_bytes()).hexdigest()


@pytest.fixture(scope="module")
def system(upf_si):
    return dscf.with_upf(dscf.si8_small(), [upf_si[0]])


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-5)])
def test_hpsi_from_upf_tables_vs_oracle(dz, ctx, system, prec, tol):
    w = dscf.kpoint_workload(system, 1, nbands=8)
    ham = dz.Hamiltonian.from_workload(ctx, w)
    try:
        got = ham.apply(w.psi, precision=prec)
    finally:
        ham.close()
    octx = adapters.context_for_k(system, system.kpoints[1], w.local_r)
    ref = hamiltonian.apply_hamiltonian_batched(octx, w.psi)
    err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    print("UPF H.psi", prec, err)
    assert err < tol
    # The file's tables differ from the closed-form Gaussian transforms by their interpolation
    # error only, so the apply stays close to the closed-form species' one.
    w0 = dscf.kpoint_workload(dscf.si8_small(), 1, nbands=8, local_r=w.local_r)
    h0 = dz.Hamiltonian.from_workload(ctx, w0)
    try:
        y0 = h0.apply(w0.psi, precision=0)
    finally:
        h0.close()
    assert 0 < float(np.max(np.abs(got - y0)) / np.max(np.abs(y0))) < 1e-5


def test_scf_from_upf_matches_oracle(dz, ctx, system, upf_si):
    gold = np.load(GOLDEN / "scf_si8_upf.npz")
    assert str(gold["upf_sha256"]) == upf_si[1]
    drv = dscf.ScfDriver(system)
    try:
        run = drv.run(max_iter=40, conv=1e-10, eig_tol=1e-10, eig_max_iter=400, warmup_steps=0, final_tol=1e-10)
    finally:
        drv.close()
    e = np.array(run.energies)
    print("iterations", len(e), "oracle", len(gold["energies"]), "dE", e[:len(gold["energies"])] - gold["energies"][:len(e)])
    assert run.converged
    # The last iterations' density changes (1e-10 and below) sit at the eigensolver's tolerance
    # (1e-10), so which iteration first falls under conv = 1e-10 is decided by solver noise: the
    # iteration counts may differ by a step or two, the energies of the common iterations may not.
    n = min(len(e), len(gold["energies"]))
    assert abs(len(e) - len(gold["energies"])) <= 2
    assert np.max(np.abs(e[:n] - gold["energies"][:n])) < 1e-8
    assert np.max(np.abs(np.array(run.eigenvalues) - gold["eigenvalues"])) < 2e-8
    l1 = float(np.sum(np.abs(run.rho - gold["rho"]))) * system.volume / run.rho.shape[0]
    assert l1 < 1e-7
