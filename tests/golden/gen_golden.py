"""Writes the golden H.psi / eigen / FFT vectors used by the GPU parity tests.

Inputs are the seeded synthetic workloads (dzhip.workloads); outputs come from
the FP64 oracle (oracle/), itself pinned by tests/test_oracle_*.py.  Each file
stores a SHA-256 of its inputs so a drift in the generator is detected.

Run:  python tests/golden/gen_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "dft-zig_amd"))

from dzhip import workloads  # noqa: E402
from oracle import fftconv, hamiltonian  # noqa: E402

sys.path.insert(0, str(HERE.parent))
from test_oracle_physics import oracle_ctx  # noqa: E402
from golden_util import input_hash  # noqa: E402


def hpsi_file(name: str, nb: int | None = None):
    w = workloads.BUILDERS[name]()
    ctx = oracle_ctx(w)
    psi = w.psi if nb is None else w.psi[:nb]
    y = hamiltonian.apply_hamiltonian_batched(ctx, psi)
    ynl = hamiltonian.apply_nonlocal_batched(ctx, psi)
    out = {"y": y, "y_nonlocal": ynl, "nbands": np.array(psi.shape[0]), "sha256": np.array(input_hash(w))}
    if ctx.npw <= 400:
        h = hamiltonian.dense_hamiltonian(ctx)
        ev, _ = np.linalg.eigh(h)
        out["eigvals"] = ev[:12]
    np.savez_compressed(HERE / f"hpsi_{name}.npz", **out)
    print(name, ctx.npw, psi.shape[0], input_hash(w)[:16])


def fft_file():
    rng = np.random.default_rng(20261202)
    out = {}
    for grid in ((12, 10, 8), (6, 5, 4)):
        n = int(np.prod(grid))
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        key = "x".join(map(str, grid))
        out[f"in_{key}"] = x
        out[f"fwd_{key}"] = fftconv.forward3d(x, grid)
        out[f"inv_{key}"] = fftconv.inverse3d(x, grid)
    np.savez_compressed(HERE / "fft_grids.npz", **out)
    print("fft", list(out)[:3])


def scf_file():
    """Oracle SCF of the small Si case: energies per iteration, final eigenvalues, density."""
    from dzhip import scf as dscf
    from oracle import scf as oscf

    system = dscf.si8_small()
    res = oscf.run_scf(system, n_iter=40, beta=0.3, conv=1e-10, pulay_history=8, pulay_start=4)
    assert res.converged
    key = np.concatenate([np.asarray(system.atom_pos).ravel(), np.ravel(system.kpoints),
                          [system.ecut_ry, system.nbands, *system.grid]])
    np.savez_compressed(HERE / "scf_si8_small.npz", energies=np.array(res.energies), drho=np.array(res.drho),
                        eigenvalues=np.array(res.eigenvalues), rho=res.rho, key=key)
    print("scf", len(res.energies), res.energies[-1])


def scf_al_file():
    """Oracle SCF of the small Al case (displaced atom) with Fermi-Dirac smearing (0.01 Ry)."""
    from dzhip import scf as dscf
    from oracle import scf as oscf

    system = dscf.al4_small()
    res = oscf.run_scf(system, n_iter=60, beta=0.3, conv=1e-12, pulay_history=8, pulay_start=4)
    assert res.converged
    np.savez_compressed(HERE / "scf_al4_small.npz", energies=np.array(res.energies), drho=np.array(res.drho),
                        eigenvalues=np.array(res.eigenvalues), rho=res.rho)
    print("scf al", len(res.energies), res.energies[-1])


def scf_al_pbe_file():
    """Oracle SCF of the small Al case (displaced atom) with PBE and Fermi-Dirac smearing."""
    from dzhip import scf as dscf
    from oracle import scf as oscf

    system = dscf.al4_small(xc="pbe")
    res = oscf.run_scf(system, n_iter=60, beta=0.3, conv=1e-12, pulay_history=8, pulay_start=4)
    assert res.converged
    np.savez_compressed(HERE / "scf_al4_pbe_small.npz", energies=np.array(res.energies), drho=np.array(res.drho),
                        eigenvalues=np.array(res.eigenvalues), rho=res.rho)
    print("scf al pbe", len(res.energies), res.energies[-1])


def scf_upf_file():
    """Oracle SCF of the small Si case with its species read from a UPF file (the GTH-form Si set
    written as a UPF v2 file on a 1200-point log mesh, si_gth.upf): radial tables and the
    short-range local form factor from the file's arrays."""
    import hashlib

    from dzhip import scf as dscf
    from dzhip import upf
    from oracle import scf as oscf

    g = dscf.SI_GTH
    path = HERE / "si_gth.upf"
    upf.write_upf(path, "Si", g["z"], g["rloc"], g["c1"], [0, 1], [g["r_s"], g["r_p"]],
                  np.diag([2 * g["h_s"], 2 * g["h_p"]]))
    u = upf.read_upf(path)
    system = dscf.with_upf(dscf.si8_small(), [u])
    res = oscf.run_scf(system, n_iter=40, beta=0.3, conv=1e-10, pulay_history=8, pulay_start=4)
    assert res.converged
    sha = hashlib.sha256(path.read_bytes()).hexdigest()
    np.savez_compressed(HERE / "scf_si8_upf.npz", energies=np.array(res.energies), drho=np.array(res.drho),
                        eigenvalues=np.array(res.eigenvalues), rho=res.rho, upf_sha256=np.array(sha))
    print("scf upf", len(res.energies), res.energies[-1])


if __name__ == "__main__":
    which = sys.argv[1:] or ["hpsi", "fft", "scf", "scf_al", "scf_al_pbe", "scf_upf"]
    if "hpsi" in which:
        hpsi_file("tiny")
        hpsi_file("si8")
    if "fft" in which:
        fft_file()
    if "scf" in which:
        scf_file()
    if "scf_al" in which:
        scf_al_file()
    if "scf_al_pbe" in which:
        scf_al_pbe_file()
    if "scf_upf" in which:
        scf_upf_file()
