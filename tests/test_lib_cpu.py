"""CPU-side checks of the product library: it builds, loads, exports every symbol the
header declares, and its host-side argument checks fire without a GPU."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    text = (ROOT / "include" / "dzhip.h").read_text()
    return sorted(set(re.findall(r"\b(dzhip_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_exports(dz):
    assert sorted(dz.EXPORTS) == header_symbols()


def test_library_exports_every_declared_symbol(dz):
    lib = dz.L()
    for name in header_symbols():
        assert hasattr(lib, name), name
    assert lib.dzhip_version() >= 100


def test_null_handles_are_rejected(dz):
    lib = dz.L()
    x = np.zeros(4, complex)
    rc = lib.dzhip_ham_apply(None, dz._dp(x.view(np.float64)), dz._dp(x.view(np.float64)), 2, 2, 0)
    assert rc == -4 and "NULL" in dz.last_error()
    rc = lib.dzhip_fft_forward(None, dz._dp(x.view(np.float64)), 4)
    assert rc == -4


def test_fft_and_radix_tables_cover_reference_sizes():
    """Sizes the reference's auto_grid produces for the configs (32, 30, 72, 75, 90) factor into
    radices the kernels implement; the digit-reversal used by the in-place transforms is a
    permutation whose DIT/DIF model reproduces numpy's FFT."""
    import sys

    sys.path.insert(0, str(ROOT / "tests"))
    from fft_model import check_sizes

    check_sizes([2, 3, 4, 5, 6, 8, 12, 30, 32, 72, 75, 90, 96, 128])


def test_fft_core_runs_on_host_against_dft(tmp_path):
    """The kernels' own LDS FFT stage code (fft_lds.h), compiled for the host and run
    sequentially: every size 1..256 used by the sizing rules, radices 2/3/4/5/8/9, DIT and
    DIF, both signs, strided and listed lines; FP64 <= 1e-13, FP32 <= 2e-6 relative."""
    import shutil
    import subprocess

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "fft_core_host"
    src = [str(ROOT / "tests" / "native" / "fft_core_host.hip"), str(ROOT / "dft-zig_amd" / "csrc" / "fft.hip")]
    r = subprocess.run([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-o", str(exe), *src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "FFT core ok" in r.stdout, r.stdout[-3000:]


def test_herm3_jacobi_host(tmp_path):
    """The CG solver's 3x3 Hermitian eigensolver (csrc/herm3.h, host code) against numpy eigh,
    including degenerate and already-diagonal matrices."""
    import subprocess

    exe = tmp_path / "herm3_host"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "dft-zig_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "herm3_host.cpp"), "-o", str(exe)], check=True)
    rng = np.random.default_rng(3)
    mats = []
    for _ in range(200):
        a = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
        mats.append(a + a.conj().T)
    mats.append(np.diag([3.0, -1.0, 2.0]).astype(complex))
    mats.append(np.diag([1.0, 1.0, 2.0]).astype(complex))
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3)))
    mats.append(q @ np.diag([0.5, 0.5, 4.0]) @ q.conj().T)
    inp = "\n".join(" ".join(f"{z.real:.17g} {z.imag:.17g}" for z in m.ravel()) for m in mats)
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for m, line in zip(mats, out):
        vals = [float(t) for t in line.split()]
        lam = vals[0]
        v = np.array([vals[1] + 1j * vals[2], vals[3] + 1j * vals[4], vals[5] + 1j * vals[6]])
        ev = np.linalg.eigvalsh(m)
        assert abs(lam - ev[0]) < 1e-12 * max(1.0, abs(ev[0]))
        assert abs(np.linalg.norm(v) - 1.0) < 1e-12
        assert np.linalg.norm(m @ v - lam * v) < 1e-11 * max(1.0, np.max(np.abs(ev)))
