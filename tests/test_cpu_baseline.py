"""The timed CPU baseline (oracle/cpu_baseline.py) runs the oracle's algorithm: its scipy-FFT
H.psi equals the oracle's numpy-FFT apply to 1e-12 on the golden workloads, and its k-thread and
all-cores drivers apply every band they report."""
import sys
from pathlib import Path

import numpy as np

from dzhip import workloads
from oracle import cpu_baseline, hamiltonian

sys.path.insert(0, str(Path(__file__).resolve().parent))
from test_oracle_physics import oracle_ctx  # noqa: E402


def test_scipy_fft_apply_equals_oracle():
    for make in (workloads.tiny, workloads.si8):
        w = make()
        ctx = oracle_ctx(w)
        ref = hamiltonian.apply_hamiltonian_batched(ctx, w.psi)
        got = cpu_baseline.apply_hamiltonian_batched(ctx, w.psi)
        assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref)), w.name


def test_thread_drivers_count_bands():
    w = workloads.tiny(nbands=6)
    ctx = oracle_ctx(w)
    r = cpu_baseline.kthread_hpsi([ctx, ctx, ctx], [w.psi[:2], w.psi[2:5], w.psi[5:]], threads=2)
    assert r["bands"] == 6 and r["seconds"] > 0
    r = cpu_baseline.allcores_hpsi(ctx, w.psi, cores=2)
    assert r["bands"] == 6
    d = cpu_baseline.davidson_iteration_algebra(200, 24, 8, cores=2)
    assert d["total_s"] > 0
    h = cpu_baseline.host_description()
    assert h["nproc"] >= 1 and h["cores_available"] >= 1
