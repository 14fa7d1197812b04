"""Timed CPU baseline: the oracle's FP64 H.psi and Davidson-iteration algebra on the host cores.

TEST INFRASTRUCTURE ONLY (the `cpu_baseline` leg of bench.py and its test); the product never
imports it.  It runs the same restated algorithm as ``hamiltonian.apply_hamiltonian_batched``
(src/features/scf/apply.zig:1584-1618: the local part column by column through a 3D FFT pair,
the nonlocal part per atom with real dgemms), with the FFTs taken by scipy.fft (pocketfft with
SIMD kernels, the closest FFTW stand-in here: the reference's CPU path runs FFTW, which cannot
be built in this image) and the dgemms by numpy's BLAS.

Threading follows the reference's k-point workers (src/features/scf/kpoint/workers.zig:182-188):
min(n_k, cores) k-threads, each applying H on its own k-point with one FFT/BLAS thread.  The
all-cores variant instead gives one k-point's FFTs and BLAS every core.

Pinned by tests/test_cpu_baseline.py: the scipy-FFT apply equals the oracle's numpy-FFT apply to
1e-12 on the golden workloads.
"""
from __future__ import annotations

import os
import platform
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.fft as sfft

from . import hamiltonian, pw


def apply_local_scipy(ctx: hamiltonian.HamContext, x: np.ndarray, workers: int = 1) -> np.ndarray:
    """hamiltonian.apply_local (apply.zig:820-852) with the FFT pair from scipy.fft."""
    nx, ny, nz = ctx.grid
    n = float(ctx.total)
    work = pw.scatter_fft(ctx.fft_idx, x, ctx.total, n).reshape(nz, ny, nx)
    work = sfft.ifftn(work, workers=workers, overwrite_x=True)
    work *= ctx.local_r.reshape(nz, ny, nx)
    work = sfft.fftn(work, workers=workers, overwrite_x=True)
    return pw.gather_fft(ctx.fft_idx, work.ravel(), 1.0 / n)


def apply_hamiltonian_batched(ctx: hamiltonian.HamContext, xb: np.ndarray, workers: int = 1) -> np.ndarray:
    """apply_hamiltonian_batched (apply.zig:1584-1618): local per column, nonlocal batched."""
    xb = np.asarray(xb, dtype=np.complex128)
    yb = np.empty_like(xb)
    for c in range(xb.shape[0]):
        yb[c] = ctx.basis.kinetic * xb[c] + apply_local_scipy(ctx, xb[c], workers)
    if ctx.species:
        yb += hamiltonian.apply_nonlocal_batched(ctx, xb)
    return yb


def host_description() -> dict:
    """CPU model and core counts of this host (the threads used are reported per measurement)."""
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "cores_available": avail}


def kthread_hpsi(contexts: list, blocks: list, threads: int) -> dict:
    """Each of `threads` k-threads applies H to its k-points' blocks (one FFT/BLAS thread each);
    returns the wall time and the bands applied."""
    from threadpoolctl import threadpool_limits

    jobs = list(zip(contexts, blocks))

    def run(i):
        done = 0
        for j in range(i, len(jobs), threads):
            ctx, xb = jobs[j]
            apply_hamiltonian_batched(ctx, xb, workers=1)
            done += xb.shape[0]
        return done

    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as pool:
            bands = sum(pool.map(run, range(threads)))
        wall = time.perf_counter() - t0
    return {"seconds": wall, "bands": bands}


def allcores_hpsi(ctx: hamiltonian.HamContext, xb: np.ndarray, cores: int) -> dict:
    """One k-point's block with every FFT and BLAS call on `cores` threads."""
    from threadpoolctl import threadpool_limits

    with threadpool_limits(limits=cores):
        t0 = time.perf_counter()
        apply_hamiltonian_batched(ctx, xb, workers=cores)
        wall = time.perf_counter() - t0
    return {"seconds": wall, "bands": int(xb.shape[0])}


def davidson_iteration_algebra(npw: int, m: int, nb: int, cores: int, seed: int = 5) -> dict:
    """The dense algebra of one block-Davidson iteration at subspace size m on the host, with
    `cores` BLAS/LAPACK threads (lobpcg/serial.zig:55-307 with common.zig:231-259, 440-508):
    the projected matrix of the new columns V^H W[:, m-nb:m], the Hermitian eigen-decomposition of
    the m x m projected matrix (zheev), and the Ritz vectors and their H images V c, W c.  Random
    data of the right shapes; the H apply of the new block is timed separately."""
    from threadpoolctl import threadpool_limits

    rng = np.random.default_rng(seed)
    v = rng.standard_normal((npw, m)) + 1j * rng.standard_normal((npw, m))
    w = rng.standard_normal((npw, m)) + 1j * rng.standard_normal((npw, m))
    t = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    t = t + t.conj().T
    with threadpool_limits(limits=cores):
        t0 = time.perf_counter()
        tnew = v.conj().T @ w[:, m - nb:]
        t1 = time.perf_counter()
        _, c = np.linalg.eigh(t)
        t2 = time.perf_counter()
        x = v @ c[:, :nb]
        hx = w @ c[:, :nb]
        t3 = time.perf_counter()
    del tnew, x, hx
    return {"gram_s": t1 - t0, "eigh_s": t2 - t1, "ritz_s": t3 - t2, "total_s": t3 - t0}
