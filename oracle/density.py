"""Band density accumulation rho(r) += w |psi(r)|^2 / Omega (FP64).

Restates (paths relative to the dft-zig source tree):
  accumulate_band_density_fft   src/features/scf/kpoint/density.zig:444-474
    (scatter, remap with scale N, inverse FFT with 1/N => psi(r) = sum_G c_G e^{iG.r})
  occupations / weights         src/features/scf/kpoint/density.zig:98-107
    nocc = ceil(nelec/2), last band nelec/2 - (nocc-1), weight = w_k * occ * 2
"""
from __future__ import annotations

import math

import numpy as np

from . import fftconv, pw


def band_weights(nelec: float, nbands: int, kweight: float) -> np.ndarray:
    nocc = int(math.ceil(nelec / 2.0))
    occ_last = 0.0 if nocc == 0 else nelec / 2.0 - (nocc - 1)
    w = np.zeros(nbands)
    for b in range(min(nocc, nbands)):
        occ = occ_last if b + 1 == nocc else 1.0
        if occ > 0.0:
            w[b] = kweight * occ * 2.0
    return w


def psi_real(fft_idx: np.ndarray, grid: tuple, coeffs: np.ndarray) -> np.ndarray:
    total = int(np.prod(grid))
    work = pw.scatter_fft(fft_idx, coeffs, total, float(total))
    return fftconv.inverse3d(work, grid)


def accumulate_density(fft_idx: np.ndarray, grid: tuple, psi: np.ndarray, weights: np.ndarray,
                       inv_volume: float, rho: np.ndarray | None = None) -> np.ndarray:
    """psi shape (nbands, npw); bands with weight 0 are skipped as the reference does."""
    total = int(np.prod(grid))
    if rho is None:
        rho = np.zeros(total)
    for b in range(psi.shape[0]):
        if weights[b] <= 0.0:
            continue
        pr = psi_real(fft_idx, grid, psi[b])
        rho += weights[b] * (pr.real * pr.real + pr.imag * pr.imag) * inv_volume
    return rho
