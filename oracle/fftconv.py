"""3D complex FFT with the reference's conventions (FP64).

Reference conventions (paths relative to the dft-zig source tree):
  forward  X_k = sum_j x_j e^{-2 pi i jk/n}, unnormalised   src/lib/fft/mixed_radix.zig:47-61
  inverse  x_j = (1/n) sum_k X_k e^{+2 pi i jk/n}            src/lib/fft/fft.zig:70-78,
                                                             src/lib/fft/fftw_fft.zig:151-166
  3D layout index x + nx*(y + ny*z); FFTW planned with dims (nz, ny, nx)
                                                             src/lib/fft/fftw_fft.zig:91-112
numpy.fft uses exactly these conventions (fftn unnormalised, ifftn 1/N) when
the flat grid is viewed as an array of shape (nz, ny, nx).
"""
from __future__ import annotations

import numpy as np


def forward3d(data: np.ndarray, grid: tuple[int, int, int]) -> np.ndarray:
    nx, ny, nz = grid
    return np.fft.fftn(np.asarray(data, dtype=np.complex128).reshape(nz, ny, nx)).ravel()


def inverse3d(data: np.ndarray, grid: tuple[int, int, int]) -> np.ndarray:
    nx, ny, nz = grid
    return np.fft.ifftn(np.asarray(data, dtype=np.complex128).reshape(nz, ny, nx)).ravel()


def dft_direct_1d(x: np.ndarray, inverse: bool) -> np.ndarray:
    """O(n^2) textbook DFT, unnormalised in both directions (mixed_radix.zig:47-61)."""
    x = np.asarray(x, dtype=np.complex128)
    n = x.shape[0]
    sign = 1.0 if inverse else -1.0
    j = np.arange(n)
    w = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
    return w @ x


def dft_direct_3d(data: np.ndarray, grid: tuple[int, int, int], inverse: bool) -> np.ndarray:
    """Brute-force separable DFT along x, y, z; inverse includes the 1/N factor."""
    nx, ny, nz = grid
    a = np.asarray(data, dtype=np.complex128).reshape(nz, ny, nx)
    sign = 1.0 if inverse else -1.0

    def mat(n):
        j = np.arange(n)
        return np.exp(sign * 2j * np.pi * np.outer(j, j) / n)

    a = np.einsum("kx,zyx->zyk", mat(nx), a)
    a = np.einsum("ky,zyx->zkx", mat(ny), a)
    a = np.einsum("kz,zyx->kyx", mat(nz), a)
    out = a.ravel()
    if inverse:
        out = out / (nx * ny * nz)
    return out
