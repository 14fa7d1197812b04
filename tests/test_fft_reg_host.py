"""The register-resident line FFT (dft-zig_amd/csrc/fft_reg.h: mixed-radix Stockham with
radix-2/3/4/5/8/9 butterflies) built for the host with g++ and compared with numpy's FFT for every
factorisation the x-line kernel uses, both signs, FP64 (1e-13) and FP32 (2e-6 of the largest
output), with the twiddles from a table and as compile-time constants (the device kernel's
form); the compile-time cos / sin table against numpy to 2 units in the last place."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SIZES = [8, 9, 12, 15, 16, 18, 24, 25, 27, 30, 32, 36, 45, 64, 72, 75, 81, 90, 96, 128]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("fftreg") / "fft_reg_driver"
    cmd = ["g++", "-O1", "-std=c++17", "-I", str(ROOT / "dft-zig_amd" / "csrc"),
           str(ROOT / "tests" / "fft_reg_driver.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_register_fft_matches_numpy(driver):
    rng = np.random.default_rng(11)
    lines, cases = [], []
    for n in SIZES:
        for sign in (-1, 1):
            x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            cases.append((n, sign, x))
            lines.append(f"{n} {sign} " + " ".join(f"{v.real:.17g} {v.imag:.17g}" for v in x))
    out = subprocess.run([str(driver)], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    vals = np.array([list(map(float, l.split())) for l in out.stdout.strip().splitlines()])
    pos = 0
    for n, sign, x in cases:
        ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
        scale = np.max(np.abs(ref))
        for form in ("table", "constant"):
            d64 = vals[pos:pos + n, 0] + 1j * vals[pos:pos + n, 1]
            d32 = vals[pos + n:pos + 2 * n, 0] + 1j * vals[pos + n:pos + 2 * n, 1]
            pos += 2 * n
            assert np.max(np.abs(d64 - ref)) < 1e-13 * scale, (n, sign, form)
            assert np.max(np.abs(d32 - ref)) < 2e-6 * scale, (n, sign, form)
    assert pos == vals.shape[0]


def test_compile_time_roots(driver):
    lines = "\n".join(f"{n} 0" for n in SIZES) + "\n"
    out = subprocess.run([str(driver)], input=lines, capture_output=True, text=True, check=True)
    vals = np.array([list(map(float, l.split())) for l in out.stdout.strip().splitlines()])
    pos = 0
    for n in SIZES:
        k = np.arange(n)
        # The angle in long double (with pi to long-double accuracy), so the reference itself is
        # good to well below an ulp of the double values.
        pi = np.longdouble("3.14159265358979323846264338327950288")
        ang = 2 * pi * k.astype(np.longdouble) / n
        c, s = np.cos(ang), np.sin(ang)
        got = vals[pos:pos + n]
        pos += n
        # Two units in the last place of the double value (the zeros are exact, below).
        for col, ref in ((0, c), (1, s)):
            ulp = np.spacing(np.abs(ref.astype(np.float64)))
            assert np.all(np.abs(got[:, col] - ref) <= 2.0 * ulp + 1e-18), (n, col)
        # Exact at the quadrant points.
        for q in range(4):
            if (q * n) % 4 == 0:
                kk = q * n // 4
                assert (got[kk, 0], got[kk, 1]) == ((1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, -1.0))[q]
    assert pos == vals.shape[0]


def lane_split_xline(b, v, nx, lanes):
    """Model of the lane-split register x-line kernels (k_xline_pair for lanes = 2, k_xline_tri for
    lanes = 3): rows b[u], u < np, of a line of length nx; V sampled at every real-space point.  Lane
    t folds and twists the rows, c_t[u0] = sum_j b[u0 + jM] w^{(u0 + jM) t} with w = e^{2 pi i / nx},
    runs an inverse length-M transform, multiplies by V(lanes m + t), a forward length-M transform,
    twists by w^{-k0 t}, and the lanes combine by a lanes-point transform across them."""
    m = nx // lanes
    npl = b.shape[0]
    w = np.exp(2j * np.pi / nx)
    h = []
    for t in range(lanes):
        c = np.zeros(m, complex)
        for u in range(npl):
            c[u % m] += b[u] * w ** (u * t)
        g = np.fft.ifft(c) * m
        f = np.fft.fft(g * v[lanes * np.arange(m) + t])
        h.append(f * w ** (-np.arange(m) * t))
    out = np.zeros(npl, complex)
    for k in range(npl):
        k0, j = k % m, k // m
        out[k] = sum(np.exp(-2j * np.pi * j * t / lanes) * h[t][k0] for t in range(lanes))
    return out


@pytest.mark.parametrize("nx,lanes,npl", [(72, 2, 35), (72, 2, 36), (72, 3, 35), (75, 3, 35), (75, 3, 25),
                                          (75, 3, 12), (90, 3, 41), (90, 2, 41), (27, 3, 27)])
def test_lane_split_xline_model(nx, lanes, npl):
    """The folded, twisted length-nx/lanes transforms and the cross-lane combination give the direct
    x-part of H (forward FFT of V times the inverse FFT of the zero-padded rows) on the np rows."""
    rng = np.random.default_rng(nx * 10 + lanes + npl)
    b = rng.standard_normal(npl) + 1j * rng.standard_normal(npl)
    v = rng.standard_normal(nx)
    full = np.zeros(nx, complex)
    full[:npl] = b
    ref = np.fft.fft(v * np.fft.ifft(full) * nx)[:npl]
    got = lane_split_xline(b, v, nx, lanes)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
