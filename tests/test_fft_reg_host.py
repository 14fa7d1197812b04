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
