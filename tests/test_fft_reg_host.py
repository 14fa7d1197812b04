"""The register-resident line FFT (dft-zig_amd/csrc/fft_reg.h: mixed-radix Stockham with
radix-2/3/4/5/8/9 butterflies) built for the host with g++ and compared with numpy's FFT for every
factorisation the x-line kernel uses, both signs, FP64 (1e-13) and FP32 (2e-6 of the largest
output).  The device kernel runs the same code."""
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
        d64 = vals[pos:pos + n, 0] + 1j * vals[pos:pos + n, 1]
        d32 = vals[pos + n:pos + 2 * n, 0] + 1j * vals[pos + n:pos + 2 * n, 1]
        pos += 2 * n
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(d64 - ref)) < 1e-13 * scale, (n, sign)
        assert np.max(np.abs(d32 - ref)) < 2e-6 * scale, (n, sign)
    assert pos == vals.shape[0]
