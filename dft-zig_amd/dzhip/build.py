"""Build libdzhip.so (all HIP sources) in-tree for gfx950 with hipcc.

Usage: python dft-zig_amd/dzhip/build.py   (or call build()).
Each source is compiled to an object under dft-zig_amd/_build/ (in parallel), then linked; the
library lands next to this file so that it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR.parent / "csrc"
OBJ_DIR = PKG_DIR.parent / "_build"
LIB_PATH = PKG_DIR / "libdzhip.so"
SOURCES = ["api.hip", "fft.hip", "hpsi.hip", "eig.hip", "precond.hip", "potential.hip", "heev.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Packed FP32 (v_pk_*_f32) is kept out of the generated code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    root = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(root) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    OBJ_DIR.mkdir(exist_ok=True)
    srcs = list(SOURCES)
    objs = [OBJ_DIR / (Path(s).stem + ".o") for s in srcs]
    cmds = [[hipcc(), *HIPCC_FLAGS, "-c", "-o", str(o), str(CSRC / s)] for s, o in zip(srcs, objs)]
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB_PATH.with_suffix(".so.tmp")
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-lrocblas", "-lrocsolver"], verbose)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
