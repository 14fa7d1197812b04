"""GPU parity of the HIP H.psi path and FFT backend against the FP64 oracle and golden vectors.

Tolerances (relative to the largest |H psi| entry of the block):
  FP64 path: 1e-12  (north star: H.psi checked against an FP64 oracle; FP64 mode must be exact
             to rounding so that final eigenvalues match to 1e-8 Ha)
  FP32 path: 2e-5   (FP32 FFT / V_loc / projector arithmetic with FP32 accumulation)
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from dzhip import workloads
import dzhip.tables as tables
from oracle import fftconv, hamiltonian

sys.path.insert(0, str(Path(__file__).resolve().parent))
from test_oracle_physics import oracle_ctx  # noqa: E402
from golden_util import input_hash  # noqa: E402

GOLDEN = Path(__file__).resolve().parent / "golden"
TOL64 = 1e-12
TOL32 = 2e-5

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def make_ham(dz, ctx, w):
    return dz.Hamiltonian.from_workload(ctx, w)


@pytest.mark.parametrize("grid", [(12, 10, 8), (6, 5, 4)])
@pytest.mark.parametrize("prec", [0, 1])
def test_fft_golden(dz, ctx, grid, prec):
    g = np.load(GOLDEN / "fft_grids.npz")
    key = "x".join(map(str, grid))
    tol = TOL64 if prec == 0 else TOL32
    plan = dz.FftPlan3d(ctx, *grid, precision=prec)
    x = g[f"in_{key}"].copy()
    plan.forward(x)
    assert relerr(x, g[f"fwd_{key}"]) < tol
    x = g[f"in_{key}"].copy()
    plan.inverse(x)
    assert relerr(x, g[f"inv_{key}"]) < tol


@pytest.mark.parametrize("grid", [(32, 32, 32), (30, 30, 30), (72, 72, 72), (75, 75, 75), (90, 90, 90),
                                  (24, 36, 40)])
def test_fft_sizes_vs_numpy(dz, ctx, grid):
    rng = np.random.default_rng(sum(grid))
    n = int(np.prod(grid))
    x0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for prec, tol in ((0, TOL64), (1, TOL32)):
        plan = dz.FftPlan3d(ctx, *grid, precision=prec)
        x = x0.copy()
        plan.forward(x)
        assert relerr(x, fftconv.forward3d(x0, grid)) < tol, (grid, prec)
        plan.inverse(x)
        assert relerr(x, x0) < tol * 4, (grid, prec)


def test_fft_size_mismatch_and_bad_grid(dz, ctx):
    plan = dz.FftPlan3d(ctx, 8, 8, 8)
    with pytest.raises(dz.InvalidMatrixSize):
        plan.forward(np.zeros(8 * 8 * 7, complex))
    with pytest.raises(dz.BadArgument):
        dz.FftPlan3d(ctx, 7, 8, 8)  # 7 is not 2/3/5-smooth


@pytest.mark.parametrize("name", ["tiny", "si8"])
@pytest.mark.parametrize("prec", [0, 1])
def test_hpsi_golden(dz, ctx, name, prec):
    w = workloads.BUILDERS[name]()
    g = np.load(GOLDEN / f"hpsi_{name}.npz")
    assert str(g["sha256"]) == input_hash(w), "workload generator drifted from the golden inputs"
    ham = make_ham(dz, ctx, w)
    tol = TOL64 if prec == 0 else TOL32
    y = ham.apply(w.psi, precision=prec)
    assert relerr(y, g["y"]) < tol
    ynl = ham.apply_nonlocal(w.psi, precision=prec)
    assert relerr(ynl, g["y_nonlocal"]) < tol
    # Single-vector callback (apply_hamiltonian) equals the batched one.
    y1 = ham.apply(w.psi[1], precision=prec)
    assert relerr(y1, g["y"][1]) < tol


def test_hpsi_bench_shape_vs_oracle(dz, ctx):
    """Full bench shape (72^3, 23,847 PWs, 64 atoms x 4 projectors), 8 bands against the oracle."""
    w = workloads.bench(nbands=8)
    ctx_o = oracle_ctx(w)
    ref = hamiltonian.apply_hamiltonian_batched(ctx_o, w.psi)
    ham = make_ham(dz, ctx, w)
    assert relerr(ham.apply(w.psi, 0), ref) < TOL64
    e32 = relerr(ham.apply(w.psi, 1), ref)
    assert e32 < TOL32
    print("bench fp32 max rel err", e32)


def test_hpsi_hermitian_and_linear_full_block(dz, ctx):
    """Size-independent properties at 160 bands: <a|H b> = <H a|b>, linearity."""
    w = workloads.bench(nbands=160)
    ham = make_ham(dz, ctx, w)
    y = ham.apply(w.psi, 0)
    a, b = w.psi[3], w.psi[77]
    lhs = np.vdot(a, y[77])
    rhs = np.vdot(y[3], b)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    comb = 0.3 * w.psi[5] - 1.7j * w.psi[9]
    yc = ham.apply(comb, 0)
    assert relerr(yc, 0.3 * y[5] - 1.7j * y[9]) < 1e-12


def test_hpsi_column_chunking_and_odd_counts(dz, ctx):
    """ncols above one pass (256) and odd counts give the same columns as small calls."""
    w = workloads.si8(nbands=300)
    ham = make_ham(dz, ctx, w)
    y = ham.apply(w.psi, 0)
    y7 = ham.apply(w.psi[290:297], 0)
    assert relerr(y7, y[290:297]) < 1e-14


def test_hpsi_errors(dz, ctx):
    w = workloads.tiny()
    ham = make_ham(dz, ctx, w)
    with pytest.raises(dz.InvalidMatrixSize):
        ham.apply(w.psi[0][:-1])
    with pytest.raises(dz.InvalidMatrixSize):
        ham.set_local(w.local_r[:-1])
    s = ham.apply_s(w.psi)
    np.testing.assert_array_equal(s, w.psi)
    # An empty block (0 columns) is a no-op in every precision and entry point; with n wrong it is
    # still InvalidMatrixSize.
    empty = np.empty((0, w.npw), dtype=np.complex128)
    for prec in (0, 1):
        assert ham.apply(empty, prec).shape == (0, w.npw)
        assert ham.apply_nonlocal(empty, prec).shape == (0, w.npw)
    with pytest.raises(dz.InvalidMatrixSize):
        ham.apply(np.empty((0, w.npw - 1), dtype=np.complex128))
    # No local potential yet -> BadState.
    ham2 = dz.Hamiltonian(ctx, w.grid, tables.fft_indices(w.basis, w.grid), w.basis.kinetic,
                          w.basis.cart, None, w.inv_volume)
    with pytest.raises(dz.BadState):
        ham2.apply(w.psi[0])
    ham2.set_local(w.local_r)
    ctx_o = oracle_ctx(w)
    ctx_o.species = []
    ref = np.array([hamiltonian.apply_hamiltonian(ctx_o, p) for p in w.psi])
    assert relerr(ham2.apply(w.psi, 0), ref) < TOL64


def test_apply_dev_with_torch_buffers(dz, ctx):
    import torch

    w = workloads.si8(nbands=12)
    ham = make_ham(dz, ctx, w)
    ref = ham.apply(w.psi, 0)
    x = torch.from_numpy(w.psi.view(np.float64).copy()).cuda()
    y = torch.empty_like(x)
    stream = torch.cuda.Stream()
    assert stream.cuda_stream != 0
    ham.set_stream(stream.cuda_stream)
    ham.apply_dev(x.data_ptr(), y.data_ptr(), 12, 0)
    stream.synchronize()
    got = y.cpu().numpy().view(np.complex128).reshape(12, -1)
    assert relerr(got, ref) < 1e-15
    x32 = torch.from_numpy(w.psi.astype(np.complex64).view(np.float32).copy()).cuda()
    y32 = torch.empty_like(x32)
    ham.apply_dev(x32.data_ptr(), y32.data_ptr(), 12, 1)
    stream.synchronize()
    got32 = y32.cpu().numpy().view(np.complex64).reshape(12, -1)
    assert relerr(got32.astype(np.complex128), ref) < TOL32
    ham.set_stream(None)


@pytest.mark.parametrize("prec", [0, 1])
def test_phase_timers(dz, ctx, prec):
    """The ScfProfile-shaped phase timers (src/features/scf/logging.zig:114-181): a profiled apply
    gives bit for bit the unprofiled result, every phase is timed on every call, and the phase
    sum of the 160-band bench block is within 10 % of the block's own time (HIP events around the
    unprofiled applies on the same stream)."""
    import torch

    w = workloads.bench(nbands=160)
    ham = make_ham(dz, ctx, w)
    dt, cdt = (np.float64, np.complex128) if prec == 0 else (np.float32, np.complex64)
    x = torch.from_numpy(w.psi.astype(cdt).view(dt).copy()).cuda()
    y0, y1 = torch.empty_like(x), torch.empty_like(x)
    stream = torch.cuda.Stream()
    ham.set_stream(stream.cuda_stream)
    try:
        for _ in range(3):
            ham.apply_dev(x.data_ptr(), y0.data_ptr(), 160, prec)
        steps = 10
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            ham.apply_dev(x.data_ptr(), y0.data_ptr(), 160, prec)
        e1.record(stream)
        stream.synchronize()
        block_ms = e0.elapsed_time(e1) / steps
        ham.profile(True)
        for _ in range(steps):
            ham.apply_dev(x.data_ptr(), y1.data_ptr(), 160, prec)
        stream.synchronize()
        phases, calls = ham.phase_times()
        ham.profile(False)
        assert torch.equal(y0, y1)
        assert calls == steps and all(v > 0.0 for v in phases.values()), (calls, phases)
        total = sum(phases.values()) / calls
        print(prec, block_ms, total, phases)
        assert abs(total - block_ms) <= 0.1 * block_ms
    finally:
        ham.set_stream(None)
        ham.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_hpsi_every_xline_split(dz, ctx, prec):
    """Every y-slab count the x-line tuner may pick gives the golden H.psi (Si-8, ny = 32)."""
    w = workloads.si8()
    g = np.load(GOLDEN / "hpsi_si8.npz")
    ham = make_ham(dz, ctx, w)
    tol = TOL64 if prec == 0 else TOL32
    errs = {}
    for d in (1, 2, 4):
        ham.set_tuning(d)
        errs[d] = relerr(ham.apply(w.psi, precision=prec), g["y"])
        assert ham.info()["xline_split_fp32" if prec else "xline_split_fp64"] == d
    # -1..-3: the register-resident x-line kernel (one lane per line), -4..-6 the lane-pair kernel
    # (the Si-8 sphere spans 15 of the 32 x values), in 256 / 128 / 64-thread workgroups.
    for v, threads, lanes in ((-1, 256, 1), (-2, 128, 1), (-3, 64, 1), (-4, 256, 2), (-5, 128, 2), (-6, 64, 2)):
        ham.set_tuning(v)
        errs[f"register{threads}x{lanes}"] = relerr(ham.apply(w.psi, precision=prec), g["y"])
        # nx = 32: both register kernels exist in both precisions.
        info = ham.info()
        assert info["xline_register_fp32" if prec else "xline_register_fp64"] == threads
        assert info["xline_lanes_fp32" if prec else "xline_lanes_fp64"] == lanes
    ham.set_tuning(0)
    errs["auto"] = relerr(ham.apply(w.psi, precision=prec), g["y"])
    print(prec, errs, ham.info())
    assert all(e < tol for e in errs.values()), errs


# Which register kernels run per workload and precision: one lane per line for nx <= 100 (FP32)
# or <= 50 (FP64); the lane-pair kernel for even nx when the sphere spans at most half the line
# (tiny: 5 of 8 x values, no; nx = 75: odd, no); the lane-triple kernel for nx divisible by 3
# whose third is 2^a 3^b 5^c (72, 75, 90), any sphere width; the residue-sequential kernel for the
# same lines when the sphere spans at most two thirds of the line (35 planes at 72, 75 and 90).
REGISTER_KERNELS = {"tiny": {"one": (0, 1), "pair": (), "tri": (), "seq": ()},
                    "si8": {"one": (0, 1), "pair": (0, 1), "tri": (), "seq": ()},
                    "bench": {"one": (1,), "pair": (0, 1), "tri": (0, 1), "seq": (0, 1)},
                    "bench75": {"one": (1,), "pair": (), "tri": (0, 1), "seq": (0, 1)},
                    "bench80": {"one": (1,), "pair": (0, 1), "tri": (), "seq": ()},
                    "bench90": {"one": (1,), "pair": (0, 1), "tri": (0, 1), "seq": (0, 1)},
                    "bench100": {"one": (1,), "pair": (0, 1), "tri": (), "seq": ()}}
KIND = {"one": 1, "pair": 2, "tri": 3, "seq": 4, "conv": 5}


@pytest.mark.parametrize("which,prec", [("tiny", 1), ("si8", 1), ("bench", 1), ("bench75", 1), ("bench90", 1),
                                        ("bench80", 1), ("bench100", 1), ("tiny", 0), ("si8", 0), ("bench", 0),
                                        ("bench75", 0), ("bench90", 0), ("bench100", 0)])
def test_hpsi_register_xline_vs_oracle(dz, ctx, which, prec):
    """The register-resident x-line kernels (one x-line per thread, or per lane pair or lane triple
    with half- or third-length transforms; Stockham in VGPRs) on grids of different factorisations
    (nx = 8, 32, 72, 75, 80, 90 and 100, the largest built, on a 100 x 72 x 72 grid; radices
    4/2/3/5/9) against the FP64 oracle apply: FP32 tolerance, and 1e-12 in FP64.  A variant that
    cannot run is refused (both precisions) or runs the LDS kernel (one precision only)."""
    make = {"tiny": workloads.tiny, "si8": workloads.si8, "bench": workloads.bench,
            "bench75": lambda: workloads.bench(nbands=8, grid=(75, 75, 75)),
            "bench80": lambda: workloads.bench(nbands=8, grid=(80, 72, 75)),
            "bench90": lambda: workloads.bench(nbands=8, grid=(90, 90, 90)),
            "bench100": lambda: workloads.bench(nbands=8, grid=(100, 72, 72))}
    w = make[which]()
    ham = make_ham(dz, ctx, w)
    key = "fp32" if prec else "fp64"
    try:
        nb = min(w.nbands, 8)
        ref = hamiltonian.apply_hamiltonian_batched(oracle_ctx(w), w.psi[:nb])
        ran = 0
        for variant, kind, lanes in ((-1, "one", 1), (-4, "pair", 2), (-7, "tri", 3), (-10, "seq", 1)):
            runs = prec in REGISTER_KERNELS[which][kind]
            if not REGISTER_KERNELS[which][kind] or (kind == "one" and 1 not in REGISTER_KERNELS[which][kind]):
                with pytest.raises(dz.BadArgument):
                    ham.set_tuning(variant)
                continue
            ham.set_tuning(variant)
            got = ham.apply(w.psi[:nb], precision=prec)
            info = ham.info()
            assert info[f"xline_register_{key}"] == (256 if runs else 0)
            assert info[f"xline_lanes_{key}"] == (lanes if runs else 0)
            assert info[f"xline_kind_{key}"] == (KIND[kind] if runs else 0)
            err = relerr(got, ref)
            print(which, prec, w.grid, variant, runs, err)
            assert err < (TOL32 if prec else TOL64)
            ran += runs
        assert ran >= 1
    finally:
        ham.close()


@pytest.mark.parametrize("which,nbands", [("tiny", 4), ("si8", 8), ("bench", 20), ("bench75", 20), ("bench90", 8),
                                          ("bench100", 8)])
@pytest.mark.parametrize("prec", [0, 1])
def test_hpsi_conv_xline_vs_oracle(dz, ctx, which, nbands, prec):
    """The convolution x-line kernel (tuning code -13: the x-part of H on each line's present
    x-frequencies as a Toeplitz matrix over all bands, MFMA products, 16-band x 8- or 16-line
    workgroup tiles) against the FP64 oracle apply on grids with x-lines of 8 to 100 points, band
    counts that leave a partial 16-band tile (20) or a single partial one (4, 8), and line counts
    that leave a partial line tile (75 x 75, 90 x 90)."""
    make = {"tiny": lambda: workloads.tiny(), "si8": lambda: workloads.si8(nbands=nbands),
            "bench": lambda: workloads.bench(nbands=nbands),
            "bench75": lambda: workloads.bench(nbands=nbands, grid=(75, 75, 75)),
            "bench90": lambda: workloads.bench(nbands=nbands, grid=(90, 90, 90)),
            "bench100": lambda: workloads.bench(nbands=nbands, grid=(100, 72, 72))}
    w = make[which]()
    ham = make_ham(dz, ctx, w)
    key = "fp32" if prec else "fp64"
    try:
        nb = min(w.nbands, nbands)
        ref = hamiltonian.apply_hamiltonian_batched(oracle_ctx(w), w.psi[:nb])
        ham.set_tuning(-13)
        got = ham.apply(w.psi[:nb], precision=prec)
        info = ham.info()
        assert info[f"xline_kind_{key}"] == 5 and info[f"xline_lanes_{key}"] == 0
        err = relerr(got, ref)
        print(which, prec, w.grid, nb, err)
        assert err < (TOL32 if prec else TOL64)
    finally:
        ham.close()


@pytest.mark.parametrize("grid,nbands", [((72, 72, 72), 12), ((75, 75, 75), 12), ((90, 90, 90), 6),
                                         ((75, 64, 90), 6), ((100, 72, 45), 6), ((40, 120, 48), 4)])
@pytest.mark.parametrize("prec", [0, 1])
def test_hpsi_zreg_plane_vs_oracle(dz, ctx, grid, nbands, prec):
    """The z-register plane kernels (plane kernel 2: the y-transforms on the plane's active
    z-frequency rows in LDS, the z-transform of every y-line in registers on three lanes per line)
    against the FP64 oracle apply, and against the full-tile kernels (plane kernel 1) on the same
    context: z-lines of 45 to 90 points (thirds of 15 to 30), ny from 64 to 120 (four to six
    wavefronts of 21 lines, the last partly idle), non-cubic grids, and the auto choice."""
    w = workloads.bench(nbands=nbands, grid=grid)
    ham = make_ham(dz, ctx, w)
    key = "fp32" if prec else "fp64"
    tol = TOL32 if prec else TOL64
    try:
        ref = hamiltonian.apply_hamiltonian_batched(oracle_ctx(w), w.psi)
        ham.set_plane_kernel(2)
        got = ham.apply(w.psi, precision=prec)
        assert ham.info()[f"plane_kind_{key}"] == 2
        ham.set_plane_kernel(1)
        tile = ham.apply(w.psi, precision=prec)
        assert ham.info()[f"plane_kind_{key}"] == 1
        ham.set_plane_kernel(0)
        auto = ham.apply(w.psi, precision=prec)
        errs = (relerr(got, ref), relerr(tile, ref), relerr(auto, ref), relerr(got, tile))
        print(grid, prec, errs, ham.info()[f"plane_kind_{key}"])
        assert max(errs) < tol, errs
    finally:
        ham.close()
    # No projectors: plane_fwd writes y instead of adding to the nonlocal result.
    ham2 = dz.Hamiltonian(ctx, w.grid, tables.fft_indices(w.basis, w.grid), w.basis.kinetic,
                          w.basis.cart, None, w.inv_volume)
    try:
        ham2.set_local(w.local_r)
        ham2.set_plane_kernel(2)
        ctx_o = oracle_ctx(w)
        ctx_o.species = []
        ref_l = hamiltonian.apply_hamiltonian_batched(ctx_o, w.psi[:2])
        assert relerr(ham2.apply(w.psi[:2], precision=prec), ref_l) < tol
    finally:
        ham2.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_hpsi_zreg_plane_scattered_basis(dz, ctx, prec):
    """A basis whose x-planes are not runs of the basis order (a seeded permutation of the 75^3
    bench sphere) takes the indexed gathers of the z-register plane kernels; the h-major order
    takes the direct reads.  Both against the oracle on the same permuted problem."""
    import dataclasses

    w = workloads.bench(nbands=6, grid=(75, 75, 75))
    perm = np.random.default_rng(5).permutation(w.npw)
    b = w.basis
    wp = dataclasses.replace(w, basis=tables.Basis(b.h[perm], b.k[perm], b.l[perm], b.cart[perm],
                                                   b.kpg[perm], b.kinetic[perm]),
                             psi=np.ascontiguousarray(w.psi[:, perm]))
    tol = TOL32 if prec else TOL64
    for wk in (w, wp):
        ham = make_ham(dz, ctx, wk)
        try:
            ham.set_plane_kernel(2)
            got = ham.apply(wk.psi, precision=prec)
            ref = hamiltonian.apply_hamiltonian_batched(oracle_ctx(wk), wk.psi)
            assert relerr(got, ref) < tol
        finally:
            ham.close()


def test_plane_kernel_choice_errors(dz, ctx):
    """Plane kernel 2 needs nz divisible by 3 with a register-FFT third (8^3: no); 1 and 0 always
    run; other codes are rejected."""
    w = workloads.tiny()
    ham = make_ham(dz, ctx, w)
    try:
        with pytest.raises(dz.BadArgument):
            ham.set_plane_kernel(2)
        with pytest.raises(dz.BadArgument):
            ham.set_plane_kernel(3)
        ham.set_plane_kernel(0)
        g = ham.apply(w.psi)
        assert ham.info()["plane_kind_fp64"] == 1
        ref = hamiltonian.apply_hamiltonian_batched(oracle_ctx(w), w.psi)
        assert relerr(g, ref) < TOL64
    finally:
        ham.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_hpsi_tri_xline_workgroup_sizes(dz, ctx, prec):
    """The lane-triple x-line kernel (21 lines per wavefront, its last lane idle) in 256-, 128- and
    64-thread workgroups on the 75^3 grid (odd line: 3 x 25-point transforms; the 30 Ry sphere
    spans 35 of the 75 x values, so each lane folds rows from two thirds of the line), 12 bands:
    the last workgroup holds a partial set of lines at every size."""
    w = workloads.bench(nbands=12, grid=(75, 75, 75))
    ham = make_ham(dz, ctx, w)
    key = "fp32" if prec else "fp64"
    try:
        ref = hamiltonian.apply_hamiltonian_batched(oracle_ctx(w), w.psi)
        for variant, threads in ((-7, 256), (-8, 128), (-9, 64)):
            ham.set_tuning(variant)
            got = ham.apply(w.psi, precision=prec)
            info = ham.info()
            assert info[f"xline_register_{key}"] == threads and info[f"xline_lanes_{key}"] == 3
            err = relerr(got, ref)
            print(prec, variant, err)
            assert err < (TOL32 if prec else TOL64)
    finally:
        ham.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_hpsi_seq3_xline_workgroup_sizes_and_wide_sphere(dz, ctx, prec):
    """The residue-sequential x-line kernel (one thread per line, the three residues x mod 3 one
    after another, rows k + 25 accumulated in LDS) in 256-, 128- and 64-thread workgroups on the
    75^3 grid, 12 bands; and a sphere wider than two thirds of the line (nx = 48: the 15 Ha
    sphere spans 35 > 32 x values) where it is refused."""
    w = workloads.bench(nbands=12, grid=(75, 75, 75))
    ham = make_ham(dz, ctx, w)
    key = "fp32" if prec else "fp64"
    try:
        ref = hamiltonian.apply_hamiltonian_batched(oracle_ctx(w), w.psi)
        for variant, threads in ((-10, 256), (-11, 128), (-12, 64)):
            ham.set_tuning(variant)
            got = ham.apply(w.psi, precision=prec)
            info = ham.info()
            assert info[f"xline_register_{key}"] == threads and info[f"xline_kind_{key}"] == 4
            err = relerr(got, ref)
            print(prec, variant, err)
            assert err < (TOL32 if prec else TOL64)
    finally:
        ham.close()
    wide = workloads.bench(nbands=4, grid=(48, 72, 72))
    hw = make_ham(dz, ctx, wide)
    try:
        assert hw.info()["planes"] > 32
        with pytest.raises(dz.BadArgument):
            hw.set_tuning(-10)
    finally:
        hw.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_hpsi_overlap_is_bitwise_identical(dz, ctx, prec):
    """The overlapped apply (nonlocal kernels on the context's second stream, joined before
    plane_fwd) gives bit for bit the one-stream result, repeatedly and over two column passes
    (300 bands: 256 + 44)."""
    w = workloads.bench(nbands=300)
    ham = make_ham(dz, ctx, w)
    try:
        ham.set_overlap(False)
        ref = ham.apply(w.psi, precision=prec)
        ham.set_overlap(True)
        for _ in range(3):
            got = ham.apply(w.psi, precision=prec)
            assert np.array_equal(got, ref)
    finally:
        ham.close()
