"""Batched multi-k apply (dzhip_hamset_*): all bands of all k-points of the uneven Al-4 mesh
(4-atom fcc cubic cell, 30 Ry, 30^3 grid, 8x8x8 reduced by symmetry to 35 k-points with 1,208 to
1,260 plane waves, 12 bands) in one launch sequence, against 35 per-k applies
(apply_hamiltonian_batched per k-point, src/features/scf/apply.zig:1584-1618, over the k-point
queue of src/features/scf/kpoint/workers.zig:133-180).  With every member on the same x-line
kernel and plane kernels (full-tile or z-register) the set's result equals the per-k results bit for
bit, in both precisions, eagerly and as a captured graph replayed between other work; the per-k loop
and the set are timed."""
import time

import numpy as np
import pytest

from dzhip import scf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def al35(dz, ctx):
    al = scf.aluminium(symmetry=True, kmesh=(8, 8, 8))
    local_r = scf.start_potential(al)
    ws = [scf.kpoint_workload(al, k, nbands=al.nbands, local_r=local_r) for k in range(len(al.kpoints))]
    hams = [dz.Hamiltonian.from_workload(ctx, w) for w in ws]
    yield al, ws, hams
    for h in hams:
        h.close()


@pytest.mark.parametrize("plane", [1, 2])
@pytest.mark.parametrize("code", [-4, -10, -13])
@pytest.mark.parametrize("prec", [0, 1])
def test_hamset_equals_per_k_applies(dz, ctx, al35, prec, code, plane):
    import torch

    al, ws, hams = al35
    for h in hams:
        h.set_tuning(code)
        h.set_plane_kernel(plane)
    hs = dz.HamSet(hams)
    try:
        info = hs.info()
        assert info["k_points"] == 35 and info["n_pw_total"] == sum(w.npw for w in ws)
        ref = [h.apply(w.psi, prec) for h, w in zip(hams, ws)]
        got = hs.apply([w.psi for w in ws], prec)
        for k in range(35):
            assert np.array_equal(got[k], ref[k]), (k, float(np.max(np.abs(got[k] - ref[k]))))
        # Device layout, eager and graph: replays between other work equal the eager result.
        cdt, rdt = (np.complex128, np.float64) if prec == 0 else (np.complex64, np.float32)
        xpk = hs.pack([w.psi.astype(cdt) for w in ws])
        x = torch.from_numpy(xpk.view(rdt).copy()).cuda()
        y_eager = torch.empty_like(x)
        y_graph = torch.empty_like(x)
        stream = torch.cuda.Stream()
        hs.set_stream(stream.cuda_stream)
        hs.apply_dev(x.data_ptr(), y_eager.data_ptr(), al.nbands, prec, use_graph=False)
        stream.synchronize()
        eager = y_eager.cpu().numpy().view(cdt)
        for k, blk in enumerate(hs.unpack(eager)):
            if prec == 0:
                assert np.array_equal(blk, ref[k])
            else:
                assert np.array_equal(blk.astype(np.complex128), ref[k])
        for rep in range(3):
            with torch.cuda.stream(stream):
                y_graph.zero_()
            hs.apply_dev(x.data_ptr(), y_graph.data_ptr(), al.nbands, prec, use_graph=True)
            # Other work on the same stream between replays.
            hs.apply_dev(x.data_ptr(), y_eager.data_ptr(), al.nbands, prec, use_graph=False)
            stream.synchronize()
            assert torch.equal(y_graph, y_eager), rep
        assert hs.info()["graphs"] == 1
        hs.set_stream(None)
    finally:
        hs.close()


def test_hamset_timing_vs_per_k_loop(dz, ctx, al35):
    """Time of one apply of all 35 k-points x 12 bands: per-k loop of device applies, the set's
    launch sequence, and its graph (printed; the set must not be slower than the loop)."""
    import torch

    al, ws, hams = al35
    for h in hams:
        h.set_tuning(0)
        h.set_plane_kernel(0)
    hs = dz.HamSet(hams)
    stream = torch.cuda.Stream()
    try:
        xs = [torch.from_numpy(w.psi.view(np.float64).copy()).cuda() for w in ws]
        ys = [torch.empty_like(x) for x in xs]
        for h in hams:
            h.set_stream(stream.cuda_stream)
        hs.set_stream(stream.cuda_stream)
        xpk = torch.from_numpy(hs.pack([w.psi for w in ws]).view(np.float64).copy()).cuda()
        ypk = torch.empty_like(xpk)
        res = {}
        for prec in (0, 1):
            if prec == 1:
                xs = [x.view(torch.complex128).to(torch.complex64).view(torch.float32) for x in xs]
                ys = [torch.empty_like(x) for x in xs]
                xpk = xpk.view(torch.complex128).to(torch.complex64).view(torch.float32)
                ypk = torch.empty_like(xpk)
                torch.cuda.synchronize()
            runs = {"per_k_loop": lambda: [h.apply_dev(x.data_ptr(), y.data_ptr(), al.nbands, prec)
                                           for h, x, y in zip(hams, xs, ys)],
                    "set": lambda: hs.apply_dev(xpk.data_ptr(), ypk.data_ptr(), al.nbands, prec, False),
                    "set_graph": lambda: hs.apply_dev(xpk.data_ptr(), ypk.data_ptr(), al.nbands, prec, True)}
            for name, fn in runs.items():
                for _ in range(3):
                    fn()
                stream.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                t0 = time.perf_counter()
                e0.record(stream)
                for _ in range(20):
                    fn()
                e1.record(stream)
                stream.synchronize()
                res[(prec, name)] = (e0.elapsed_time(e1) / 20, (time.perf_counter() - t0) / 20 * 1e3)
        for key, (gpu_ms, wall_ms) in sorted(res.items()):
            print(f"al4 35k x 12 bands precision {key[0]} {key[1]}: {gpu_ms:.3f} ms stream, {wall_ms:.3f} ms host wall")
        for prec in (0, 1):
            assert res[(prec, "set_graph")][0] <= res[(prec, "per_k_loop")][0]
    finally:
        for h in hams:
            h.set_stream(None)
        hs.set_stream(None)
        hs.close()
