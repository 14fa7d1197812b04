#!/usr/bin/env python3
"""H.psi block-apply and SCF benchmark on one MI355X, on the Si-64 workload.

Headline (`value`): H.psi throughput of one k-point of the Si-64 SCF system in FP64 (the
reference's precision): 2x2x2 conventional Si cells (64 atoms displaced by N(0, 0.02 A)), 30 Ry,
75^3 grid, k = (1/2, 1/2, 1/2) of the 2x2x2 mesh (N_pw = 23,984, 36 x-planes), 256 KB projectors,
160 bands, local_r = V_ion + V_Hxc of the atomic start density, psi random (seeded) with
coefficients damped by 1/(1 + |k+G|^2).  One step = one apply of H to the 160-band block with psi
resident in HBM; the FP32 apply is timed the same way and reported beside it, with its error.

value = algorithmic H.psi bytes per block / time per block (GB/s), where
    B = nb*s_c*(2*N_pw + 2*N) + s_r*N + 4*N_pw + s_r*4*N_pw
with s_c/s_r = 16/8 (FP64) or 8/4 (FP32): psi read + H psi written, one psi(r) round trip per band,
V once, the int32 index map and the projector table (a lower bound for any design that
materialises psi(r) once per band).

Also reported: the isolated H.psi bench (L = 20.5 bohr cube, 15 Ha, 72^3, N_pw = 23,847, 160 bands,
64 Gaussian s/p atoms) in FP32 and FP64; the CPU baseline (the oracle's FP64 H.psi with
min(n_k, cores) k-threads over the 8 Si-64 k-points, an all-cores single-k variant, and the dense
algebra of one Davidson iteration); and (`--scf si64`, the default) the SCF time to 1e-6 Ha of
Si-64 for plain FP64 (k-points in sequence), FP64 with concurrent k-points, mixed precision with
concurrent k-points, and the same with the learned preconditioner.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--scf si64|si8|none]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "dft-zig_amd"))

# Hardware queues of the HIP runtime (read once, when HIP initialises).  The concurrent-k SCF runs
# eight k-point streams; the runtime's default of four queues puts several streams on one queue,
# whose packets execute in order, so one stream's eigensolver holds back another's H applies.
try:
    _queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "0"))
except ValueError:
    _queues = 0
if _queues < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

METRIC = "SCF time-to-1e-6 Ha & H·ψ GB/s, Si-64: mixed+batched+learned vs plain FP64"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}
TRAFFIC_FILE = ROOT / "profiles" / "traffic_si64.txt"
SI64_K = 7  # k = (1/2, 1/2, 1/2): the largest sphere of the Si-64 mesh


def hpsi_bytes(nb, npw, ngrid, prec):
    sc, sr = (8, 4) if prec == "fp32" else (16, 8)
    return nb * sc * (2 * npw + 2 * ngrid) + sr * ngrid + 4 * npw + sr * 4 * npw


def kernel_model(phase, nb, npw, ngrid, planes, plane_elems, nproj, prec):
    """(bound, algorithmic amount per launch, unit) for each phase of the apply."""
    sc, sr = (8, 4) if prec == "fp32" else (16, 8)
    half = planes * plane_elems
    if phase == "plane_inv":
        return "hbm", nb * (npw + half) * sc, "B"
    if phase == "xline":
        return "hbm", nb * 2 * half * sc + ngrid * sr, "B"
    if phase == "plane_fwd":
        return "hbm", nb * (half + 2 * npw) * sc + npw * sr, "B"
    # complex GEMM, structure factor folded into the projector matrix: 8 flops per complex MAC
    return "mfma", 8.0 * nproj * npw * nb, "flop"


def read_traffic():
    """HBM bytes per launch of each phase from the committed PMC summary, if present
    (keys: phase name, with a _fp64 suffix for the FP64 apply)."""
    if not TRAFFIC_FILE.exists():
        return {}
    out = {}
    for line in TRAFFIC_FILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch"):
            _, kern, val = line.split()[:3]
            out[kern] = float(val)
    return out


def time_gpu(ham, x, y, ncols, prec_code, steps, warmup, torch):
    """(ms per block apply as run in production, per-phase ms of a profiled pass).

    The block is timed as the library runs it (no per-phase events); a second, profiled pass of
    the same steps times each phase with HIP events on the same stream, which is what the
    roofline of a kernel needs."""
    # An explicit (non-default) stream: the library launches on it and torch's events bracket
    # exactly those launches.  (The default stream's handle is 0, which the library reads as
    # "use the context's own stream".)  The overlapped pass forks from and joins back into it.
    stream = torch.cuda.Stream()
    ham.set_stream(stream.cuda_stream)
    for _ in range(warmup):
        ham.apply_dev(x.data_ptr(), y.data_ptr(), ncols, prec_code)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        ham.apply_dev(x.data_ptr(), y.data_ptr(), ncols, prec_code)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / steps
    ham.profile(True)
    for _ in range(steps):
        ham.apply_dev(x.data_ptr(), y.data_ptr(), ncols, prec_code)
    torch.cuda.synchronize()
    phases, calls = ham.phase_times()
    ham.profile(False)
    ham.set_stream(None)
    return ms, {k: v / max(calls, 1) for k, v in phases.items()}


def cpu_baseline(system, bands_per_k=16, cores_cap=16):
    """The oracle's FP64 H.psi on the host (oracle/cpu_baseline.py): min(n_k, cores) k-threads over
    the Si-64 k-points, `bands_per_k` bands of each, one FFT/BLAS thread per k-thread (the
    reference's k-point worker scheme); an all-cores variant on one k-point; and the dense algebra
    of one Davidson iteration at the full subspace (m = 648) with every core."""
    from dzhip import scf as dscf
    from oracle import adapters, cpu_baseline as cb

    host = cb.host_description()
    cores = max(1, min(host["cores_available"], cores_cap))
    nk = len(system.kpoints)
    kthreads = min(nk, cores)
    ctxs, blocks, nbytes = [], [], 0.0
    ngrid = int(np.prod(system.grid))
    local_r = dscf.start_potential(system)
    for k in range(nk):
        wk = dscf.kpoint_workload(system, k, nbands=bands_per_k, local_r=local_r)
        ctxs.append(adapters.context_from_workload(wk))
        blocks.append(wk.psi)
        nbytes += hpsi_bytes(bands_per_k, wk.npw, ngrid, "fp64")
    r = cb.kthread_hpsi(ctxs, blocks, kthreads)
    gbs = nbytes / r["seconds"] / 1e9
    one = cb.allcores_hpsi(ctxs[SI64_K], blocks[SI64_K], cores)
    gbs_all = hpsi_bytes(one["bands"], ctxs[SI64_K].npw, ngrid, "fp64") / one["seconds"] / 1e9
    npw = ctxs[SI64_K].npw
    dav = cb.davidson_iteration_algebra(npw, 4 * system.nbands + 8, system.nbands + 4, cores)
    per_band = r["seconds"] * kthreads / r["bands"]
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": kthreads, "kind": "port",
            "sample": f"oracle FP64 H.psi (scipy pocketfft + numpy BLAS, one thread per k-thread) on "
                      f"{bands_per_k} of 160 bands at each of the {nk} Si-64 k-points, {kthreads} k-threads "
                      f"(min(n_k, cores)), {r['seconds']:.2f} s wall; about {per_band * 160:.1f} s per 160-band "
                      f"block per k-thread",
            "host": host,
            "all_cores_single_k": {"value": round(gbs_all, 3), "unit": "GB/s", "cores": cores,
                                   "sample": f"{one['bands']} bands of k = {SI64_K}, FFTs and BLAS on {cores} "
                                             f"threads, {one['seconds']:.2f} s"},
            "davidson_iteration_algebra_s": {k: round(v, 4) for k, v in dav.items()},
            "davidson_iteration_note": f"m = {4 * system.nbands + 8} subspace, {system.nbands + 4} new columns, "
                                       f"N_pw = {npw}, {cores} threads: V^H W, zheev (numpy eigh), V c and W c; "
                                       f"add the H.psi of {system.nbands + 4} bands for a full iteration"}


# SCF settings shared by every variant (only precision, k-point concurrency and the
# preconditioner differ): Pulay history 8 from the first iteration (the reference's pulay_start
# setting; from iteration 4, its default, the start density sloshes for some ten iterations),
# beta 0.3, model-dielectric preconditioner diemac 12 (a Si setting of the reference's config),
# atomic start density, diagonal start vectors, eigensolver tolerance min(1e-4, drho / 10) >=
# eig_tol_min on the occupied bands, converged when the RMS density change is below conv = 1e-7
# (ten times tighter than the reference's default 1e-6, config.zig:137; the total energy is then
# settled to about 1e-9 Ry, far inside the 1e-6 Ha target).
SCF_COMMON = dict(conv=1e-7, pulay_start=1, diemac=12.0, start="atomic", init_diagonal=True, eig_tol_ratio=0.1,
                  eig_tol_min=1e-9)


def scf_leg(name, max_iter, arms, log=sys.stderr):
    """Time-to-1e-6 Ha of the SCF for the forms in `arms`: plain FP64 ("fp64": k-points one after
    another on one stream), FP64 with every k-point on its own stream and host thread
    ("fp64_concurrent"), mixed precision with concurrent k-points ("mixed_batched": FP32 H apply,
    FP64 subspace algebra, FP64 apply below a 1e-5 residual) and the same with the learned
    preconditioner on the kinetic-shift base ("mixed_batched_learned": trained first, on the
    device, on residual/correction pairs of the warm-started solves of a 12-step FP64 SCF of the
    8-atom Si cell at the same cutoff; collection and training time reported, not counted).
    E_inf is the plain FP64 run's final energy when that run converged (RMS density change below
    conv); otherwise no time-to-1e-6 Ha or speedup is reported."""
    import dzhip
    from dzhip import precond
    from dzhip import scf as dscf

    system = {"si64": dscf.si64, "si8": dscf.si8_small, "gaas64": dscf.gaas64,
              "al4": lambda: dscf.aluminium(symmetry=True, kmesh=(8, 8, 8), xc="pbe")}[name]()
    nocc = int(np.ceil(system.nelec / 2)) if system.smearing == 0.0 else 0
    settings = dict(SCF_COMMON)
    if name == "al4":
        # The perfect fcc cell: LCG start vectors (diagonal ones can miss symmetry-forbidden states),
        # every band converged (Fermi-Dirac occupations), the metal's own dielectric constant.
        settings.update(init_diagonal=False, diemac=1.0)
    out = {"system": system.name, "grid": list(system.grid), "k_points": len(system.kpoints),
           "bands": system.nbands, "settings": settings, "max_iter": max_iter,
           "gpu_max_hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "0"))}
    params, learned_base = None, dzhip.PRECOND_KINETIC
    if any(a.endswith("learned") for a in arms):
        t0 = time.perf_counter()
        tctx = dzhip.Context(0)
        if name == "gaas64":
            # Residual/correction pairs of Si-8 and Al-4 runs at the cutoff of the run (so the
            # features cover its |k+G|^2 range), on the TPA base the GaAs comparison measures against.
            learned_base = dzhip.PRECOND_TPA
            train_systems = [dscf.silicon(ecut_ry=system.ecut_ry, nbands=24,
                                          kpoints=([np.zeros(3), np.array([0.11, 0.23, 0.05])], [0.5, 0.5])),
                             dscf.aluminium(ecut_ry=30.0, kmesh=(2, 2, 2), nbands=10,
                                            displace=dscf.AL_SMALL_DISPLACEMENT)]
        else:
            # Training system at the cutoff of the run, so the features cover its |k+G|^2 range.
            train_systems = [dscf.silicon(ecut_ry=system.ecut_ry, nbands=24,
                                          kpoints=([np.zeros(3), np.array([0.11, 0.23, 0.05])], [0.5, 0.5]))]
        params, hist, nsamp = precond.train_on_scf(tctx, train_systems, base=learned_base, steps=12,
                                                   diemac=12.0, start="atomic", init_diagonal=True)
        tctx.close()
        train_s = time.perf_counter() - t0
        print(f"learned preconditioner: {nsamp} samples, loss {hist[0]:.4f} -> {hist[-1]:.4f}, {train_s:.1f} s",
              file=log, flush=True)
        out["learned_training"] = {"samples": nsamp,
                                   "base": "TPA" if learned_base == dzhip.PRECOND_TPA else "kinetic shift",
                                   "systems": [t.name for t in train_systems], "loss_start": float(hist[0]),
                                   "loss_end": float(hist[-1]), "seconds": round(train_s, 2)}
    forms = {"fp64": (dzhip.FP64, False, dzhip.PRECOND_KINETIC),
             "fp64_concurrent": (dzhip.FP64, True, dzhip.PRECOND_KINETIC),
             "mixed_batched": (dzhip.FP32, True, dzhip.PRECOND_KINETIC),
             "mixed_batched_learned": (dzhip.FP32, True, dzhip.PRECOND_MLP),
             # GaAs-64 (Gamma only): the TPA preconditioner against the network on the TPA base.
             "tpa": (dzhip.FP32, False, dzhip.PRECOND_TPA),
             "tpa_learned": (dzhip.FP32, False, dzhip.PRECOND_MLP)}
    runs, setup, retries = {}, {}, {}
    for tag in arms:
        prec, conc, pc = forms[tag]
        drv = dscf.ScfDriver(system, concurrent_k=conc, device_grid=True)
        setup[tag] = drv.setup_s
        fallbacks = []
        if pc == dzhip.PRECOND_MLP:
            drv.set_precond_model(params, precond.ENERGY_SCALE, learned_base)
            drv.solve_hook = lambda k, ham, st: fallbacks.append(st["precond_fallback"])

        def show(r, tag=tag):
            sec = " ".join(f"{k[2:-3]} {v:.0f}" for k, v in r.sections_ms.items())
            print(f"scf {name} {tag} it {r.iteration} E {r.energy_ry:.10f} drho {r.drho:.3e} t {r.wall_s:.2f}s "
                  f"eig {r.eig_s:.2f}s eig_it {r.eig_iterations} applies {r.applies} fp32 {r.fp32_applies} "
                  f"gpu_ms[{sec}]", file=log, flush=True)
        retries0 = dzhip.heev_retries()
        try:
            runs[tag] = drv.run(max_iter=max_iter, precision=prec, preconditioner=pc, nconv=nocc, progress=show,
                                **settings)
        finally:
            drv.close()
        retries[tag] = dzhip.heev_retries() - retries0
        if pc == dzhip.PRECOND_MLP:
            out["learned_training"]["scf_solves_fallback"] = {"solves": len(fallbacks),
                                                              "fell_back": sum(1 for f in fallbacks if f > 0)}
    base_arm = "fp64" if "fp64" in runs else arms[0]
    ref = runs.get(base_arm)
    e_inf = ref.energies[-1] if ref is not None and ref.converged else None
    out["e_inf_ry"] = e_inf
    out["e_inf_source"] = (f"{base_arm} GPU run, converged (RMS density change below conv)" if e_inf is not None
                           else f"none: the {base_arm} run did not converge, so no time-to-1e-6 Ha is reported")
    for tag, run in runs.items():
        t = run.time_to(e_inf) if e_inf is not None else None
        out[tag] = {"t_1e-6_Ha_s": None if t is None else round(t, 3), "iterations": len(run.records),
                    "converged": run.converged,
                    "final_dE_Ha": None if e_inf is None else (run.energies[-1] - e_inf) * dscf.HA_PER_RY,
                    "final_drho": run.records[-1].drho,
                    "wall_s": round(run.records[-1].wall_s, 3), "setup_s": round(setup[tag], 3),
                    "eig_s": round(sum(r.eig_s for r in run.records), 3),
                    "davidson_iterations": sum(r.eig_iterations for r in run.records),
                    "applies": sum(r.applies for r in run.records),
                    "solver_gpu_ms": {k[2:-3]: round(sum(r.sections_ms.get(k, 0.0) for r in run.records), 1)
                                      for k in dscf.SECTIONS},
                    "fp32_applies": sum(r.fp32_applies for r in run.records),
                    # Rayleigh-Ritz tridiagonalisations whose ordinary launch gave up waiting for its
                    # workgroups beside other k-points' work and was repeated cooperatively.
                    "heev_cooperative_repeats": retries[tag]}
    a = out.get(base_arm, {}).get("t_1e-6_Ha_s")
    for tag in arms:
        if tag == base_arm:
            continue
        b = out[tag]["t_1e-6_Ha_s"]
        out[f"speedup_{tag}"] = None if not a or not b else round(a / b, 3)
    return out


def heev_profile(dzhip, torch, ctx, sizes=(164, 328, 492, 648), nev=160, calls=10):
    """The library's dense eigensolver (dzhip_heev_dev) against torch.linalg.eigh (rocSOLVER) on
    random Hermitian matrices with a prescribed spectrum (eigenvalues uniform in [-2, 30], random
    unitary eigenvectors, seeded) at the Rayleigh-Ritz sizes of a 160-band Davidson solve; one line
    per size with both times (HIP events, 3 warm-up calls) and the eigensolver's largest eigenvalue
    difference to numpy's eigvalsh and largest |Z^H Z - I|."""
    st = torch.cuda.Stream()
    rng = np.random.default_rng(0)
    for m in sizes:
        q, _ = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
        lam = np.sort(rng.uniform(-2.0, 30.0, m))
        a = (q * lam) @ q.conj().T
        a = 0.5 * (a + a.conj().T)
        at = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
        w = torch.zeros(nev, dtype=torch.float64, device="cuda")
        z = torch.zeros((nev, m), dtype=torch.complex128, device="cuda")
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for _ in range(3):
            ctx.heev_dev(at.data_ptr(), m, m, nev, w.data_ptr(), z.data_ptr(), m, st.cuda_stream)
        ctx.heev_check(st.cuda_stream)
        e0.record(st)
        for _ in range(calls):
            ctx.heev_dev(at.data_ptr(), m, m, nev, w.data_ptr(), z.data_ptr(), m, st.cuda_stream)
        e1.record(st)
        ctx.heev_check(st.cuda_stream)
        ms = e0.elapsed_time(e1) / calls
        err = float(np.max(np.abs(w.cpu().numpy() - np.linalg.eigvalsh(a)[:nev])))
        zz = z.cpu().numpy().T
        orth = float(np.max(np.abs(zz.conj().T @ zz - np.eye(nev))))
        with torch.cuda.stream(st):
            for _ in range(2):
                torch.linalg.eigh(at)
            e0.record(st)
            for _ in range(calls):
                torch.linalg.eigh(at)
            e1.record(st)
        torch.cuda.synchronize()
        ms_t = e0.elapsed_time(e1) / calls
        print(json.dumps({"heev": {"m": m, "nev": nev, "ms": round(ms, 3), "torch_eigh_ms": round(ms_t, 3),
                                   "max_dlambda": err, "max_orth": orth}}), flush=True)


def eig_profile(ham):
    """Section times of one cold and two warm FP64 Davidson solves (128 bands, 12 iterations,
    diagonal start) on the Si-64 k-point context, one JSON line per solve."""
    for rep in ("cold", "warm", "warm"):
        t0 = time.perf_counter()
        _, _, st = ham.eig_solve(128, tol=1e-4, max_iter=12, init_diagonal=True, want_vectors=False)
        wall = time.perf_counter() - t0
        print(json.dumps({"solve": rep, "wall_ms": round(wall * 1e3, 1), "iterations": st["iterations"],
                          "sections_ms": {k[2:-3]: round(v, 1) for k, v in st.items() if k.startswith("t_")}}),
              flush=True)


def hpsi_block(dzhip, torch, ham, w, steps, warmup):
    """FP64 and FP32 timings of one block apply of workload w (psi resident in HBM), the phases of
    each, and the FP32-vs-FP64 error."""
    nb, npw = w.nbands, w.npw
    ngrid = int(np.prod(w.grid))
    x64 = torch.from_numpy(w.psi.view(np.float64).copy()).cuda()
    y64 = torch.empty_like(x64)
    x32 = torch.from_numpy(w.psi.astype(np.complex64).view(np.float32).copy()).cuda()
    y32 = torch.empty_like(x32)
    ms64, ph64 = time_gpu(ham, x64, y64, nb, dzhip.FP64, steps, warmup, torch)
    ms32, ph32 = time_gpu(ham, x32, y32, nb, dzhip.FP32, steps, warmup, torch)
    h32 = y32.cpu().numpy().view(np.complex64).reshape(nb, npw).astype(np.complex128)
    h64 = y64.cpu().numpy().view(np.complex128).reshape(nb, npw)
    err = float(np.max(np.linalg.norm(h32 - h64, axis=1) / np.linalg.norm(h64, axis=1)))
    abserr = float(np.max(np.abs(h32 - h64)))
    return {"ms64": ms64, "ph64": ph64, "ms32": ms32, "ph32": ph32,
            "gbs64": hpsi_bytes(nb, npw, ngrid, "fp64") / (ms64 * 1e-3) / 1e9,
            "gbs32": hpsi_bytes(nb, npw, ngrid, "fp32") / (ms32 * 1e-3) / 1e9,
            "err": {"max_band_rel_l2": err, "max_abs": abserr}}


def batched_al4(dzhip, torch, ctx, steps, warmup):
    """Uneven multi-k batching (fcc Al, 4-atom cubic cell, 30 Ry, 30^3 grid, 8x8x8 mesh reduced by
    symmetry to 35 k-points of 1,208-1,260 plane waves, 12 bands): one apply of all k-points x bands
    as a per-k loop of device applies, as the set's launch sequence, and as its captured graph."""
    from dzhip import scf as dscf

    al = dscf.aluminium(symmetry=True, kmesh=(8, 8, 8))
    local_r = dscf.start_potential(al)
    ws = [dscf.kpoint_workload(al, k, nbands=al.nbands, local_r=local_r) for k in range(len(al.kpoints))]
    hams = [dzhip.Hamiltonian.from_workload(ctx, w) for w in ws]
    hs = dzhip.HamSet(hams)
    stream = torch.cuda.Stream()
    ngrid = int(np.prod(al.grid))
    out = {"workload": "fcc Al 4-atom cubic cell, 30 Ry, 30^3 grid, 35 IBZ k-points (8x8x8, 48 ops + time "
                       "reversal), N_pw 1208-1260, 12 bands per k",
           "k_points": len(ws), "n_pw_total": hs.total}
    try:
        for h in hams:
            h.set_stream(stream.cuda_stream)
        hs.set_stream(stream.cuda_stream)
        for prec, tag, cdt, rdt in ((dzhip.FP64, "fp64", np.complex128, np.float64),
                                    (dzhip.FP32, "fp32", np.complex64, np.float32)):
            xs = [torch.from_numpy(w.psi.astype(cdt).view(rdt).copy()).cuda() for w in ws]
            ys = [torch.empty_like(x) for x in xs]
            xpk = torch.from_numpy(hs.pack([w.psi.astype(cdt) for w in ws]).view(rdt).copy()).cuda()
            ypk = torch.empty_like(xpk)
            torch.cuda.synchronize()
            runs = {"per_k_loop": lambda: [h.apply_dev(x.data_ptr(), y.data_ptr(), al.nbands, prec)
                                           for h, x, y in zip(hams, xs, ys)],
                    "set": lambda: hs.apply_dev(xpk.data_ptr(), ypk.data_ptr(), al.nbands, prec, False),
                    "set_graph": lambda: hs.apply_dev(xpk.data_ptr(), ypk.data_ptr(), al.nbands, prec, True)}
            nbytes = sum(hpsi_bytes(al.nbands, w.npw, ngrid, tag) for w in ws)
            res = {}
            for name, fn in runs.items():
                for _ in range(max(warmup, 1)):
                    fn()
                stream.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(steps):
                    fn()
                e1.record(stream)
                stream.synchronize()
                ms = e0.elapsed_time(e1) / steps
                res[name] = {"ms": round(ms, 4), "gbs": round(nbytes / (ms * 1e-3) / 1e9, 1)}
            res["speedup_graph_vs_loop"] = round(res["per_k_loop"]["ms"] / res["set_graph"]["ms"], 2)
            out[tag] = res
        out["graphs_captured"] = hs.info()["graphs"]
    finally:
        for h in hams:
            h.set_stream(None)
        hs.set_stream(None)
        hs.close()
        for h in hams:
            h.close()
    return out


def roofline(w, info, phases, prec):
    """Roofline of the dominant kernel of an apply (per-phase HIP events on the launch stream)."""
    nb, npw = w.nbands, w.npw
    ngrid = int(np.prod(w.grid))
    dom = max(phases, key=phases.get)
    plane_elems = w.grid[1] * w.grid[2]
    bound, amount, _ = kernel_model(dom, nb, npw, ngrid, info["planes"], plane_elems, info["projectors"], prec)
    t_dom = phases[dom] * 1e-3
    traffic = read_traffic().get(dom + ("_fp64" if prec == "fp64" else ""))
    if bound == "hbm":
        achieved = amount / t_dom / 1e9
        return {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes": amount, "avg_ms": round(phases[dom], 5), "precision": prec}
    achieved = amount / t_dom / 1e12
    peak = MFMA_PEAK_TFLOPS[prec]
    return {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 2), "peak": peak,
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic,
            "algorithmic_flops": amount, "avg_ms": round(phases[dom], 5), "precision": prec}


def tuning_info(ham):
    t = ham.info()
    return {k: t[k] for k in ("xline_split_fp32", "xline_split_fp64", "xline_register_fp32", "xline_register_fp64",
                              "xline_lanes_fp32", "xline_lanes_fp64", "xline_kind_fp32", "xline_kind_fp64",
                              "plane_kind_fp32", "plane_kind_fp64")}


SCF_ARMS = ("fp64", "fp64_concurrent", "mixed_batched", "mixed_batched_learned")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--scf", default="si64", choices=("none", "si8", "si64", "gaas64", "al4"),
                    help="also time the SCF to 1e-6 Ha (plain FP64 vs the concurrent and mixed forms)")
    ap.add_argument("--scf-max-iter", type=int, default=70)
    ap.add_argument("--scf-arms", default=",".join(SCF_ARMS),
                    help="comma-separated SCF forms to run (fp64 gives E_inf)")
    ap.add_argument("--no-isolated", action="store_true", help="skip the isolated 72^3 H.psi bench")
    ap.add_argument("--no-batched", action="store_true", help="skip the 35-k Al-4 batched apply")
    ap.add_argument("--no-gaas", action="store_true", help="skip the GaAs-64 H.psi block")
    ap.add_argument("--eig", action="store_true",
                    help="instead of the bench line: one cold and two warm FP64 Davidson solves of 128 bands "
                         "on the Si-64 k-point (12 iterations), with the solver's section times")
    ap.add_argument("--heev", action="store_true",
                    help="instead of the bench line: the dense eigensolver against torch.linalg.eigh at the "
                         "Rayleigh-Ritz sizes m = 164..648 (nev = 160)")
    ap.add_argument("--overlap", action="store_true",
                    help="run the nonlocal kernels on a second stream beside the local ones (timed pass only)")
    ap.add_argument("--plane-kernel", type=int, default=0,
                    help="force the plane kernels of the headline block: 1 full-tile, 2 z-register (0: autotune)")
    ap.add_argument("--xline-split", type=int, default=0,
                    help="fix the x-line variant: d > 0 LDS y-slabs, -1..-3 one-lane, -4..-6 lane-pair, "
                         "-7..-9 lane-triple, -10..-12 residue-sequential register kernel (0 = autotune on the "
                         "first apply)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (k-point sharding is "
              "design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import dzhip
    from dzhip import scf as dscf
    from dzhip import workloads

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.heev:
        hctx = dzhip.Context(0)
        heev_profile(dzhip, torch, hctx)
        hctx.close()
        return
    system = dscf.si64()
    w = dscf.kpoint_workload(system, SI64_K, nbands=system.nbands)
    ctx = dzhip.Context(0)
    ham = dzhip.Hamiltonian.from_workload(ctx, w)
    if args.xline_split:
        ham.set_tuning(args.xline_split)
    if args.plane_kernel:
        ham.set_plane_kernel(args.plane_kernel)
    if args.overlap:
        ham.set_overlap(True)
    if args.eig:
        eig_profile(ham)
        return
    info = ham.info()
    r = hpsi_block(dzhip, torch, ham, w, args.steps, args.warmup)
    tuned = tuning_info(ham)
    ham.close()

    # The SCF leg runs next, before the other H.psi workloads and the CPU baseline: its concurrent
    # arms (eight host threads driving eight streams) ran 0.5-0.8 s slower after those legs.
    arms = [a for a in args.scf_arms.split(",") if a]
    if args.scf == "gaas64" and args.scf_arms == ",".join(SCF_ARMS):
        arms = ["tpa", "tpa_learned"]
    if args.scf == "al4" and args.scf_arms == ",".join(SCF_ARMS):
        arms = ["fp64", "fp64_concurrent", "mixed_batched"]
    scf_result = None if args.scf == "none" else scf_leg(args.scf, args.scf_max_iter, arms)

    isolated = None
    if not args.no_isolated:
        wb = workloads.bench(nbands=160)
        hb = dzhip.Hamiltonian.from_workload(ctx, wb)
        rb = hpsi_block(dzhip, torch, hb, wb, args.steps, args.warmup)
        isolated = {"workload": "L=20.5 bohr cube, ecut 15 Ha, 72^3 grid, N_pw=23847, 160 bands, 64 atoms x "
                                "(s+3p) Gaussian projectors, V_loc = 64 Gaussians (seeded PCG64 20261202)",
                    "fp64": {"value": round(rb["gbs64"], 2), "unit": "GB/s", "ms_per_step": round(rb["ms64"], 4),
                             "frac_hbm": round(rb["gbs64"] / HBM_PEAK_GBS, 4)},
                    "fp32": {"value": round(rb["gbs32"], 2), "unit": "GB/s", "ms_per_step": round(rb["ms32"], 4),
                             "frac_hbm": round(rb["gbs32"] / HBM_PEAK_GBS, 4)},
                    "fp32_vs_fp64_err": rb["err"],
                    "phases_ms_fp64": {k: round(v, 5) for k, v in rb["ph64"].items()},
                    "phases_ms_fp32": {k: round(v, 5) for k, v in rb["ph32"].items()},
                    "tuning": tuning_info(hb)}
        hb.close()

    batched = None if args.no_batched else batched_al4(dzhip, torch, ctx, args.steps, args.warmup)

    gaas = None
    if not args.no_gaas:
        wg = dscf.kpoint_workload(dscf.gaas64(), 0, nbands=160)
        hg = dzhip.Hamiltonian.from_workload(ctx, wg)
        rg = hpsi_block(dzhip, torch, hg, wg, args.steps, args.warmup)
        gaas = {"workload": "GaAs-64 zincblende 2x2x2 supercell (a = 11.306 A, N(0, 0.05 A) displacements), 35 Ry, "
                            "90^3 grid, Gamma, N_pw = 34049, 160 bands, 256 projectors",
                "fp64": {"value": round(rg["gbs64"], 2), "unit": "GB/s", "ms_per_step": round(rg["ms64"], 4),
                         "frac_hbm": round(rg["gbs64"] / HBM_PEAK_GBS, 4)},
                "fp32": {"value": round(rg["gbs32"], 2), "unit": "GB/s", "ms_per_step": round(rg["ms32"], 4),
                         "frac_hbm": round(rg["gbs32"] / HBM_PEAK_GBS, 4)},
                "fp32_vs_fp64_err": rg["err"],
                "phases_ms_fp64": {k: round(v, 5) for k, v in rg["ph64"].items()},
                "phases_ms_fp32": {k: round(v, 5) for k, v in rg["ph32"].items()},
                "tuning": tuning_info(hg)}
        hg.close()

    cpu = None if args.no_cpu else cpu_baseline(system)

    out = {
        "metric": METRIC,
        "value": round(r["gbs64"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms64"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Si-64 structure from seeded displacements, GTH-form local part and Gaussian s/p "
                "KB projectors, random seeded psi)",
        "config": {"workload": "Si-64 H.psi block apply at one k-point: 2x2x2 Si cubic cells (64 atoms), 30 Ry, "
                               "75^3 grid, k = (1/2,1/2,1/2), N_pw = 23984, 160 bands, 256 projectors",
                   "n_pw": w.npw, "grid": list(w.grid), "bands": w.nbands, "projectors": info["projectors"],
                   "x_planes": info["planes"], "k_index": SI64_K, "parallelism": "single GPU"},
        "roofline": roofline(w, info, r["ph64"], "fp64"),
        "cpu_baseline": cpu,
        "frac_hbm": round(r["gbs64"] / HBM_PEAK_GBS, 4),
        "fp32": {"value": round(r["gbs32"], 2), "unit": "GB/s", "ms_per_step": round(r["ms32"], 4),
                 "frac_hbm": round(r["gbs32"] / HBM_PEAK_GBS, 4), "roofline": roofline(w, info, r["ph32"], "fp32")},
        "fp32_vs_fp64_err": r["err"],
        "tuning": tuned,
        # Per-phase times of the profiled pass (HIP events between the phases).
        "phases_ms_fp64": {k: round(v, 5) for k, v in r["ph64"].items()},
        "phases_ms_fp32": {k: round(v, 5) for k, v in r["ph32"].items()},
        "isolated_hpsi_bench": isolated,
        "batched_multi_k_al4": batched,
        "gaas64_hpsi": gaas,
        "scf_time_to_1e-6_Ha": scf_result,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
