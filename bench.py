#!/usr/bin/env python3
"""H.psi block-apply benchmark on one MI355X (the isolated H.psi kernel configuration).

Workload (synthetic, seeded): cubic box L = 20.5 bohr, ecut 15 Ha, 72^3 grid, N_pw = 23,847,
160 bands, V_loc = 64 Gaussians, 64 atoms x (s + 3p) Gaussian KB projectors.  One step = one
apply of H to the whole 160-band block with psi already resident in HBM.

value = algorithmic H.psi bytes per block / time per block (GB/s), for the FP32 apply (the
mixed-precision SCF path); the FP64 apply is timed the same way and reported beside it.
Algorithmic bytes per block (lower bound for any design that materialises psi(r) once per band):
    B = nb*s_c*(2*N_pw + 2*N) + s_r*N + 4*N_pw + s_r*4*N_pw
with s_c/s_r = 8/4 (FP32) or 16/8 (FP64).

The SCF metric (`--scf si64`, the default; `--scf none` skips it): time to 1e-6 Ha of the Si-64
SCF (2x2x2 cubic cells, 8 k-points, 160 bands, 75^3 grid) for plain FP64 and the mixed-precision,
k-concurrent forms with and without the learned preconditioner.

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

METRIC = "SCF time-to-1e-6 Ha & H·ψ GB/s, Si-64: mixed+batched+learned vs plain FP64"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}
TRAFFIC_FILE = ROOT / "profiles" / "traffic_bench.txt"


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
    """Bytes per launch of the dominant kernel from the committed PMC summary, if present."""
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


def cpu_baseline(w, budget_s=12.0):
    """Oracle (numpy FP64 restatement of the reference apply) on one host core."""
    from threadpoolctl import threadpool_limits

    from oracle import adapters, hamiltonian

    ctx_o = adapters.context_from_workload(w)
    done, t_used = 0, 0.0
    with threadpool_limits(limits=1):
        while t_used < budget_s and done < w.nbands:
            nb = min(4, w.nbands - done)
            t = time.perf_counter()
            hamiltonian.apply_hamiltonian_batched(ctx_o, w.psi[done:done + nb])
            t_used += time.perf_counter() - t
            done += nb
    gbs = hpsi_bytes(done, w.npw, int(np.prod(w.grid)), "fp64") / t_used / 1e9
    return {"value": round(gbs, 4), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP64 H.psi (numpy pocketfft + 1-thread BLAS) on {done} of {w.nbands} bands "
                      f"of the same workload, {t_used:.1f} s; "
                      f"{t_used / done * 1e3:.1f} ms/band -> {t_used / done * w.nbands:.1f} s per 160-band block"}


# SCF settings shared by every variant (only precision, k-point concurrency and the
# preconditioner differ): Pulay history 8 from iteration 4, beta 0.3, model-dielectric
# preconditioner diemac 12 (a Si setting of the reference's config), atomic start density,
# diagonal start vectors, eigensolver tolerance min(1e-4, drho / 10) >= 1e-7 on the occupied bands.
SCF_COMMON = dict(conv=1e-8, diemac=12.0, start="atomic", init_diagonal=True, eig_tol_ratio=0.1, eig_tol_min=1e-7)


def scf_leg(name, max_iter, log=sys.stderr):
    """Time-to-1e-6 Ha of the SCF for three forms: plain FP64 (k-points one after another on one
    stream), mixed + batched (FP32 H apply with FP64 subspace algebra and an FP64 apply below a
    1e-5 residual, every k-point on its own stream and host thread), and the same with the
    learned preconditioner on the kinetic-shift base (trained first, on the device, on
    residual/correction pairs of the warm-started solves of a 12-step FP64 SCF of the 8-atom Si
    cell at the same cutoff; its collection and training time is reported, not counted; the
    warm-up SCF steps run the base to give the network's fallback its yardstick).  E_inf is the
    plain FP64 run's final energy (the GPU SCF is tied to the FP64 oracle SCF on the small Si
    cell by the parity tests)."""
    import dzhip
    from dzhip import precond
    from dzhip import scf as dscf

    system = {"si64": dscf.si64, "si8": dscf.si8_small}[name]()
    nocc = int(np.ceil(system.nelec / 2))
    t0 = time.perf_counter()
    tctx = dzhip.Context(0)
    # Training system at the cutoff of the run, so the features cover its |k+G|^2 range.
    train_systems = [dscf.silicon(ecut_ry=system.ecut_ry, nbands=24,
                                  kpoints=([np.zeros(3), np.array([0.11, 0.23, 0.05])], [0.5, 0.5]))]
    params, hist, nsamp = precond.train_on_scf(tctx, train_systems, base=dzhip.PRECOND_KINETIC, steps=12,
                                               diemac=12.0, start="atomic", init_diagonal=True)
    tctx.close()
    train_s = time.perf_counter() - t0
    print(f"learned preconditioner: {nsamp} samples, loss {hist[0]:.4f} -> {hist[-1]:.4f}, {train_s:.1f} s",
          file=log, flush=True)
    variants = [("fp64", dzhip.FP64, False, dzhip.PRECOND_KINETIC),
                ("mixed_batched", dzhip.FP32, True, dzhip.PRECOND_KINETIC),
                ("mixed_batched_learned", dzhip.FP32, True, dzhip.PRECOND_MLP)]
    runs, setup = {}, {}
    mlp_solves = None
    for tag, prec, conc, pc in variants:
        drv = dscf.ScfDriver(system, concurrent_k=conc, device_grid=True)
        setup[tag] = drv.setup_s
        fallbacks = []
        if pc == dzhip.PRECOND_MLP:
            drv.set_precond_model(params, precond.ENERGY_SCALE, dzhip.PRECOND_KINETIC)
            drv.solve_hook = lambda k, ham, st: fallbacks.append(st["precond_fallback"])

        def show(r, tag=tag):
            sec = " ".join(f"{k[2:-3]} {v:.0f}" for k, v in r.sections_ms.items())
            print(f"scf {name} {tag} it {r.iteration} E {r.energy_ry:.10f} drho {r.drho:.3e} t {r.wall_s:.2f}s "
                  f"eig {r.eig_s:.2f}s eig_it {r.eig_iterations} applies {r.applies} fp32 {r.fp32_applies} "
                  f"gpu_ms[{sec}]", file=log, flush=True)
        try:
            runs[tag] = drv.run(max_iter=max_iter, precision=prec, preconditioner=pc, nconv=nocc, progress=show,
                                **SCF_COMMON)
        finally:
            drv.close()
        if pc == dzhip.PRECOND_MLP:
            mlp_solves = {"solves": len(fallbacks), "fell_back": sum(1 for f in fallbacks if f > 0)}
    e_inf = runs["fp64"].energies[-1]
    out = {"system": system.name, "grid": list(system.grid), "k_points": len(system.kpoints), "bands": system.nbands,
           "e_inf_ry": e_inf, "e_inf_source": "plain FP64 GPU run, final iteration",
           "settings": {k: v for k, v in SCF_COMMON.items()},
           "learned_training": {"samples": nsamp, "base": "kinetic shift", "loss_start": float(hist[0]),
                                "loss_end": float(hist[-1]), "seconds": round(train_s, 2),
                                "scf_solves_fallback": mlp_solves}}
    for tag, run in runs.items():
        t = run.time_to(e_inf)
        out[tag] = {"t_1e-6_Ha_s": None if t is None else round(t, 3), "iterations": len(run.records),
                    "converged": run.converged, "final_dE_Ha": (run.energies[-1] - e_inf) * dscf.HA_PER_RY,
                    "wall_s": round(run.records[-1].wall_s, 3), "setup_s": round(setup[tag], 3),
                    "eig_s": round(sum(r.eig_s for r in run.records), 3),
                    "davidson_iterations": sum(r.eig_iterations for r in run.records),
                    "applies": sum(r.applies for r in run.records),
                    "solver_gpu_ms": {k[2:-3]: round(sum(r.sections_ms.get(k, 0.0) for r in run.records), 1)
                                      for k in dscf.SECTIONS},
                    "fp32_applies": sum(r.fp32_applies for r in run.records)}
    a = out["fp64"]["t_1e-6_Ha_s"]
    for tag in ("mixed_batched", "mixed_batched_learned"):
        b = out[tag]["t_1e-6_Ha_s"]
        out[f"speedup_{tag}"] = None if not a or not b else round(a / b, 3)
    return out


def eig_profile(ham):
    """Section times of one cold and two warm FP64 Davidson solves (128 bands, 12 iterations,
    diagonal start) on the bench context, one JSON line per solve."""
    for rep in ("cold", "warm", "warm"):
        t0 = time.perf_counter()
        _, _, st = ham.eig_solve(128, tol=1e-4, max_iter=12, init_diagonal=True, want_vectors=False)
        wall = time.perf_counter() - t0
        print(json.dumps({"solve": rep, "wall_ms": round(wall * 1e3, 1), "iterations": st["iterations"],
                          "sections_ms": {k[2:-3]: round(v, 1) for k, v in st.items() if k.startswith("t_")}}),
              flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--scf", default="si64", choices=("none", "si8", "si64"),
                    help="also time the SCF to 1e-6 Ha (plain FP64 vs mixed+batched)")
    ap.add_argument("--scf-max-iter", type=int, default=45)
    ap.add_argument("--eig", action="store_true",
                    help="instead of the bench line: one cold and two warm FP64 Davidson solves of 128 bands "
                         "on the bench workload (12 iterations), with the solver's section times")
    ap.add_argument("--overlap", action="store_true",
                    help="run the nonlocal kernels on a second stream beside the local ones (timed pass only)")
    ap.add_argument("--xline-split", type=int, default=0,
                    help="fix the x-line variant: d > 0 LDS y-slabs, -1..-3 register kernel, -4..-6 lane-pair register kernel "
                         "(0 = autotune on the first apply)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (k-point sharding is "
              "design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import dzhip
    from dzhip import workloads

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    w = workloads.bench(nbands=160)
    ngrid = int(np.prod(w.grid))
    ctx = dzhip.Context(0)
    ham = dzhip.Hamiltonian.from_workload(ctx, w)
    if args.xline_split:
        ham.set_tuning(args.xline_split)
    if args.overlap:
        ham.set_overlap(True)
    if args.eig:
        eig_profile(ham)
        return
    info = ham.info()
    nb, npw = w.nbands, w.npw

    x64 = torch.from_numpy(w.psi.view(np.float64).copy()).cuda()
    y64 = torch.empty_like(x64)
    x32 = torch.from_numpy(w.psi.astype(np.complex64).view(np.float32).copy()).cuda()
    y32 = torch.empty_like(x32)

    ms32, ph32 = time_gpu(ham, x32, y32, nb, dzhip.FP32, args.steps, args.warmup, torch)
    ms64, ph64 = time_gpu(ham, x64, y64, nb, dzhip.FP64, args.steps, args.warmup, torch)

    tuned = ham.info()
    h32 = y32.cpu().numpy().view(np.complex64).reshape(nb, npw).astype(np.complex128)
    h64 = y64.cpu().numpy().view(np.complex128).reshape(nb, npw)
    err = float(np.max(np.linalg.norm(h32 - h64, axis=1) / np.linalg.norm(h64, axis=1)))
    abserr = float(np.max(np.abs(h32 - h64)))

    gbs32 = hpsi_bytes(nb, npw, ngrid, "fp32") / (ms32 * 1e-3) / 1e9
    gbs64 = hpsi_bytes(nb, npw, ngrid, "fp64") / (ms64 * 1e-3) / 1e9

    # Roofline of the dominant kernel of the FP32 apply (per-phase HIP events, same stream).
    dom = max(ph32, key=ph32.get)
    plane_elems = w.grid[1] * w.grid[2]
    bound, amount, _ = kernel_model(dom, nb, npw, ngrid, info["planes"], plane_elems, info["projectors"], "fp32")
    t_dom = ph32[dom] * 1e-3
    traffic = read_traffic().get(dom)
    if bound == "hbm":
        achieved = amount / t_dom / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes": amount, "avg_ms": round(ph32[dom], 5)}
    else:
        achieved = amount / t_dom / 1e12
        peak = MFMA_PEAK_TFLOPS["fp32"]
        roof = {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 2), "peak": peak,
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic,
                "algorithmic_flops": amount, "avg_ms": round(ph32[dom], 5)}

    cpu = None if args.no_cpu else cpu_baseline(w)

    out = {
        "metric": METRIC,
        "value": round(gbs32, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded PCG64 20261202: Gaussian V_loc, Gaussian s/p KB projectors, random psi)",
        "config": {"workload": "H.psi block apply: L=20.5 bohr cube, ecut 15 Ha, 72^3 grid, N_pw=23847, "
                               "160 bands, 64 atoms x (s+3p) projectors", "n_pw": npw, "grid": list(w.grid),
                   "bands": nb, "projectors": info["projectors"], "x_planes": info["planes"],
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp64": {"value": round(gbs64, 2), "unit": "GB/s", "ms_per_step": round(ms64, 4),
                 "frac_hbm": round(gbs64 / HBM_PEAK_GBS, 4)},
        "fp32_frac_hbm": round(gbs32 / HBM_PEAK_GBS, 4),
        "fp32_vs_fp64_err": {"max_band_rel_l2": err, "max_abs": abserr},
        "tuning": {"xline_split_fp32": tuned["xline_split_fp32"], "xline_split_fp64": tuned["xline_split_fp64"],
                   "xline_register_fp32": tuned["xline_register_fp32"],
                   "xline_register_fp64": tuned["xline_register_fp64"],
                   "xline_lanes_fp32": tuned["xline_lanes_fp32"], "xline_lanes_fp64": tuned["xline_lanes_fp64"]},
        # Per-phase times of the profiled pass (HIP events between the phases).
        "phases_ms_fp32": {k: round(v, 5) for k, v in ph32.items()},
        "phases_ms_fp64": {k: round(v, 5) for k, v in ph64.items()},
        "scf_time_to_1e-6_Ha": None if args.scf == "none" else scf_leg(args.scf, args.scf_max_iter),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
