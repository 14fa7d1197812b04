"""Python binding of libdzhip (ctypes over the C ABI in include/dzhip.h).

Mirrors the reference's operator interface for the H.psi path:
  FftPlan3d.forward/inverse      <- Fft3dPlan (src/features/fft/fft.zig:44-263)
  Hamiltonian.apply / apply_batch <- Operator.apply / apply_batch (src/features/linalg/lobpcg/common.zig:12-41)
                                    backed by apply_hamiltonian(_batched) (src/features/scf/apply.zig:643-658, 1584-1618)
  Hamiltonian.apply_s            <- apply_overlap (apply.zig:1041-1072), S = I for norm-conserving species
Errors: InvalidMatrixSize (reference error.InvalidMatrixSize) and HipFailure.

There is no CPU fallback: the library is loaded on first use and a missing
build raises ImportError; every compute call goes to the HIP library.
"""
from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libdzhip.so"

FP64 = 0
FP32 = 1

PHASES = ("nl_project", "nl_back", "plane_inv", "xline", "plane_fwd")

EXPORTS = (
    "dzhip_version", "dzhip_last_error", "dzhip_ctx_create", "dzhip_ctx_destroy",
    "dzhip_fft_plan_create", "dzhip_fft_plan_destroy", "dzhip_fft_forward", "dzhip_fft_inverse",
    "dzhip_ham_create", "dzhip_ham_destroy", "dzhip_ham_set_local", "dzhip_ham_apply",
    "dzhip_ham_apply_dev", "dzhip_ham_apply_nonlocal", "dzhip_ham_apply_overlap",
    "dzhip_ham_set_stream", "dzhip_ham_sync", "dzhip_ham_profile", "dzhip_ham_phase_times",
    "dzhip_ham_info", "dzhip_eig_solve", "dzhip_density_accumulate",
    "dzhip_density_accumulate_solution", "dzhip_ham_set_precond_model", "dzhip_ham_training_samples",
    "dzhip_mlp_loss_grad", "dzhip_mlp_train", "dzhip_cg_solve",
    "dzhip_scf_grid_create", "dzhip_scf_grid_destroy", "dzhip_scf_grid_set_ion", "dzhip_scf_grid_init",
    "dzhip_ham_set_local_grid", "dzhip_scf_grid_add_solution_density", "dzhip_scf_grid_hxc",
    "dzhip_scf_grid_mix", "dzhip_scf_grid_read", "dzhip_ham_set_tuning", "dzhip_ham_set_plane_kernel", "dzhip_scf_grid_set_dielectric",
    "dzhip_scf_grid_set_xc", "dzhip_ham_set_overlap",
    "dzhip_scf_grid_init_rho", "dzhip_heev", "dzhip_heev_dev", "dzhip_heev_check", "dzhip_heev_retries",
    "dzhip_hamset_create", "dzhip_hamset_destroy", "dzhip_hamset_offsets", "dzhip_hamset_apply_dev",
    "dzhip_hamset_apply", "dzhip_hamset_set_stream", "dzhip_hamset_info", "dzhip_heev_dev_noalloc",
)


class DzhipError(RuntimeError):
    code = -2


class InvalidMatrixSize(DzhipError):
    code = -1


class HipFailure(DzhipError):
    code = -2


class OutOfMemory(DzhipError):
    code = -3


class BadState(DzhipError):
    code = -4


class BadArgument(DzhipError):
    code = -5


_ERRS = {-1: InvalidMatrixSize, -2: HipFailure, -3: OutOfMemory, -4: BadState, -5: BadArgument}

PRECOND_KINETIC = 0
PRECOND_TPA = 1
PRECOND_MLP = 2
XC_LDA_PZ = 0
XC_PBE = 1
MLP_PARAMS = 1217


class EigOpts(ct.Structure):
    _fields_ = [
        ("max_iter", ct.c_uint64), ("tol", ct.c_double), ("max_subspace", ct.c_uint64),
        ("block_size", ct.c_uint64), ("init_diagonal", ct.c_int), ("precision", ct.c_int),
        ("mixed_switch_tol", ct.c_double), ("preconditioner", ct.c_int), ("seed", ct.c_uint64),
        ("reuse_vectors", ct.c_int), ("collect_snapshots", ct.c_uint64), ("nconv", ct.c_uint64),
    ]


class EigStats(ct.Structure):
    _fields_ = [
        ("iterations", ct.c_uint64), ("apply_calls", ct.c_uint64), ("fp32_applies", ct.c_uint64),
        ("fp64_applies", ct.c_uint64), ("restarts", ct.c_uint64), ("precision_switch_iter", ct.c_uint64),
        ("converged", ct.c_uint64), ("max_residual", ct.c_double), ("initial_residual", ct.c_double),
        ("precond_fallback", ct.c_uint64), ("t_rr_ms", ct.c_double), ("t_ritz_ms", ct.c_double),
        ("t_precond_ms", ct.c_double), ("t_ortho_ms", ct.c_double), ("t_apply_ms", ct.c_double),
        ("t_extend_ms", ct.c_double), ("ortho_fallbacks", ct.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SpeciesDesc(ct.Structure):
    _fields_ = [
        ("n_beta", ct.c_uint64), ("l_list", ct.POINTER(ct.c_int32)), ("dij", ct.POINTER(ct.c_double)),
        ("dij_m", ct.POINTER(ct.c_double)), ("m_total", ct.c_uint64), ("phi", ct.POINTER(ct.c_double)),
    ]


class HamDesc(ct.Structure):
    _fields_ = [
        ("nx", ct.c_uint64), ("ny", ct.c_uint64), ("nz", ct.c_uint64), ("n_pw", ct.c_uint64),
        ("fft_index", ct.POINTER(ct.c_int32)), ("kinetic", ct.POINTER(ct.c_double)),
        ("gcart", ct.POINTER(ct.c_double)), ("local_r", ct.POINTER(ct.c_double)),
        ("inv_volume", ct.c_double), ("n_species", ct.c_uint64), ("species", ct.POINTER(SpeciesDesc)),
        ("n_atoms", ct.c_uint64), ("atom_pos", ct.POINTER(ct.c_double)),
        ("atom_species", ct.POINTER(ct.c_int32)),
    ]


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"libdzhip.so not built at {LIB_PATH}; run dft-zig_amd/dzhip/build.py "
                          "(there is no CPU fallback)")
    # torch ships its own libamdhip64.so.7 / libhsa-runtime64.so.1.  Loading it first lets the
    # dynamic loader bind libdzhip to that same runtime (same sonames); loading libdzhip first
    # would put a second HIP runtime into the process and the later one finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ct.CDLL(str(LIB_PATH))
    vp, u64, i32, dp = ct.c_void_p, ct.c_uint64, ct.c_int, ct.POINTER(ct.c_double)
    fp = ct.POINTER(ct.c_float)
    sig = {
        "dzhip_version": ([], i32), "dzhip_last_error": ([], ct.c_char_p),
        "dzhip_ctx_create": ([i32], vp), "dzhip_ctx_destroy": ([vp], None),
        "dzhip_fft_plan_create": ([vp, u64, u64, u64, i32], vp), "dzhip_fft_plan_destroy": ([vp], None),
        "dzhip_fft_forward": ([vp, dp, u64], i32), "dzhip_fft_inverse": ([vp, dp, u64], i32),
        "dzhip_ham_create": ([vp, ct.POINTER(HamDesc)], vp), "dzhip_ham_destroy": ([vp], None),
        "dzhip_ham_set_local": ([vp, dp, u64], i32),
        "dzhip_ham_apply": ([vp, dp, dp, u64, u64, i32], i32),
        "dzhip_ham_apply_dev": ([vp, vp, vp, u64, u64, i32], i32),
        "dzhip_ham_apply_nonlocal": ([vp, dp, dp, u64, u64, i32], i32),
        "dzhip_ham_apply_overlap": ([vp, dp, dp, u64, u64], i32),
        "dzhip_ham_set_stream": ([vp, vp], i32), "dzhip_ham_sync": ([vp], i32),
        "dzhip_ham_profile": ([vp, i32], i32),
        "dzhip_ham_phase_times": ([vp, dp, u64, ct.POINTER(ct.c_uint64)], i32),
        "dzhip_ham_info": ([vp, ct.POINTER(ct.c_uint64), u64], i32),
        "dzhip_eig_solve": ([vp, dp, u64, u64, ct.POINTER(EigOpts), dp, u64, dp, dp, ct.POINTER(EigStats)], i32),
        "dzhip_density_accumulate": ([vp, dp, u64, u64, dp, dp, u64, i32], i32),
        "dzhip_density_accumulate_solution": ([vp, u64, dp, dp, u64, i32], i32),
        "dzhip_ham_set_precond_model": ([vp, fp, u64, ct.c_double, i32], i32),
        "dzhip_ham_training_samples": ([vp, ct.c_double, i32, u64, fp, u64, ct.POINTER(ct.c_uint64)], i32),
        "dzhip_mlp_loss_grad": ([vp, fp, u64, fp, dp, dp], i32),
        "dzhip_mlp_train": ([vp, fp, u64, fp, u64, u64, ct.c_double, u64, dp], i32),
        "dzhip_cg_solve": ([vp, dp, u64, u64, ct.POINTER(EigOpts), dp, u64, dp, dp, ct.POINTER(EigStats)], i32),
        "dzhip_scf_grid_create": ([vp, u64, u64, u64, dp, ct.c_double, u64], vp),
        "dzhip_scf_grid_destroy": ([vp], None),
        "dzhip_scf_grid_set_ion": ([vp, dp, u64], i32),
        "dzhip_scf_grid_init": ([vp, ct.c_double], i32),
        "dzhip_ham_set_local_grid": ([vp, vp], i32),
        "dzhip_scf_grid_add_solution_density": ([vp, vp, dp, u64], i32),
        "dzhip_scf_grid_hxc": ([vp, dp, u64], i32),
        "dzhip_scf_grid_mix": ([vp, ct.c_double, i32], i32),
        "dzhip_scf_grid_read": ([vp, i32, dp, u64], i32),
        "dzhip_ham_set_tuning": ([vp, i32], i32),
        "dzhip_ham_set_plane_kernel": ([vp, i32], i32),
        "dzhip_scf_grid_set_dielectric": ([vp, ct.c_double, ct.c_double], i32),
        "dzhip_scf_grid_set_xc": ([vp, i32], i32),
        "dzhip_ham_set_overlap": ([vp, i32], i32),
        "dzhip_scf_grid_init_rho": ([vp, dp, u64], i32),
        "dzhip_heev": ([vp, dp, u64, u64, u64, dp, dp, u64], i32),
        "dzhip_heev_dev": ([vp, vp, u64, u64, u64, vp, vp, u64, vp], i32),
        "dzhip_heev_check": ([vp, vp], i32),
        "dzhip_heev_retries": ([], u64),
        "dzhip_heev_dev_noalloc": ([vp, vp, u64, u64, u64, vp, vp, u64, vp], i32),
        "dzhip_hamset_create": ([ct.POINTER(vp), u64], vp), "dzhip_hamset_destroy": ([vp], None),
        "dzhip_hamset_offsets": ([vp, ct.POINTER(ct.c_uint64), u64], i32),
        "dzhip_hamset_apply_dev": ([vp, vp, vp, u64, i32, i32], i32),
        "dzhip_hamset_apply": ([vp, ct.POINTER(vp), ct.POINTER(vp), u64, i32], i32),
        "dzhip_hamset_set_stream": ([vp, vp], i32),
        "dzhip_hamset_info": ([vp, ct.POINTER(ct.c_uint64), u64], i32),
    }
    missing = set(EXPORTS) - set(sig)
    if missing:
        raise ImportError(f"no ctypes signature for {sorted(missing)}")
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_LIB = None


def L():
    """The loaded library (loaded on first use; raises if it was never built)."""
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def last_error() -> str:
    return L().dzhip_last_error().decode(errors="replace")


def heev_retries() -> int:
    """Rayleigh-Ritz tridiagonalisations repeated cooperatively in this process (their ordinary
    launch beside other k-points' work gave up waiting for its workgroups)."""
    return int(L().dzhip_heev_retries())


def check(rc: int) -> None:
    if rc != 0:
        raise _ERRS.get(rc, DzhipError)(f"dzhip error {rc}: {last_error()}")


def _fp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_float))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ct.POINTER(ct.c_int32))


def _as_c128(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.complex128)


class Context:
    """One device context (metal_create_context counterpart)."""

    def __init__(self, device: int = 0):
        self.handle = L().dzhip_ctx_create(int(device))
        if not self.handle:
            raise HipFailure(f"dzhip_ctx_create({device}) failed: {last_error()}")
        self.device = device

    def mlp_loss_grad(self, samples, params):
        """Full-batch loss and gradient (FP64) of the learned preconditioner on (6, n) samples."""
        smp = np.ascontiguousarray(samples, dtype=np.float32)
        p = np.ascontiguousarray(params, dtype=np.float32)
        g = np.zeros(MLP_PARAMS)
        loss = ct.c_double(0.0)
        check(L().dzhip_mlp_loss_grad(self.handle, _fp(smp), smp.shape[1], _fp(p), _dp(g), ct.byref(loss)))
        return loss.value, g

    def mlp_train(self, samples, params, epochs: int, batch: int = 4096, lr: float = 1e-3, seed: int = 1):
        """Adam on the device; returns (params, loss after each epoch)."""
        smp = np.ascontiguousarray(samples, dtype=np.float32)
        p = np.ascontiguousarray(params, dtype=np.float32).copy()
        hist = np.zeros(max(1, epochs))
        check(L().dzhip_mlp_train(self.handle, _fp(smp), smp.shape[1], _fp(p), epochs, batch, lr, seed, _dp(hist)))
        return p, hist[:epochs]

    def heev(self, a, nev=None):
        """Lowest nev eigenpairs of the Hermitian matrix a (full storage) on the device:
        (values ascending, vectors as columns of an m x nev array)."""
        a = np.asarray(a, dtype=np.complex128)
        m = a.shape[0]
        if a.shape != (m, m):
            raise InvalidMatrixSize("heev needs a square matrix")
        nev = m if nev is None else int(nev)
        at = np.ascontiguousarray(a.T)  # column-major a
        w = np.zeros(nev)
        zt = np.zeros((nev, m), dtype=np.complex128)  # column-major m x nev
        check(L().dzhip_heev(self.handle, _dp(at.view(np.float64)), m, m, nev, _dp(w), _dp(zt.view(np.float64)), m))
        return w, zt.T

    def heev_dev(self, a_ptr, lda, m, nev, w_ptr, z_ptr, ldz, stream=None):
        """Device-pointer form (enqueue only); heev_check() synchronises and reports a failure."""
        check(L().dzhip_heev_dev(self.handle, a_ptr, lda, m, nev, w_ptr, z_ptr, ldz, stream))

    def heev_dev_noalloc(self, a_ptr, lda, m, nev, w_ptr, z_ptr, ldz, stream=None):
        """The allocation-free form (m <= 256, the library's static scratch; enqueue only)."""
        check(L().dzhip_heev_dev_noalloc(self.handle, a_ptr, lda, m, nev, w_ptr, z_ptr, ldz, stream))

    def heev_check(self, stream=None):
        check(L().dzhip_heev_check(self.handle, stream))

    def close(self):
        if self.handle:
            L().dzhip_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FftPlan3d:
    """In-place 3D FFT on host arrays with the reference conventions (forward unnormalised,
    inverse 1/N), index x + nx*(y + ny*z)."""

    def __init__(self, ctx: Context, nx: int, ny: int, nz: int, precision: int = FP64):
        self.ctx = ctx
        self.shape = (nx, ny, nz)
        self.handle = L().dzhip_fft_plan_create(ctx.handle, nx, ny, nz, precision)
        if not self.handle:
            raise BadArgument(f"dzhip_fft_plan_create failed: {last_error()}")

    def forward(self, data: np.ndarray) -> None:
        check(L().dzhip_fft_forward(self.handle, _dp(data.view(np.float64)), data.size))

    def inverse(self, data: np.ndarray) -> None:
        check(L().dzhip_fft_inverse(self.handle, _dp(data.view(np.float64)), data.size))

    def close(self):
        if self.handle:
            L().dzhip_fft_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Hamiltonian:
    """Per-k H.psi context on the device (ApplyContext counterpart).

    Arguments mirror what the reference's SCF driver holds for one k-point:
    grid (nx, ny, nz), fft_index (PwGridMap.fft_indices), kinetic (|k+G|^2, Ry),
    gcart (G, bohr^-1), local_r (Ry, FFT order), inv_volume, species list of
    dicts {l_list, dij, phi, [dij_m]}, atom positions (bohr) and species indices.
    """

    def __init__(self, ctx: Context, grid, fft_index, kinetic, gcart, local_r, inv_volume,
                 species=(), atom_pos=None, atom_species=None):
        self.ctx = ctx
        self.grid = tuple(int(g) for g in grid)
        self.n = int(np.asarray(kinetic).shape[0])
        keep = []

        def arr(a, dt):
            v = np.ascontiguousarray(a, dtype=dt)
            keep.append(v)
            return v

        fidx = arr(fft_index, np.int32)
        kin = arr(kinetic, np.float64)
        self._kinetic = kin
        gc = arr(gcart, np.float64).reshape(-1)
        sp_arr = (SpeciesDesc * max(1, len(species)))()
        for i, sp in enumerate(species):
            ls = arr(sp["l_list"], np.int32)
            dij = arr(sp["dij"], np.float64).reshape(-1)
            phi = arr(sp["phi"], np.float64).reshape(-1)
            mt = int(sum(2 * int(l) + 1 for l in ls))
            sp_arr[i].n_beta = ls.shape[0]
            sp_arr[i].l_list = _ip(ls)
            sp_arr[i].dij = _dp(dij)
            if sp.get("dij_m") is not None:
                dm = arr(sp["dij_m"], np.float64).reshape(-1)
                sp_arr[i].dij_m = _dp(dm)
            else:
                sp_arr[i].dij_m = ct.POINTER(ct.c_double)()
            sp_arr[i].m_total = mt
            sp_arr[i].phi = _dp(phi)
        d = HamDesc()
        d.nx, d.ny, d.nz = self.grid
        d.n_pw = self.n
        d.fft_index = _ip(fidx)
        d.kinetic = _dp(kin)
        d.gcart = _dp(gc)
        if local_r is not None:
            lr = arr(local_r, np.float64)
            d.local_r = _dp(lr)
        else:
            d.local_r = ct.POINTER(ct.c_double)()
        d.inv_volume = float(inv_volume)
        d.n_species = len(species)
        d.species = sp_arr if species else ct.POINTER(SpeciesDesc)()
        if species and atom_pos is not None and len(atom_pos) > 0:
            ap = arr(atom_pos, np.float64).reshape(-1)
            asp = arr(atom_species, np.int32)
            d.n_atoms = asp.shape[0]
            d.atom_pos = _dp(ap)
            d.atom_species = _ip(asp)
        else:
            d.n_atoms = 0
            d.atom_pos = ct.POINTER(ct.c_double)()
            d.atom_species = ct.POINTER(ct.c_int32)()
        self.handle = L().dzhip_ham_create(ctx.handle, ct.byref(d))
        if not self.handle:
            raise BadArgument(f"dzhip_ham_create failed: {last_error()}")

    @classmethod
    def from_workload(cls, ctx: Context, w, local_r=None):
        from . import tables

        species = [{"l_list": sp.l_list, "dij": sp.dij, "phi": w.phi(s)} for s, sp in enumerate(w.species)]
        return cls(ctx, w.grid, tables.fft_indices(w.basis, w.grid), w.basis.kinetic, w.basis.cart,
                   w.local_r if local_r is None else local_r, w.inv_volume, species, w.atom_pos,
                   w.atom_species)

    # -- reference callback shapes ------------------------------------------------
    def set_local(self, local_r: np.ndarray) -> None:
        lr = np.ascontiguousarray(local_r, dtype=np.float64)
        check(L().dzhip_ham_set_local(self.handle, _dp(lr), lr.size))

    def _host_call(self, fn, x, precision, *extra):
        x = _as_c128(x)
        if x.ndim == 1:
            ncols, n = 1, x.shape[0]
        else:
            ncols, n = x.shape
        y = np.empty_like(x)
        check(fn(self.handle, _dp(x.view(np.float64)), _dp(y.view(np.float64)), n, ncols, *extra))
        return y

    def apply(self, x, precision: int = FP64) -> np.ndarray:
        """y = H x for one vector (apply_hamiltonian) or (ncols, n) block (apply_hamiltonian_batched)."""
        return self._host_call(L().dzhip_ham_apply, x, precision, precision)

    apply_batch = apply

    def apply_nonlocal(self, x, precision: int = FP64) -> np.ndarray:
        return self._host_call(L().dzhip_ham_apply_nonlocal, x, precision, precision)

    def apply_s(self, x) -> np.ndarray:
        return self._host_call(L().dzhip_ham_apply_overlap, x, FP64)

    def apply_dev(self, x_ptr: int, y_ptr: int, ncols: int, precision: int = FP64) -> None:
        """Device pointers (e.g. torch tensors' data_ptr()), enqueued without waiting."""
        check(L().dzhip_ham_apply_dev(self.handle, ct.c_void_p(x_ptr), ct.c_void_p(y_ptr), self.n,
                                      ncols, precision))

    def set_stream(self, stream_handle: int | None) -> None:
        check(L().dzhip_ham_set_stream(self.handle, ct.c_void_p(stream_handle or 0)))

    def sync(self) -> None:
        check(L().dzhip_ham_sync(self.handle))

    def profile(self, enable: bool) -> None:
        check(L().dzhip_ham_profile(self.handle, 1 if enable else 0))

    def phase_times(self):
        out = (ct.c_double * len(PHASES))()
        calls = ct.c_uint64()
        check(L().dzhip_ham_phase_times(self.handle, out, len(PHASES), ct.byref(calls)))
        return dict(zip(PHASES, list(out))), int(calls.value)

    def eig_solve(self, nbands: int, diag=None, tol: float = 1e-6, max_iter: int = 40, precision: int = FP64,
                  preconditioner: int = PRECOND_KINETIC, init=None, max_subspace: int = 0, block_size: int = 0,
                  init_diagonal: bool = False, mixed_switch_tol: float = 0.0, seed: int = 0,
                  want_vectors: bool = True, reuse: bool = False, collect_snapshots: int = 0, nconv: int = 0):
        """Lowest eigenpairs (hermitian_eigen_decomp_iterative counterpart).  diag defaults to the
        kinetic energies the context was built with; init is an (init_cols, n) block of start vectors;
        reuse (without init) starts from the previous solve's vectors kept on the device.
        Returns (values[nbands], vectors (nbands, n) or None, stats dict)."""
        if diag is None:
            diag = self._kinetic
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        o = EigOpts(max_iter, tol, max_subspace, block_size, 1 if init_diagonal else 0, precision,
                    mixed_switch_tol, preconditioner, seed, 1 if reuse else 0, collect_snapshots, nconv)
        vals = np.zeros(nbands)
        vecs = np.empty((nbands, self.n), dtype=np.complex128) if want_vectors else None
        st = EigStats()
        if init is not None:
            init = np.ascontiguousarray(init, dtype=np.complex128)
            ip, ic = _dp(init.view(np.float64)), init.shape[0]
        else:
            ip, ic = ct.POINTER(ct.c_double)(), 0
        vp = _dp(vecs.view(np.float64)) if vecs is not None else ct.POINTER(ct.c_double)()
        check(L().dzhip_eig_solve(self.handle, _dp(diag), diag.shape[0], nbands, ct.byref(o), ip, ic,
                                  _dp(vals), vp, ct.byref(st)))
        return vals, vecs, st.as_dict()

    def density(self, psi, weights, rho=None):
        """rho(r) (+)= (1/Omega) sum_b w_b |psi_b(r)|^2 in FFT order (accumulate_band_density_fft)."""
        psi = _as_c128(psi)
        if psi.ndim == 1:
            psi = psi[None, :]
        w = np.ascontiguousarray(weights, dtype=np.float64)
        if w.shape[0] != psi.shape[0]:
            raise InvalidMatrixSize("one weight per band")
        total = int(np.prod(self.grid))
        acc = rho is not None
        out = np.ascontiguousarray(rho, dtype=np.float64).copy() if acc else np.zeros(total)
        check(L().dzhip_density_accumulate(self.handle, _dp(psi.view(np.float64)), psi.shape[1], psi.shape[0],
                                           _dp(w), _dp(out), total, 1 if acc else 0))
        return out

    def cg_solve(self, nbands: int, diag=None, tol: float = 0.0, max_iter: int = 0, init=None, seed: int = 0,
                 want_vectors: bool = True, reuse: bool = False):
        """Band-by-band CG (cg.solve counterpart; tol 0 = 1e-6, max_iter 0 = 20).
        Returns (values[nbands], vectors (nbands, n) or None, stats dict)."""
        if diag is None:
            diag = self._kinetic
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        o = EigOpts(max_iter, tol, 0, 0, 0, FP64, 0.0, PRECOND_KINETIC, seed, 1 if reuse else 0, 0, 0)
        vals = np.zeros(nbands)
        vecs = np.empty((nbands, self.n), dtype=np.complex128) if want_vectors else None
        st = EigStats()
        if init is not None:
            init = np.ascontiguousarray(init, dtype=np.complex128)
            ip, ic = _dp(init.view(np.float64)), init.shape[0]
        else:
            ip, ic = ct.POINTER(ct.c_double)(), 0
        vp = _dp(vecs.view(np.float64)) if vecs is not None else ct.POINTER(ct.c_double)()
        check(L().dzhip_cg_solve(self.handle, _dp(diag), diag.shape[0], nbands, ct.byref(o), ip, ic,
                                 _dp(vals), vp, ct.byref(st)))
        return vals, vecs, st.as_dict()

    def set_precond_model(self, params, energy_scale: float, base: int = PRECOND_TPA):
        """Install the learned preconditioner's parameters (MLP_PARAMS floats) and the base
        preconditioner they rescale (PRECOND_KINETIC or PRECOND_TPA)."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        check(L().dzhip_ham_set_precond_model(self.handle, _fp(p), p.shape[0], float(energy_scale), int(base)))

    def training_samples(self, energy_scale: float, stride: int = 1, capacity: int = 1 << 20,
                         base: int = PRECOND_TPA):
        """Residual/correction training samples of the last solve (collect_snapshots > 0), as a
        (6, count) float32 array: x1, x2, x3, base factor B, target, weight."""
        out = np.zeros((6, capacity), dtype=np.float32)
        cnt = ct.c_uint64(0)
        check(L().dzhip_ham_training_samples(self.handle, float(energy_scale), int(base), stride, _fp(out),
                                             capacity, ct.byref(cnt)))
        return np.ascontiguousarray(out[:, :cnt.value])

    def density_solution(self, weights, rho=None):
        """density() of the first len(weights) vectors of the last eig_solve, read on the device."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        total = int(np.prod(self.grid))
        acc = rho is not None
        out = np.ascontiguousarray(rho, dtype=np.float64).copy() if acc else np.zeros(total)
        check(L().dzhip_density_accumulate_solution(self.handle, w.shape[0], _dp(w), _dp(out), total,
                                                    1 if acc else 0))
        return out

    def set_tuning(self, xline_split: int = 0):
        """Force the x-line variant: d > 0 LDS y-slabs, -1..-3 / -4..-6 / -7..-9 the one-lane /
        lane-pair / lane-triple register kernel in 256 / 128 / 64-thread workgroups (0: autotune)."""
        check(L().dzhip_ham_set_tuning(self.handle, int(xline_split)))

    def set_plane_kernel(self, kind: int = 0):
        """Force the plane kernels: 1 full-tile LDS, 2 z-register (0: autotune)."""
        check(L().dzhip_ham_set_plane_kernel(self.handle, int(kind)))

    def set_overlap(self, on: bool = True):
        """Run the nonlocal kernels on a second stream, concurrently with the local ones."""
        check(L().dzhip_ham_set_overlap(self.handle, 1 if on else 0))

    def info(self):
        out = (ct.c_uint64 * 15)()
        check(L().dzhip_ham_info(self.handle, out, 15))
        keys = ("n_pw", "grid_points", "planes", "projectors", "device_bytes", "xline_split_fp32",
                "xline_split_fp64", "xline_register_fp32", "xline_register_fp64", "xline_lanes_fp32",
                "xline_lanes_fp64", "xline_kind_fp32", "xline_kind_fp64", "plane_kind_fp32",
                "plane_kind_fp64")
        return dict(zip(keys, [int(v) for v in out]))

    def close(self):
        if self.handle:
            L().dzhip_ham_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


GRID_RHO, GRID_VIN, GRID_VOUT, GRID_LOCAL, GRID_VION = range(5)


class HamSet:
    """Several k-point Hamiltonians of one grid applied in one launch sequence (dzhip_hamset_*).
    Device layout: band-major over the concatenated spheres, (k, b, G) at b * total + offsets[k] + G."""

    def __init__(self, hams):
        self.hams = list(hams)
        arr = (ct.c_void_p * len(self.hams))(*[h.handle for h in self.hams])
        self.handle = L().dzhip_hamset_create(arr, len(self.hams))
        if not self.handle:
            raise _ERRS.get(-5, DzhipError)(f"dzhip_hamset_create failed: {last_error()}")
        off = (ct.c_uint64 * (len(self.hams) + 1))()
        check(L().dzhip_hamset_offsets(self.handle, off, len(self.hams) + 1))
        self.offsets = [int(o) for o in off]
        self.total = self.offsets[-1]

    def apply(self, blocks, precision: int = FP64):
        """Per-k host blocks (ncols, n_pw(k)) -> per-k H blocks, all k-points in one launch sequence."""
        xs = [_as_c128(b) for b in blocks]
        ncols = xs[0].shape[0]
        ys = [np.empty_like(x) for x in xs]
        xa = (ct.c_void_p * len(xs))(*[x.ctypes.data for x in xs])
        ya = (ct.c_void_p * len(ys))(*[y.ctypes.data for y in ys])
        check(L().dzhip_hamset_apply(self.handle, xa, ya, ncols, precision))
        return ys

    def apply_dev(self, x_ptr: int, y_ptr: int, ncols: int, precision: int = FP64, use_graph: bool = False):
        check(L().dzhip_hamset_apply_dev(self.handle, ct.c_void_p(x_ptr), ct.c_void_p(y_ptr), ncols, precision,
                                         1 if use_graph else 0))

    def set_stream(self, stream_handle):
        check(L().dzhip_hamset_set_stream(self.handle, ct.c_void_p(stream_handle or 0)))

    def info(self):
        out = (ct.c_uint64 * 4)()
        check(L().dzhip_hamset_info(self.handle, out, 4))
        return dict(zip(("k_points", "n_pw_total", "planes_total", "graphs"), [int(v) for v in out]))

    def pack(self, blocks):
        """Per-k (ncols, n_pw(k)) blocks -> the set's band-major (ncols, total) layout."""
        ncols = blocks[0].shape[0]
        out = np.zeros((ncols, self.total), dtype=np.result_type(*[b.dtype for b in blocks]))
        for k, b in enumerate(blocks):
            out[:, self.offsets[k]:self.offsets[k + 1]] = b
        return out

    def unpack(self, arr):
        return [arr[:, self.offsets[k]:self.offsets[k + 1]] for k in range(len(self.hams))]

    def close(self):
        if self.handle:
            L().dzhip_hamset_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ScfGrid:
    """Device-resident SCF grid state (dzhip_scf_grid): density sum, V_in/V_out, Pulay history."""

    def __init__(self, ctx: Context, grid, recip, volume: float, pulay_history: int = 8):
        self.grid = tuple(int(v) for v in grid)
        self.n = int(np.prod(self.grid))
        r = np.ascontiguousarray(np.asarray(recip, dtype=np.float64).reshape(9))
        self.handle = L().dzhip_scf_grid_create(ctx.handle, *self.grid, _dp(r), float(volume), pulay_history)
        if not self.handle:
            raise BadArgument(f"dzhip_scf_grid_create failed: {last_error()}")
        self._ctx = ctx

    def set_ion(self, v_ion):
        v = np.ascontiguousarray(v_ion, dtype=np.float64)
        check(L().dzhip_scf_grid_set_ion(self.handle, _dp(v), v.shape[0]))

    def init(self, nelec: float):
        check(L().dzhip_scf_grid_init(self.handle, float(nelec)))

    def init_rho(self, rho):
        r = np.ascontiguousarray(rho, dtype=np.float64)
        check(L().dzhip_scf_grid_init_rho(self.handle, _dp(r), r.shape[0]))

    def set_local(self, ham: "Hamiltonian"):
        check(L().dzhip_ham_set_local_grid(ham.handle, self.handle))

    def add_density(self, ham: "Hamiltonian", weights):
        w = np.ascontiguousarray(weights, dtype=np.float64)
        check(L().dzhip_scf_grid_add_solution_density(self.handle, ham.handle, _dp(w), w.shape[0]))

    def hxc(self):
        """(E_H, E_xc, int V_in rho, RMS drho) of the density sum, which becomes rho."""
        out = np.zeros(4)
        check(L().dzhip_scf_grid_hxc(self.handle, _dp(out), 4))
        return tuple(float(v) for v in out)

    def set_dielectric(self, diemac: float = 1.0, dielng: float = 1.0):
        check(L().dzhip_scf_grid_set_dielectric(self.handle, float(diemac), float(dielng)))

    def set_xc(self, functional: str):
        """Exchange-correlation functional of V_xc and E_xc: "lda_pz" (default) or "pbe"."""
        code = {"lda_pz": XC_LDA_PZ, "pbe": XC_PBE}[functional]
        check(L().dzhip_scf_grid_set_xc(self.handle, code))

    def mix(self, beta: float, pulay: bool):
        check(L().dzhip_scf_grid_mix(self.handle, float(beta), 1 if pulay else 0))

    def read(self, which: int):
        out = np.zeros(self.n)
        check(L().dzhip_scf_grid_read(self.handle, which, _dp(out), self.n))
        return out

    def close(self):
        if self.handle:
            L().dzhip_scf_grid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
