"""SCF driver around the device eigensolver (host orchestration).

Follows the reference's SCF loop (src/features/scf/scf.zig:489-521: compute_density ->
build_potential_grid -> mix) with its defaults for the measured systems: fixed occupations
(src/features/scf/kpoint/density.zig:98-107), LDA-PZ (src/features/xc/xc.zig:195-210, 344-384),
Hartree (src/features/scf/potential.zig:34-44), potential mixing, linear (mixing.zig:10-18) then
Pulay/DIIS with history 8 from iteration 4 (mixing.zig:183-234, scf.zig:147-188, defaults
config/defaults.zig:42-45), convergence on the RMS density change (util.zig:17-24), iterative
eigensolver schedule of kpoint/solve.zig:294-402 (warm-up: 2 SCF steps at tol 1e-3 and 10
iterations, then tol 1e-4 and 20 iterations, vectors reused from the previous SCF step;
config/defaults.zig:32-40).

Per k-point the eigenproblem and the |psi|^2 density run on the GPU through libdzhip
(dzhip_eig_solve with the device warm-start cache, dzhip_density_accumulate_solution), so no
wave function crosses PCIe inside an SCF iteration; the potentials and the mixing run on the host
in FP64 as in the reference.  Each k-point may get its own context (own HIP stream and rocBLAS
handle) and run on its own host thread, the k-point worker scheme of kpoint/workers.zig:82-188.

Systems use a GTH-form synthetic pseudopotential (no UPF files ship with the reference): local
part v(q) = 2[-4 pi Z/q^2 + C1 (2 pi)^{3/2} rloc^3] e^{-q^2 rloc^2/2} Ry per atom, G = 0 dropped
(hamiltonian.zig:457-461), and Gaussian s/p Kleinman-Bylander projectors with D in Ry.  Total
energies leave out the ion-ion and psp-core constants of the structure.
"""
from __future__ import annotations

import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import FP64, PRECOND_KINETIC, PRECOND_MLP, PRECOND_TPA, Context, Hamiltonian, ScfGrid, tables, upf
from .workloads import SpeciesDesc

HA_PER_RY = 0.5
SECTIONS = ("t_rr_ms", "t_ritz_ms", "t_precond_ms", "t_ortho_ms", "t_apply_ms", "t_extend_ms")


@dataclass
class ScfSystem:
    name: str
    cell: np.ndarray
    recip: np.ndarray
    volume: float
    grid: tuple
    ecut_ry: float
    atom_pos: np.ndarray
    atom_species: np.ndarray
    species: list
    z_ion: float
    rloc: float
    c1: float  # Ha
    nelec: float
    nbands: int
    kpoints: list  # Cartesian, bohr^-1
    kweights: list
    smearing: float = 0.0  # Fermi-Dirac width (Ry); 0 = fixed occupations
    atom_rho_width: float = 1.2  # bohr: Gaussian valence charge per atom of the atomic start density
    xc: str = "lda_pz"  # "lda_pz" or "pbe" (PBE runs on the device grid)
    # Per-species GTH local parameters (Z, rloc, C1) for structures of several species; None = the
    # single set (z_ion, rloc, c1) for every atom.
    species_local: list | None = None
    # Per-species UPF data (dzhip.upf.Upf) when the species were read from UPF files: V_loc(q) is
    # then the file's short-range local form factor (Z from species_local).
    species_upf: list | None = None

    def local_params(self):
        """(Z, rloc, C1) of every species."""
        if self.species_local is not None:
            return list(self.species_local)
        return [(self.z_ion, self.rloc, self.c1)] * len(self.species)


# GTH-form parameters of a four-electron Si-like pseudo-atom (Hartree units for C1 and h): the
# local part, s-projector radius and strength of the Goedecker-Teter-Hutter Si LDA set, one
# Gaussian p projector of radius r_p with h_p = -0.3 Ha.  With these the diamond cell is an
# insulator with 16 valence bands (degeneracies 1, 6, 6, 3 at the cubic-cell Gamma point) and a
# gap of 0.08 Ry over a 4x4x4 mesh of the cubic cell, so fixed occupations are well defined.
SI_GTH = {"z": 4.0, "rloc": 0.44, "c1": -7.336103, "r_s": 0.422738, "h_s": 5.906928,
          "r_p": 0.484278, "h_p": -0.3}
SI_A_ANGSTROM = 5.431
DIAMOND_FRAC = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                         [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])


def mp_mesh(recip, mesh):
    """Unshifted Monkhorst-Pack points i/n (Cartesian), equal weights; without symmetry reduction
    (what remains of the mesh for a structure with only the identity, kpoints/mesh.zig:5-47)."""
    m1, m2, m3 = mesh
    pts = [np.array([i / m1, j / m2, k / m3]) @ recip for i in range(m1) for j in range(m2) for k in range(m3)]
    return pts, [1.0 / (m1 * m2 * m3)] * len(pts)


def silicon(cells=(1, 1, 1), ecut_ry=20.0, kmesh=(1, 1, 1), nbands=None, displacement_angstrom=0.0,
            seed=64, grid=None, kpoints=None) -> ScfSystem:
    """Diamond Si in n1 x n2 x n3 conventional cubic cells (8 atoms each, a = 5.431 A), atoms
    optionally displaced by N(0, displacement) (numpy PCG64(seed)); k-points from an unshifted
    MP mesh unless given explicitly as (Cartesian points, weights)."""
    a = SI_A_ANGSTROM * tables.ANGSTROM_TO_BOHR
    n1, n2, n3 = cells
    cell = np.diag([a * n1, a * n2, a * n3])
    recip = tables.reciprocal(cell)
    pos = np.array([(f + np.array([i, j, k])) * a for i in range(n1) for j in range(n2) for k in range(n3)
                    for f in DIAMOND_FRAC])
    if displacement_angstrom > 0:
        rng = np.random.Generator(np.random.PCG64(seed))
        pos = pos + rng.normal(0.0, displacement_angstrom * tables.ANGSTROM_TO_BOHR, size=pos.shape)
    g = SI_GTH
    sp = SpeciesDesc([0, 1], [g["r_s"], g["r_p"]], np.diag([2.0 * g["h_s"], 2.0 * g["h_p"]]))
    nat = pos.shape[0]
    nelec = g["z"] * nat
    kpts, kw = kpoints if kpoints is not None else mp_mesh(recip, kmesh)
    grid = tuple(grid) if grid is not None else tuple(tables.auto_grid(ecut_ry, recip))
    nb = nbands or int(math.ceil(nelec / 2.0))
    return ScfSystem(f"si{nat}", cell, recip, abs(float(np.linalg.det(cell))), grid, ecut_ry, pos,
                     np.zeros(nat, dtype=np.int32), [sp], g["z"], g["rloc"], g["c1"], nelec, nb,
                     [np.asarray(k, dtype=np.float64) for k in kpts], list(kw))


def mp_mesh_time_reversal(recip, mesh):
    """Unshifted Monkhorst-Pack mesh with k and -k merged (time reversal, the only symmetry used
    here; kpoints/kpoints.zig:30-93 also applies the crystal point group)."""
    m = np.array(mesh)
    seen, pts, wts = {}, [], []
    for i in range(m[0]):
        for j in range(m[1]):
            for k in range(m[2]):
                key = (i, j, k)
                neg = tuple(int(v) for v in (-np.array(key)) % m)
                if neg in seen:
                    wts[seen[neg]] += 1.0
                    continue
                seen[key] = len(pts)
                pts.append(np.array(key) / m @ recip)
                wts.append(1.0)
    tot = float(np.prod(m))
    return pts, [w / tot for w in wts]


def lattice_symmetry_ops(cell, atom_frac, atom_species, tol=1e-6):
    """Integer matrices W (fractional coordinates, rows x -> x W + t) of the space-group operations
    of the structure: lattice automorphisms (W g W^T = g, g = cell cell^T, entries in {-1, 0, 1})
    that map the atoms onto atoms of the same species with some translation t (the role of
    symmetry.get_symmetry_ops, src/features/symmetry/symmetry.zig)."""
    import itertools

    cell = np.asarray(cell, dtype=np.float64)
    g = cell @ cell.T
    frac = np.asarray(atom_frac, dtype=np.float64) % 1.0
    spc = np.asarray(atom_species)
    ops = []
    for entries in itertools.product((-1, 0, 1), repeat=9):
        w = np.array(entries, dtype=np.int64).reshape(3, 3)
        if abs(round(np.linalg.det(w))) != 1 or not np.allclose(w @ g @ w.T, g, atol=tol * np.max(np.abs(g))):
            continue
        mapped = frac @ w
        found = False
        for b in np.nonzero(spc == spc[0])[0]:
            t = frac[b] - mapped[0]
            img = (mapped + t) % 1.0
            ok = True
            for a in range(frac.shape[0]):
                d = (img[a] - frac) % 1.0
                d = np.minimum(d, 1.0 - d)
                hit = np.nonzero((np.max(d, axis=1) < tol) & (spc == spc[a]))[0]
                if hit.size == 0:
                    ok = False
                    break
            if ok:
                found = True
                break
        if found:
            ops.append(w)
    return ops


def ibz_kmesh(cell, mesh, atom_frac, atom_species, time_reversal=True):
    """Unshifted Monkhorst-Pack mesh reduced to its irreducible part, as reduction.reduce_kmesh
    (src/features/kpoints/reduction.zig:55-120): mesh points in flat order (x index fastest), each
    not yet covered starts an orbit under the k-space images of the symmetry operations (k W^T on
    the integer indices, modulo the mesh) and time reversal; weight = orbit size / mesh size; the
    representative is the seed, k_frac = i / n wrapped to [-1/2, 1/2) (mesh.zig:39-47).
    Returns (Cartesian points, weights, fractional points)."""
    recip = tables.reciprocal(np.asarray(cell, dtype=np.float64))
    n = np.array(mesh, dtype=np.int64)
    ops = lattice_symmetry_ops(cell, atom_frac, atom_species)
    total = int(np.prod(n))
    used = np.zeros(total, dtype=bool)

    def flat(idx):
        idx = idx % n
        return int(idx[0] + n[0] * (idx[1] + n[1] * idx[2]))

    pts, wts, fr = [], [], []
    for f in range(total):
        if used[f]:
            continue
        seed = np.array([f % n[0], (f // n[0]) % n[1], f // (n[0] * n[1])], dtype=np.int64)
        group = {f}
        for w in ops:
            img = seed @ w.T
            group.add(flat(img))
            if time_reversal:
                group.add(flat(-img))
        for q in group:
            used[q] = True
        kf = seed / n
        kf = kf - np.floor(kf + 0.5)
        pts.append(kf @ recip)
        wts.append(len(group) / total)
        fr.append(kf)
    return pts, wts, fr


AL_GTH = {"z": 3.0, "rloc": 0.45, "c1": -8.491351, "r_s": 0.460104, "h_s": 5.088340, "r_p": 0.536744,
          "h_p": 2.193438}


def aluminium(ecut_ry=30.0, kmesh=(4, 4, 4), nbands=12, smearing_ry=0.01, grid=None, displace=None,
              xc="lda_pz", symmetry=False) -> ScfSystem:
    """fcc Al in the 4-atom cubic cell (a = 4.05 A), GTH-form local part and s/p Gaussian
    projectors (Goedecker-Teter-Hutter Al LDA parameters), Fermi-Dirac smearing; k-points from an
    unshifted mesh reduced by time reversal (4x4x4 -> 36 points with uneven plane-wave counts), or
    with symmetry=True by the structure's space group and time reversal (8x8x8 -> 35 points).

    `displace` (bohr, 3-vector) moves the second atom off its site.  In the perfect crystal,
    partially occupied levels are symmetry-degenerate, and the density then depends at the 1e-9
    level on how an eigensolver rotates each degenerate set (occupations of numerically split
    copies differ by ~1e-7), which leaves any SCF at a noise floor of drho ~ 1e-9.  A small
    displacement lifts the degeneracies so the fixed point is defined to round-off."""
    a = 4.05 * tables.ANGSTROM_TO_BOHR
    cell = np.diag([a, a, a])
    recip = tables.reciprocal(cell)
    pos = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]]) * a
    if displace is not None:
        pos[1] += np.asarray(displace, dtype=np.float64)
    g = AL_GTH
    sp = SpeciesDesc([0, 1], [g["r_s"], g["r_p"]], np.diag([2.0 * g["h_s"], 2.0 * g["h_p"]]))
    if symmetry:
        kpts, kw, _ = ibz_kmesh(cell, kmesh, pos @ np.linalg.inv(cell), np.zeros(4, np.int32))
    else:
        kpts, kw = mp_mesh_time_reversal(recip, kmesh)
    grid = tuple(grid) if grid is not None else tuple(tables.auto_grid(ecut_ry, recip))
    return ScfSystem("al4", cell, recip, abs(float(np.linalg.det(cell))), grid, ecut_ry, pos, np.zeros(4, np.int32),
                     [sp], g["z"], g["rloc"], g["c1"], 12.0, nbands, kpts, kw, smearing_ry, xc=xc)


def si8_config(ecut_ry=20.0, nbands=20):
    """The 8-atom Si cell at its stated configuration: LDA, 20 Ry, auto grid (32^3), the unshifted
    2x2x2 mesh of the cubic cell reduced by the structure's operations and time reversal to Gamma,
    X, M and R with weights 1, 3, 3, 1 (/8)."""
    s = silicon((1, 1, 1), ecut_ry, nbands=nbands)
    kpts, kw, _ = ibz_kmesh(s.cell, (2, 2, 2), s.atom_pos @ np.linalg.inv(s.cell), s.atom_species)
    import dataclasses

    return dataclasses.replace(s, kpoints=[np.asarray(k, dtype=np.float64) for k in kpts], kweights=list(kw))


def si8_small(ecut_ry=8.0, nbands=20):
    """Small SCF parity case: the 8-atom cubic cell at 8 Ry (24^3 grid, about 420 plane waves),
    k = Gamma and (1/4, 1/4, 1/4) in reciprocal-lattice units, weights 1/2 each."""
    a = SI_A_ANGSTROM * tables.ANGSTROM_TO_BOHR
    recip = tables.reciprocal(np.diag([a, a, a]))
    kpts = [np.zeros(3), np.array([0.25, 0.25, 0.25]) @ recip]
    return silicon((1, 1, 1), ecut_ry, nbands=nbands, kpoints=(kpts, [0.5, 0.5]))


AL_SMALL_DISPLACEMENT = (0.11, 0.07, 0.05)  # bohr


def al4_small(ecut_ry=8.0, nbands=10, xc="lda_pz"):
    """Small smearing parity case: the 4-atom Al cell at 8 Ry (15^3 grid), 2x2x2 mesh reduced by
    time reversal (8 k-points with different plane-wave counts), second atom displaced by
    AL_SMALL_DISPLACEMENT so no partially occupied level is degenerate."""
    return aluminium(ecut_ry=ecut_ry, kmesh=(2, 2, 2), nbands=nbands, displace=AL_SMALL_DISPLACEMENT, xc=xc)


def with_upf(system, upfs):
    """The system with every species read from a UPF file (dzhip.upf.Upf, one per species): the
    projector tables are the files' radial tables over [0, sqrt(2 ecut) 1.5] (common.zig:124-147,
    RadialTable 2048 points, nonlocal.zig:250-300), D from PP_DIJ, V_loc(q) the files' short-range
    local form factor (form_factor.zig:225-256), Z their z_valence."""
    import dataclasses

    g_max = math.sqrt(2.0 * system.ecut_ry) * 1.5
    return dataclasses.replace(system, species=[upf.UpfSpecies(u, g_max) for u in upfs],
                               species_local=[(u.z_valence, 0.0, 0.0) for u in upfs], species_upf=list(upfs))


def si64(ecut_ry=30.0, nbands=160):
    """2x2x2 supercell of the cubic cell, atoms displaced by N(0, 0.02 A), unshifted 2x2x2 mesh
    (8 k-points: the displacements leave only the identity), 160 bands."""
    return silicon((2, 2, 2), ecut_ry, (2, 2, 2), nbands, displacement_angstrom=0.02, seed=64)


# GTH-form parameters of Ga (3 valence electrons) and As (5) for the zincblende cell (Hartree units
# for C1 and h): local radius of the Hartwigsen-Goedecker-Hutter LDA sets with the leading s and p
# projector of each (radius and strength); the deeper projectors and the d channel of those sets
# are left out, which keeps the structure a gapped semiconductor with 4 bonds per atom.
GA_GTH = {"z": 3.0, "rloc": 0.56, "c1": 0.0, "r_s": 0.610791, "h_s": 2.369325, "r_p": 0.704596, "h_p": 0.746305}
AS_GTH = {"z": 5.0, "rloc": 0.52, "c1": 0.0, "r_s": 0.456400, "h_s": 4.560761, "r_p": 0.550562, "h_p": 1.812247}
GAAS_A_ANGSTROM = 5.653


def gaas(cells=(2, 2, 2), ecut_ry=35.0, kpoints=None, nbands=160, displacement_angstrom=0.05, seed=5,
         grid=None, smearing_ry=0.01) -> ScfSystem:
    """Zincblende GaAs in n1 x n2 x n3 conventional cubic cells (Ga on the fcc sites, As at
    (1/4, 1/4, 1/4) + fcc; a = 5.653 A, so the 2x2x2 supercell has a = 11.306 A and 64 atoms), atoms
    displaced by N(0, displacement) (numpy PCG64(seed)), Gamma only unless k-points are given.

    Fermi-Dirac smearing (smearing_ry; 0 = fixed occupations): with these GTH-form parameters the
    64-atom cell at Gamma has no gap at the Fermi level (levels 126-128 lie within 1.5e-3 Ry and
    levels 128 and 129 within 2e-4 Ry), so with fixed occupations the highest occupied level
    changes between SCF iterations and the density never settles."""
    a = GAAS_A_ANGSTROM * tables.ANGSTROM_TO_BOHR
    n1, n2, n3 = cells
    cell = np.diag([a * n1, a * n2, a * n3])
    recip = tables.reciprocal(cell)
    fcc = DIAMOND_FRAC[:4]
    pos, spc = [], []
    for i in range(n1):
        for j in range(n2):
            for k in range(n3):
                for f in fcc:
                    pos.append((f + np.array([i, j, k])) * a)
                    spc.append(0)
                for f in fcc:
                    pos.append((f + 0.25 + np.array([i, j, k])) * a)
                    spc.append(1)
    pos = np.array(pos)
    if displacement_angstrom > 0:
        rng = np.random.Generator(np.random.PCG64(seed))
        pos = pos + rng.normal(0.0, displacement_angstrom * tables.ANGSTROM_TO_BOHR, size=pos.shape)
    species, local = [], []
    for g in (GA_GTH, AS_GTH):
        species.append(SpeciesDesc([0, 1], [g["r_s"], g["r_p"]], np.diag([2.0 * g["h_s"], 2.0 * g["h_p"]])))
        local.append((g["z"], g["rloc"], g["c1"]))
    spc = np.array(spc, dtype=np.int32)
    nelec = float(sum(local[s][0] for s in spc))
    kpts, kw = kpoints if kpoints is not None else ([np.zeros(3)], [1.0])
    grid = tuple(grid) if grid is not None else tuple(tables.auto_grid(ecut_ry, recip))
    return ScfSystem(f"gaas{pos.shape[0]}", cell, recip, abs(float(np.linalg.det(cell))), grid, ecut_ry, pos, spc,
                     species, 0.0, 0.0, 0.0, nelec, nbands, [np.asarray(k, dtype=np.float64) for k in kpts], list(kw),
                     smearing_ry, species_local=local)


def gaas64(ecut_ry=35.0, nbands=160, smearing_ry=0.01):
    """The 64-atom GaAs supercell (2x2x2 cubic cells, a = 11.306 A, displacements N(0, 0.05 A)),
    35 Ry (90^3 grid), Gamma point, 160 bands, Fermi-Dirac smearing 0.01 Ry."""
    return gaas((2, 2, 2), ecut_ry, nbands=nbands, smearing_ry=smearing_ry)


def start_potential(system, potential="atomic"):
    """local_r of the first SCF iteration: V_ion + V_Hxc[rho_atomic] ("atomic") or V_ion ("ion")."""
    s = system
    g, g2 = g_vectors(s.grid, s.recip)
    local_r = ionic_potential(s, g, g2)
    if potential == "atomic":
        rho = atomic_density(s, g, g2)
        local_r = local_r + hartree(rho, s.grid, g2, s.volume)[0] + lda_pz(rho)[1]
    return local_r


def kpoint_workload(system, k_index=0, nbands=None, seed=20261203, potential="atomic", local_r=None):
    """One k-point of an SCF system as an H.psi workload: its G-sphere, the SCF system's projectors
    and atoms, local_r = V_ion + V_Hxc[rho_atomic] (the first SCF iteration's input potential;
    potential="ion" keeps V_ion only; a given local_r is used as is) and a random psi block with
    coefficients damped by 1 / (1 + |k+G|^2), seeded."""
    from .workloads import Workload, _random_psi

    s = system
    k = np.asarray(s.kpoints[k_index], dtype=np.float64)
    basis = tables.generate_basis(s.recip, s.ecut_ry, k)
    if local_r is None:
        local_r = start_potential(s, potential)
    rng = np.random.Generator(np.random.PCG64(seed))
    psi = _random_psi(rng, nbands or s.nbands, basis.kinetic)
    return Workload(f"{s.name}_k{k_index}", s.cell, s.recip, s.volume, tuple(s.grid), s.ecut_ry, k, basis, local_r,
                    list(s.species), s.atom_pos, s.atom_species, psi,
                    {"seed": seed, "k_index": k_index, "potential": potential})


# ----------------------------------------------------------------- grid potentials (host, FP64)
def _freqs(n):
    i = np.arange(n)
    return np.where(i <= (n - 1) // 2, i, i - n)


def g_vectors(grid, recip):
    """Cartesian G and |G|^2 on the FFT-order grid (index x + nx (y + ny z), wrapped frequencies)."""
    nx, ny, nz = grid
    fz, fy, fx = np.meshgrid(_freqs(nz), _freqs(ny), _freqs(nx), indexing="ij")
    g = (fx.reshape(-1, 1) * recip[0] + fy.reshape(-1, 1) * recip[1] + fz.reshape(-1, 1) * recip[2])
    return g, np.einsum("ij,ij->i", g, g)


def to_g(v_r, grid):
    nx, ny, nz = grid
    return np.fft.fftn(v_r.reshape(nz, ny, nx)).reshape(-1)


def to_r(v_g, grid):
    nx, ny, nz = grid
    return np.fft.ifftn(v_g.reshape(nz, ny, nx)).reshape(-1)


def lda_pz(rho):
    """Energy density f (Ry per volume) and potential df/drho (Ry), Perdew-Zunger 1981 fit of
    Ceperley-Alder correlation plus Slater exchange; zero where rho <= 1e-12."""
    rho = np.asarray(rho, dtype=np.float64)
    f = np.zeros_like(rho)
    v = np.zeros_like(rho)
    on = rho > 1e-12
    n = rho[on]
    cbrt_n = np.cbrt(n)
    cx = -0.75 * (3.0 / math.pi) ** (1.0 / 3.0)
    rs = np.cbrt(3.0 / (4.0 * math.pi * n))
    dense = rs < 1.0
    ec = np.empty_like(n)
    vc = np.empty_like(n)
    r = rs[dense]
    ln = np.log(r)
    ec[dense] = 0.0311 * ln - 0.048 + 0.0020 * r * ln - 0.0116 * r
    vc[dense] = ec[dense] - (r / 3.0) * (0.0311 / r + 0.0020 * ln + 0.0020 - 0.0116)
    r = rs[~dense]
    sq = np.sqrt(r)
    den = 1.0 + 1.0529 * sq + 0.3334 * r
    ec[~dense] = -0.1423 / den
    vc[~dense] = ec[~dense] - (r / 3.0) * (0.1423 * (0.5 * 1.0529 / sq + 0.3334) / (den * den))
    f[on] = 2.0 * n * (cx * cbrt_n + ec)
    v[on] = 2.0 * (4.0 / 3.0 * cx * cbrt_n + vc)
    return f, v


def hartree(rho, grid, g2, volume):
    """V_H(r) (Ry, V(G) = 8 pi rho(G)/G^2, G = 0 dropped) and E_H = (1/2) int V_H rho."""
    n = rho.shape[0]
    rg = to_g(rho, grid)
    safe = np.where(g2 > 1e-12, g2, 1.0)
    vh = to_r(np.where(g2 > 1e-12, rg * (8.0 * math.pi) / safe, 0.0), grid).real
    return vh, 0.5 * float(np.dot(vh, rho)) * volume / n


def ionic_potential(system, g, g2):
    """V_ion(r) = sum_G e^{iG.r} (1/Omega) sum_s v_s(|G|) sum_{a in s} e^{-iG.tau_a}, v_s of the GTH
    local form of species s."""
    s = system
    total = np.zeros(g.shape[0], dtype=np.complex128)
    nz = g2 > 1e-20
    q2 = g2[nz]
    for sp, (z, rloc, c1) in enumerate(s.local_params()):
        sf = np.zeros(g.shape[0], dtype=np.complex128)
        for tau in s.atom_pos[np.asarray(s.atom_species) == sp]:
            sf += np.exp(-1j * (g @ tau))
        vq = np.zeros(g.shape[0])
        if s.species_upf is not None:
            # The file's local form factor at each distinct |G| (hamiltonian.zig:449-473).
            qs, inv = np.unique(np.sqrt(q2), return_inverse=True)
            vq[nz] = upf.local_form_factor(s.species_upf[sp], qs)[inv]
        else:
            vq[nz] = 2.0 * (-4.0 * math.pi * z / q2 + c1 * (2.0 * math.pi) ** 1.5 * rloc ** 3) * np.exp(-0.5 * q2 * rloc ** 2)
        total += sf * vq
    return to_r(total * (g.shape[0] / s.volume), s.grid).real


def atomic_density(system, g, g2):
    """Superposition of atomic valence charges (core_density.zig:128-178): rho(G) =
    (1/Omega) sum_a Z_a e^{-G^2 a^2 / 2} e^{-iG.tau_a}, a Gaussian of width a per atom (the
    synthetic species carry no PP_RHOATOM), so int rho = nelec."""
    s = system
    zs = [p[0] for p in s.local_params()]
    sf = np.zeros(g.shape[0], dtype=np.complex128)
    for tau, sp in zip(s.atom_pos, s.atom_species):
        sf += zs[int(sp)] * np.exp(-1j * (g @ tau))
    rho_g = np.exp(-0.5 * g2 * s.atom_rho_width ** 2) * sf / s.volume
    return to_r(rho_g * g.shape[0], s.grid).real


def band_weights(nelec, nbands, kweight):
    """w_b = w_k occ_b 2: the first ceil(nelec/2) bands are occupied, the last one partially."""
    nocc = int(math.ceil(nelec / 2.0))
    w = np.zeros(nbands)
    for b in range(min(nocc, nbands)):
        w[b] = 2.0 * kweight * ((nelec / 2.0 - (nocc - 1)) if b + 1 == nocc else 1.0)
    return w


def fermi_dirac(e, mu, sigma):
    """Occupation 1/(1 + e^{(e - mu)/sigma}) (fermi_level.zig:7-11)."""
    x = np.clip((np.asarray(e, dtype=np.float64) - mu) / sigma, -700.0, 700.0)
    return 1.0 / (1.0 + np.exp(x))


def fd_entropy(f):
    """-(f ln f + (1 - f) ln(1 - f)), zero within 1e-12 of 0 and 1 (fermi_level.zig:15-19)."""
    f = np.asarray(f, dtype=np.float64)
    inner = (f > 1e-12) & (f < 1.0 - 1e-12)
    g = np.where(inner, f, 0.5)
    return np.where(inner, -(g * np.log(g) + (1.0 - g) * np.log1p(-g)), 0.0)


def fermi_level(nelec, sigma, eigenvalues, kweights):
    """mu with sum_k 2 w_k sum_b f(e_kb) = nelec by bisection on a padded bracket (pad
    max(10 sigma, 1e-3), doubled up to 12 times), 80 halvings, |dN| < 1e-8 stops
    (fermi_level.zig:54-124)."""
    ev = np.concatenate([np.asarray(v, dtype=np.float64) for v in eigenvalues])
    wt = np.concatenate([np.full(len(v), 2.0 * w) for v, w in zip(eigenvalues, kweights)])

    def count(mu):
        return float(np.dot(wt, fermi_dirac(ev, mu, sigma)))

    e0, e1 = float(ev.min()), float(ev.max())
    pad = max(10.0 * sigma, 1e-3)
    for _ in range(13):
        lo, hi = e0 - pad, e1 + pad
        if count(lo) <= nelec <= count(hi):
            break
        pad *= 2.0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        c = count(mid)
        if abs(c - nelec) < 1e-8:
            return mid
        if c > nelec:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def occupations(system, eigenvalues):
    """Per-k band weights w_k occ 2 and the -TS energy (Ry) for fixed or Fermi-Dirac occupations."""
    s = system
    if s.smearing <= 0.0:
        return [band_weights(s.nelec, len(v), wk) for v, wk in zip(eigenvalues, s.kweights)], 0.0, None
    mu = fermi_level(s.nelec, s.smearing, eigenvalues, s.kweights)
    ws, ts = [], 0.0
    for v, wk in zip(eigenvalues, s.kweights):
        f = fermi_dirac(v, mu, s.smearing)
        ws.append(2.0 * wk * f)
        ts += float(np.sum(2.0 * wk * s.smearing * fd_entropy(f)))
    return ws, ts, mu


def model_dielectric(resid, grid, g2, diemac, dielng):
    """P(G) R(G), P = (dielng^2 G^2 + 1/diemac) / (dielng^2 G^2 + 1), P(0) = 1/diemac: the model
    dielectric preconditioner applied to the Pulay residual (mixing.zig:309-338, scf.zig:164-178)."""
    x = dielng * dielng * g2
    p = np.where(g2 < 1e-12, 1.0 / diemac, (x + 1.0 / diemac) / (x + 1.0))
    return to_r(to_g(resid, grid) * p, grid).real


class Pulay:
    """DIIS on (V_in, R = V_out - V_in): keep the last `history` pairs, linear step while fewer
    than two, else V = sum_i c_i (V_i + beta R_i) with sum_i c_i = 1 minimising |sum_i c_i R_i|
    (mixing.zig:183-234); a singular system gives equal weights (mixing.zig:340-400).

    Two deliberate departures keep the iteration converging below 1e-6 Ry residuals.  The
    overlap matrix is divided by its largest diagonal entry before the elimination: the
    reference tests its pivots against an absolute 1e-12, so once <R|R> (its G-space norm, the
    real-space sum / N) is below about 1e-12 it takes the equal-weights branch on every step.
    And a singular system keeps only the newest pair and takes the linear step, where the
    reference averages the whole history.  Whenever the reference's system is regular the
    coefficients are the same (they are invariant under a scaling of the matrix)."""

    def __init__(self, history: int):
        self.history = history
        self.hist_v: list = []
        self.hist_r: list = []

    def step(self, v_in, resid, beta):
        self.hist_v.append(np.array(v_in))
        self.hist_r.append(np.array(resid))
        if len(self.hist_v) > self.history:
            del self.hist_v[0], self.hist_r[0]
        m = len(self.hist_v)
        if m < 2:
            return v_in + beta * resid
        r = np.stack(self.hist_r)
        a = np.zeros((m + 1, m + 1))
        a[:m, :m] = r @ r.T
        a[:m, :m] /= max(float(np.max(np.diag(a)[:m])), 1e-300)
        a[:m, m] = -1.0
        a[m, :m] = -1.0
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        c = _gauss_solve(a, rhs)
        if c is None:
            del self.hist_v[:-1], self.hist_r[:-1]
            return v_in + beta * resid
        out = np.zeros_like(v_in)
        for i in range(m):
            out += c[i] * (self.hist_v[i] + beta * self.hist_r[i])
        return out


def _gauss_solve(a, b):
    """Gaussian elimination with partial pivoting; None when a pivot is below 1e-12."""
    a = a.copy()
    b = b.copy()
    size = a.shape[0]
    for col in range(size):
        p = col + int(np.argmax(np.abs(a[col:, col])))
        if p != col:
            a[[col, p]] = a[[p, col]]
            b[[col, p]] = b[[p, col]]
        if abs(a[col, col]) < 1e-12:
            return None
        f = a[col + 1:, col] / a[col, col]
        a[col + 1:, col:] -= np.outer(f, a[col, col:])
        b[col + 1:] -= f * b[col]
    x = np.zeros(size)
    for i in range(size - 1, -1, -1):
        x[i] = (b[i] - a[i, i + 1:] @ x[i + 1:]) / a[i, i]
    return x


# ----------------------------------------------------------------- driver
@dataclass
class ScfRecord:
    iteration: int
    wall_s: float  # since the start of the first iteration
    energy_ry: float
    band_ry: float
    drho: float
    eig_iterations: int
    applies: int
    fp32_applies: int
    eig_s: float  # wall time inside the k-point solves and densities
    sections_ms: dict = field(default_factory=dict)  # solver GPU time per section, summed over k


@dataclass
class ScfRun:
    records: list = field(default_factory=list)
    eigenvalues: list = field(default_factory=list)
    rho: np.ndarray | None = None
    v_in: np.ndarray | None = None
    converged: bool = False
    setup_s: float = 0.0

    @property
    def energies(self):
        return [r.energy_ry for r in self.records]

    def time_to(self, e_ref_ry: float, tol_ha: float = 1e-6):
        """Wall time to the end of the first iteration after which every iteration's total energy
        stays within tol (Ha) of e_ref; None if the run never gets there."""
        first = None
        for i, r in enumerate(self.records):
            if abs(r.energy_ry - e_ref_ry) * HA_PER_RY <= tol_ha:
                first = i if first is None else first
            else:
                first = None
        return None if first is None else self.records[first].wall_s


class ScfDriver:
    """One Hamiltonian context per k-point, built once and kept across SCF iterations."""

    def __init__(self, system: ScfSystem, concurrent_k: bool = False, device: int = 0, device_grid: bool = False,
                 pulay_history: int = 8):
        """device_grid keeps rho, V and the Pulay history on the GPU (dzhip_scf_grid); otherwise
        potentials and mixing run on the host in FP64 numpy."""
        t0 = time.perf_counter()
        s = self.sys = system
        if s.xc != "lda_pz" and not device_grid:
            raise ValueError(f"xc = {s.xc!r} runs on the device grid (device_grid=True)")
        self.concurrent = concurrent_k and len(s.kpoints) > 1
        self.ctxs = [Context(device) for _ in s.kpoints] if self.concurrent else [Context(device)]
        self.g, self.g2 = g_vectors(s.grid, s.recip)
        self.v_ion = ionic_potential(s, self.g, self.g2)
        zero = np.zeros(int(np.prod(s.grid)))
        self.hams = []
        for i, kp in enumerate(s.kpoints):
            b = tables.generate_basis(s.recip, s.ecut_ry, kp)
            species = [{"l_list": sp.l_list, "dij": sp.dij,
                        "phi": tables.build_phi(b.kpg, sp.l_list, sp.radial_fns())} for sp in s.species]
            ctx = self.ctxs[i if self.concurrent else 0]
            self.hams.append(Hamiltonian(ctx, s.grid, tables.fft_indices(b, s.grid), b.kinetic, b.cart, zero,
                                         1.0 / s.volume, species, s.atom_pos, s.atom_species))
        self.npw = [h.n for h in self.hams]
        self.grid = None
        if device_grid:
            self.grid = ScfGrid(self.ctxs[0], s.grid, s.recip, s.volume, pulay_history)
            self.grid.set_ion(self.v_ion)
            self.grid.set_xc(s.xc)
        self.pool = ThreadPoolExecutor(len(self.hams)) if self.concurrent else None
        # Base preconditioner of the installed learned model (set_precond_model) and an optional
        # callback hook(k, ham, stats) after every k-point solve.
        self.mlp_base = PRECOND_TPA
        self.solve_hook = None
        self.setup_s = time.perf_counter() - t0

    def set_precond_model(self, params, energy_scale, base=PRECOND_TPA):
        """Install a learned preconditioner model on every k-point's context."""
        for h in self.hams:
            h.set_precond_model(params, energy_scale, base)
        self.mlp_base = base

    def _solve_k(self, k, local_r, nb, opts):
        ham = self.hams[k]
        if self.grid is not None:
            self.grid.set_local(ham)
        else:
            ham.set_local(local_r)
        vals, _, st = ham.eig_solve(nb, want_vectors=False, reuse=True, **opts)
        if self.solve_hook is not None:
            self.solve_hook(k, ham, st)
        return vals, st

    def _solve_all(self, local_r, opts):
        """Every k-point's eigenproblem (concurrently when each has its own context), then the
        occupations (Fermi level when smearing) and the densities of the device solutions."""
        s = self.sys
        ks = range(len(self.hams))
        if self.pool is not None:
            res = list(self.pool.map(lambda k: self._solve_k(k, local_r, s.nbands, opts), ks))
        else:
            res = [self._solve_k(k, local_r, s.nbands, opts) for k in ks]
        vals = [r[0] for r in res]
        weights, ts, mu = occupations(s, vals)
        rho = None
        for ham, w in zip(self.hams, weights):
            if self.grid is not None:
                self.grid.add_density(ham, w)
            else:
                rho = ham.density_solution(w, rho=rho)
        e_band = float(sum(np.dot(w, v) for w, v in zip(weights, vals)))
        return vals, [r[1] for r in res], rho, e_band, ts

    def _run_device(self, max_iter, beta, conv, pulay_history, pulay_start, warmup_steps, solve_opts, progress,
                    energy_tol_ry, rho0=None):
        """The SCF loop with rho, V and mixing on the device (same algorithm as run())."""
        s = self.sys
        g = self.grid
        if rho0 is None:
            g.init(s.nelec)
        else:
            g.init_rho(rho0)
        run = ScfRun(setup_s=self.setup_s)
        t0 = time.perf_counter()
        drho_prev = None
        for it in range(max_iter):
            opts = solve_opts(it, drho_prev)
            te = time.perf_counter()
            vals, stats, _, e_band, ts = self._solve_all(None, opts)
            eig_s = time.perf_counter() - te
            e_h, e_xc, vin_rho, drho = g.hxc()
            e_tot = e_band - vin_rho + e_h + e_xc - ts
            run.records.append(self._record(it, e_tot, e_band, drho, stats, eig_s))
            run.eigenvalues = vals
            drho_prev = drho
            done = drho < conv
            if energy_tol_ry is not None and it >= warmup_steps and len(run.records) > 1:
                done = done or abs(run.records[-1].energy_ry - run.records[-2].energy_ry) < energy_tol_ry
            if not done:
                g.mix(beta, pulay_history > 0 and it >= pulay_start)
            run.records[-1].wall_s = time.perf_counter() - t0
            if progress is not None:
                progress(run.records[-1])
            if done:
                run.converged = True
                break
        run.rho = g.read(0)
        run.v_in = g.read(1)
        return run

    def run(self, max_iter=60, beta=0.3, conv=1e-9, precision=FP64, preconditioner=PRECOND_KINETIC,
            pulay_history=8, pulay_start=4, eig_tol=1e-4, eig_max_iter=20, warmup_steps=2, warmup_tol=1e-3,
            warmup_max_iter=10, mixed_switch_tol=0.0, final_tol=None, energy_tol_ry=None, eig_tol_ratio=None,
            eig_tol_min=1e-8, progress=None, diemac=1.0, dielng=1.0, max_subspace=0, init_diagonal=False,
            nconv=0, start="uniform", collect_snapshots=0):
        """SCF from the uniform start density nelec/Omega (or start="atomic": a superposition of
        atomic valence charges).  Stops when the RMS density change is
        below conv, or (with energy_tol_ry) when two successive total energies after the warm-up
        differ by less than energy_tol_ry.

        eig_tol_ratio (None = the reference's fixed eig_tol) tightens the eigensolver tolerance with
        the SCF: tol = min(eig_tol, max(eig_tol_min, eig_tol_ratio * drho of the previous step)).
        collect_snapshots > 0 makes every solve record that many snapshots for training samples
        (read them in solve_hook).
        A fixed 1e-4 Ry residual leaves band-energy errors of order N_b r^2 / gap, about 1e-5 Ry
        for 128 bands across a 0.08 Ry gap, which a converged total energy cannot carry.
        progress, if given, is called with each ScfRecord."""
        s = self.sys

        def solve_opts(it, drho_prev):
            warm = it < warmup_steps
            tol = warmup_tol if warm else eig_tol
            if not warm and eig_tol_ratio is not None and drho_prev is not None:
                tol = min(eig_tol, max(eig_tol_min, eig_tol_ratio * drho_prev))
            # The learned preconditioner's fallback compares against its base preconditioner's
            # convergence rate on each context, so the warm-up solves run the base and measure it.
            pc = self.mlp_base if (preconditioner == PRECOND_MLP and warm) else preconditioner
            return dict(tol=tol, max_iter=warmup_max_iter if warm else eig_max_iter, precision=precision,
                        preconditioner=pc, mixed_switch_tol=mixed_switch_tol,
                        max_subspace=max_subspace, init_diagonal=init_diagonal and it == 0, nconv=nconv,
                        collect_snapshots=collect_snapshots)

        rho0 = atomic_density(s, self.g, self.g2) if start == "atomic" else None
        if self.concurrent:
            self._pretune(precision)
        if self.grid is not None:
            self.grid.set_dielectric(diemac, dielng)
            run = self._run_device(max_iter, beta, conv, pulay_history, pulay_start, warmup_steps, solve_opts,
                                   progress, energy_tol_ry, rho0)
            self._final(run, final_tol, preconditioner)
            return run
        n = int(np.prod(s.grid))
        vol = s.volume
        rho = np.full(n, s.nelec / vol) if rho0 is None else rho0
        v_in = hartree(rho, s.grid, self.g2, vol)[0] + lda_pz(rho)[1]
        mixer = Pulay(pulay_history) if pulay_history > 0 else None
        run = ScfRun(setup_s=self.setup_s)
        t0 = time.perf_counter()
        drho_prev = None
        for it in range(max_iter):
            warm = it < warmup_steps
            opts = solve_opts(it, drho_prev)
            local_r = self.v_ion + v_in
            te = time.perf_counter()
            vals, stats, rho_out, e_band, ts = self._solve_all(local_r, opts)
            eig_s = time.perf_counter() - te
            vh_out, e_h = hartree(rho_out, s.grid, self.g2, vol)
            f_xc, vxc_out = lda_pz(rho_out)
            e_xc = float(np.sum(f_xc)) * vol / n
            # Kohn-Sham energy of rho_out at the input potential (equals the reference's
            # band + double counting at self-consistency).
            e_tot = e_band - float(np.dot(v_in, rho_out)) * vol / n + e_h + e_xc - ts
            drho = math.sqrt(float(np.mean((rho_out - rho) ** 2)))
            run.records.append(self._record(it, e_tot, e_band, drho, stats, eig_s))
            run.eigenvalues = vals
            rho = rho_out
            drho_prev = drho
            done = drho < conv
            if energy_tol_ry is not None and not warm and len(run.records) > 1:
                done = done or abs(run.records[-1].energy_ry - run.records[-2].energy_ry) < energy_tol_ry
            if not done:
                v_out = vh_out + vxc_out
                if mixer is not None and it >= pulay_start:
                    resid = v_out - v_in
                    if diemac > 1.0:
                        resid = model_dielectric(resid, s.grid, self.g2, diemac, dielng)
                    v_in = mixer.step(v_in, resid, beta)
                else:
                    v_in = v_in + beta * (v_out - v_in)
            run.records[-1].wall_s = time.perf_counter() - t0
            if progress is not None:
                progress(run.records[-1])
            if done:
                run.converged = True
                break
        run.rho = rho
        run.v_in = v_in
        self._final(run, final_tol, preconditioner)
        return run

    def _pretune(self, precision):
        """Tune the x-line kernels of every context (one 16-column apply per precision the solves
        use, one context after another) before the k-points run concurrently, so the tuner's
        trials are timed on an otherwise idle device."""
        precs = (FP64,) if precision == FP64 else (FP64, precision)
        rng = np.random.default_rng(0)
        for ham in self.hams:
            x = rng.standard_normal((16, ham.n)) + 1j * rng.standard_normal((16, ham.n))
            for p in precs:
                ham.apply(x, p)

    @staticmethod
    def _record(it, e_tot, e_band, drho, stats, eig_s):
        return ScfRecord(it, 0.0, e_tot, e_band, drho, sum(st["iterations"] for st in stats),
                         sum(st["apply_calls"] for st in stats), sum(st["fp32_applies"] for st in stats), eig_s,
                         {k: sum(st[k] for st in stats) for k in SECTIONS})

    def _final(self, run, final_tol, preconditioner):
        """Tight FP64 eigenvalues on the last input potential (already set on every context)."""
        if final_tol is None:
            return
        run.eigenvalues = []
        for ham in self.hams:
            vals, _, _ = ham.eig_solve(self.sys.nbands, tol=final_tol, max_iter=400, precision=FP64,
                                       preconditioner=preconditioner, want_vectors=False, reuse=True)
            run.eigenvalues.append(vals)

    def close(self):
        if self.grid is not None:
            self.grid.close()
            self.grid = None
        if self.pool is not None:
            self.pool.shutdown()
            self.pool = None
        for h in self.hams:
            h.close()
        self.hams = []
        for c in self.ctxs:
            c.close()
        self.ctxs = []
