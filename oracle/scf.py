"""SCF algebra around the eigensolver, FP64 (numpy), for small cells.

Restates (paths relative to the dft-zig source tree):
  hartree_potential_at_g: V_H(G) = 8 pi rho(G) / G^2, G = 0 dropped   src/features/scf/potential.zig:34-44
  lda_pz (Ry), PZ81 correlation energy/potential                      src/features/xc/xc.zig:195-210, 344-384
  pbe (Ry): PBE exchange, PBE correlation on PW92                     src/features/xc/xc.zig:213-342
  GGA fields: grad n by i G, b = df/d|grad n|^2 grad n, v = df/dn - 2 div b
                                                                     src/features/scf/xc_fields.zig:87-140, 306-375
  assemble_energy_terms: total = band + (-E_H - int v_xc rho + E_xc) + ion terms
                                                                     src/features/scf/energy.zig:372-396
  ionic local potential sum_a e^{-iG.tau} v(|G|) / Omega, G = 0 skipped
                                                                     src/features/hamiltonian/hamiltonian.zig:421-474
  local_r = Re IFFT(N (V_ion + V_Hxc)(G)) in FFT order              src/features/scf/potential.zig:320-334
  occupations nocc = ceil(nelec/2), weight w_k occ 2                 src/features/scf/kpoint/density.zig:98-107
  Fermi-Dirac occupations, entropy -TS, bisection for mu            src/features/scf/fermi_level.zig:7-124,
                                                                     src/features/scf/kpoint/density.zig:398-404
  linear potential mixing V_in <- V_in + beta (V_out - V_in)         src/features/scf/mixing.zig:10-18
  Pulay/DIIS on (V_in, R = V_out - V_in), history 8, from iteration 4 src/features/scf/mixing.zig:183-234,
                                                                     src/features/scf/scf.zig:147-188,
                                                                     src/features/config/defaults.zig:42-45
  model dielectric preconditioner of the Pulay residual (diemac > 1)  src/features/scf/mixing.zig:309-338
  Pulay system by Gaussian elimination, |pivot| < 1e-12 -> equal weights
                                                                     src/features/scf/mixing.zig:340-400
  convergence on the RMS density change                              src/features/scf/util.zig:17-24

The reference mixes V_Hxc(G) over the full FFT grid; with diemac = 1 (its default) the Pulay
inner products are Parseval multiples of the real-space ones used here, so the coefficients and the
mixed potential are the same.  Energies per iteration use the Kohn-Sham functional of rho_out at the
input potential, band - int V_Hxc^in rho_out + E_H + E_xc, which equals assemble_energy_terms at
self-consistency.  The ion-ion (Ewald) and psp_core terms are constants of the structure and are left out of the
energies compared here; the eigenproblems use dense numpy.linalg.eigh.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import adapters
from . import density as odens
from . import fftconv, hamiltonian, pw

PI = math.pi


def lda_pz(n: np.ndarray):
    """(f, df/dn) in Ry per volume, element-wise; zero below n = 1e-12 (xc.zig:195-210)."""
    n = np.asarray(n, dtype=np.float64)
    f = np.zeros_like(n)
    v = np.zeros_like(n)
    m = n > 1e-12
    nm = n[m]
    n13 = np.power(nm, 1.0 / 3.0)
    c_x = -0.75 * math.pow(3.0 / PI, 1.0 / 3.0)
    rs = np.power(3.0 / (4.0 * PI * nm), 1.0 / 3.0)
    eps_c = np.empty_like(nm)
    v_c = np.empty_like(nm)
    lo = rs < 1.0
    a, b, c, d = 0.0311, -0.048, 0.0020, -0.0116
    lr = np.log(rs[lo])
    e_lo = a * lr + b + c * rs[lo] * lr + d * rs[lo]
    eps_c[lo] = e_lo
    v_c[lo] = e_lo - (rs[lo] / 3.0) * (a / rs[lo] + c * lr + c + d)
    hi = ~lo
    g, b1, b2 = -0.1423, 1.0529, 0.3334
    sr = np.sqrt(rs[hi])
    den = 1.0 + b1 * sr + b2 * rs[hi]
    e_hi = g / den
    eps_c[hi] = e_hi
    v_c[hi] = e_hi - (rs[hi] / 3.0) * (-g * (0.5 * b1 / sr + b2) / (den * den))
    f[m] = 2.0 * (nm * c_x * n13 + nm * eps_c)
    v[m] = 2.0 * (4.0 / 3.0 * c_x * n13 + v_c)
    return f, v


def pw92(rs: np.ndarray):
    """PW92 unpolarised correlation: (eps_c, d eps_c / d rs, v_c) in Ha (xc.zig:320-342)."""
    a, a1, b1, b2, b3, b4 = 0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294
    sr = np.sqrt(rs)
    f = b1 * sr + b2 * rs + b3 * rs * sr + b4 * rs * rs
    g = 1.0 + 1.0 / (2.0 * a * f)
    lg = np.log(g)
    eps = -2.0 * a * (1.0 + a1 * rs) * lg
    df = 0.5 * b1 / sr + b2 + 1.5 * b3 * sr + 2.0 * b4 * rs
    dlg = -(1.0 / (2.0 * a * f * f * g)) * df
    deps = -2.0 * a * (a1 * lg + (1.0 + a1 * rs) * dlg)
    return eps, deps, eps - (rs / 3.0) * deps


def pbe(n: np.ndarray, g2: np.ndarray):
    """(f, df/dn, df/d|grad n|^2) of PBE in Ry, element-wise; zero below n = 1e-12
    (xc.zig:213-318: exchange with mu = 0.2195149727645171, kappa = 0.804; correlation
    H(t^2, A) on PW92 with beta = 0.06672455060314922, gamma = 0.031090690869654895)."""
    n = np.asarray(n, dtype=np.float64)
    g2 = np.broadcast_to(np.asarray(g2, dtype=np.float64), n.shape)
    f = np.zeros_like(n)
    vn = np.zeros_like(n)
    vg = np.zeros_like(n)
    m = n > 1e-12
    nm, gm = n[m], g2[m]
    # Exchange.
    mu, kappa = 0.2195149727645171, 0.804
    c_x = -0.75 * math.pow(3.0 / PI, 1.0 / 3.0)
    n13 = np.power(nm, 1.0 / 3.0)
    ex0 = nm * c_x * n13
    kf = np.power(3.0 * PI * PI * nm, 1.0 / 3.0)
    den_x = 4.0 * kf * kf * nm * nm
    s2 = np.where(gm > 0.0, gm / den_x, 0.0)
    t = 1.0 + mu * s2 / kappa
    fx = 1.0 + kappa - kappa / t
    dfx = mu / (t * t)
    fx_n = 4.0 / 3.0 * c_x * n13 * fx + ex0 * dfx * (-8.0 / 3.0 * s2 / nm)
    fx_g = ex0 * dfx / den_x
    # Correlation.
    beta, gamma = 0.06672455060314922, 0.031090690869654895
    rs = np.power(3.0 / (4.0 * PI * nm), 1.0 / 3.0)
    eps_c, deps_drs, v_c = pw92(rs)
    deps_dn = deps_drs * (-rs / (3.0 * nm))
    ks2 = 4.0 * kf / PI
    den_c = 4.0 * ks2 * nm * nm
    t2 = np.where(gm > 0.0, gm / den_c, 0.0)
    bg = beta / gamma
    x = bg * t2
    u = np.exp(-eps_c / gamma)
    um = u - 1.0
    big = np.abs(um) > 1e-12
    A = np.where(big, bg / np.where(big, um, 1.0), 0.0)
    y = A * t2
    d2 = 1.0 + y + y * y
    q = (1.0 + y) / d2
    arg = 1.0 + x * q
    pos = arg > 0.0
    H = np.where(pos, gamma * np.log(np.where(pos, arg, 1.0)), 0.0)
    dq_dt2 = -A * y * (2.0 + y) / (d2 * d2)
    dH_dt2 = np.where(pos, gamma * (bg * q + x * dq_dt2) / np.where(pos, arg, 1.0), 0.0)
    dH_dA = np.where(pos, -gamma * x * t2 * y * (2.0 + y) / (np.where(pos, arg, 1.0) * d2 * d2), 0.0)
    dA_dn = np.where(big, bg * u / (gamma * np.where(big, um, 1.0) ** 2) * deps_dn, 0.0)
    dH_dn = dH_dt2 * (-7.0 / 3.0 * t2 / nm) + dH_dA * dA_dn
    fc = nm * (eps_c + H)
    fc_n = v_c + H + nm * dH_dn
    fc_g = nm * dH_dt2 / den_c
    f[m] = 2.0 * (ex0 * fx + fc)
    vn[m] = 2.0 * (fx_n + fc_n)
    vg[m] = 2.0 * (fx_g + fc_g)
    return f, vn, vg


def gradient(values, grid, recip):
    """(d/dx, d/dy, d/dz) of a real grid function by i G in reciprocal space, real part kept
    (xc_fields.zig:306-339)."""
    n = values.shape[0]
    vg = fftconv.forward3d(values.astype(np.complex128), grid) / n
    g, _ = g_squared_grid(grid, recip)
    return [fftconv.inverse3d(1j * vg * g[:, c] * n, grid).real for c in range(3)]


def divergence(bx, by, bz, grid, recip):
    """div b by i G . b(G), real part kept (xc_fields.zig:341-375)."""
    n = bx.shape[0]
    g, _ = g_squared_grid(grid, recip)
    acc = np.zeros(n, dtype=np.complex128)
    for c, comp in enumerate((bx, by, bz)):
        acc += fftconv.forward3d(comp.astype(np.complex128), grid) / n * g[:, c]
    return fftconv.inverse3d(1j * acc * n, grid).real


def xc_fields(rho, grid, recip, functional="lda_pz"):
    """(f_xc, v_xc) on the grid in Ry: LDA point-wise, PBE with v = df/dn - 2 div(df/dg2 grad n)
    (xc_fields.zig:64-140, 377-386)."""
    if functional == "lda_pz":
        return lda_pz(rho)
    if functional != "pbe":
        raise ValueError(functional)
    gx, gy, gz = gradient(rho, grid, recip)
    f, vn, vg = pbe(rho, gx * gx + gy * gy + gz * gz)
    return f, vn - 2.0 * divergence(vg * gx, vg * gy, vg * gz, grid, recip)


def g_squared_grid(grid, recip):
    """|G|^2 on the FFT-order grid (frequencies wrapped, index_to_freq)."""
    nx, ny, nz = grid
    fx = np.array([pw.index_to_freq(i, nx) for i in range(nx)])
    fy = np.array([pw.index_to_freq(i, ny) for i in range(ny)])
    fz = np.array([pw.index_to_freq(i, nz) for i in range(nz)])
    Z, Y, X = np.meshgrid(fz, fy, fx, indexing="ij")
    g = X.ravel()[:, None] * recip[0] + Y.ravel()[:, None] * recip[1] + Z.ravel()[:, None] * recip[2]
    return g, np.sum(g * g, axis=1)


def hartree(rho_r, grid, recip, volume):
    """V_H(r) and E_H = (1/2) int V_H rho (Ry)."""
    n = rho_r.shape[0]
    rho_g = fftconv.forward3d(rho_r, grid) / n
    _, g2 = g_squared_grid(grid, recip)
    vg = np.zeros_like(rho_g)
    nz = g2 > 1e-12
    vg[nz] = rho_g[nz] * 8.0 * PI / g2[nz]
    vh = (fftconv.inverse3d(vg * n, grid)).real
    e_h = 0.5 * float(np.sum(vh * rho_r)) * volume / n
    return vh, e_h


def ionic_potential_r(grid, recip, volume, positions, vq_fn):
    """V_ion(r) = Re sum_G V_ion(G) e^{iG.r}, V_ion(G) = (1/Omega) sum_a e^{-iG.tau_a} v(|G|)."""
    n = int(np.prod(grid))
    g, g2 = g_squared_grid(grid, recip)
    q = np.sqrt(g2)
    vg = np.zeros(n, dtype=np.complex128)
    nz = q > 1e-10
    sf = np.zeros(n, dtype=np.complex128)
    for tau in positions:
        sf += np.exp(-1j * (g @ tau))
    # v(|G|) once per distinct |G| (bitwise the same values as one call per G).
    qs, inv = np.unique(q[nz], return_inverse=True)
    vg[nz] = sf[nz] * np.array([vq_fn(x) for x in qs])[inv] / volume
    return (fftconv.inverse3d(vg * n, grid)).real


def solve_pulay(b: np.ndarray, rhs: np.ndarray):
    """Gaussian elimination with partial pivoting; a pivot below 1e-12 gives equal weights
    (mixing.zig:340-400).  Returns (coefficients, singular)."""
    a = np.array(b, dtype=np.float64)
    r = np.array(rhs, dtype=np.float64)
    size = a.shape[0]
    for col in range(size):
        piv = col + int(np.argmax(np.abs(a[col:, col])))
        if piv != col:
            a[[col, piv]] = a[[piv, col]]
            r[[col, piv]] = r[[piv, col]]
        if abs(a[col, col]) < 1e-12:
            out = np.full(size, 1.0 / (size - 1))
            out[-1] = 0.0
            return out, True
        for row in range(col + 1, size):
            f = a[row, col] / a[col, col]
            a[row, col:] -= f * a[col, col:]
            r[row] -= f * r[col]
    x = np.zeros(size)
    for i in range(size - 1, -1, -1):
        x[i] = (r[i] - np.dot(a[i, i + 1:], x[i + 1:])) / a[i, i]
    return x, False


class PulayMixer:
    """mix_with_residual (mixing.zig:183-234): append (v_in, residual), drop the oldest beyond
    `history`, linear step while fewer than 2 entries, else v = sum_i c_i (v_i + beta R_i) with
    [B -1; -1 0] c = [0; -1], B_ij = <R_i|R_j>.

    relative=False is the reference as written: its absolute pivot threshold (1e-12) takes the
    equal-weights branch once <R|R> is below about 1e-12 (an RMS potential residual near 1e-6
    Ry), after which the iteration wanders instead of converging.  relative=True is the form
    the GPU driver uses: B is divided by its largest diagonal entry (the coefficients of a
    regular system are invariant under a scaling of B), and a singular system drops the history
    except the newest pair and takes the linear step instead of averaging the history."""

    def __init__(self, history: int, relative: bool = True):
        self.history = history
        self.relative = relative
        self.v, self.r = [], []

    def mix(self, v_in: np.ndarray, residual: np.ndarray, beta: float) -> np.ndarray:
        self.v.append(v_in.copy())
        self.r.append(residual.copy())
        while len(self.v) > self.history:
            self.v.pop(0)
            self.r.pop(0)
        m = len(self.v)
        if m < 2:
            return v_in + beta * residual
        # <R_i|R_j> over the reference's V(G) = (1/N) sum_r V(r) e^{-iG.r} on the full grid
        # (fft_grid.zig:27-36) is (1/N) sum_r R_i(r) R_j(r) by Parseval; the scale matters for
        # the absolute singular-pivot threshold of solve_pulay.
        n = self.r[0].shape[0]
        b = np.zeros((m + 1, m + 1))
        for i in range(m):
            for j in range(m):
                b[i, j] = float(np.dot(self.r[i], self.r[j])) / n
        if self.relative:
            b[:m, :m] /= max(float(np.max(np.diag(b)[:m])), 1e-300)
        b[:m, m] = -1.0
        b[m, :m] = -1.0
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        c, singular = solve_pulay(b, rhs)
        if singular and self.relative:
            del self.v[:-1], self.r[:-1]
            return v_in + beta * residual
        out = np.zeros_like(v_in)
        for i in range(m):
            out += c[i] * (self.v[i] + beta * self.r[i])
        return out


def smearing_occ(e, mu, sigma):
    """Fermi-Dirac occupation 1/(1 + e^{(e - mu)/sigma}); a step for sigma <= 0 (fermi_level.zig:7-11)."""
    e = np.asarray(e, dtype=np.float64)
    if sigma <= 0.0:
        return (e <= mu).astype(np.float64)
    return 1.0 / (1.0 + np.exp((e - mu) / sigma))


def smearing_entropy(f):
    """-(f ln f + (1 - f) ln(1 - f)), zero within 1e-12 of 0 and 1 (fermi_level.zig:15-19)."""
    f = np.asarray(f, dtype=np.float64)
    ok = (f > 1e-12) & (f < 1.0 - 1e-12)
    fs = np.where(ok, f, 0.5)
    return np.where(ok, -(fs * np.log(fs) + (1.0 - fs) * np.log(1.0 - fs)), 0.0)


def find_fermi_level(nelec, sigma, eigenvalues, kweights):
    """Bisection for mu with sum_k 2 w_k sum_b f(e_kb) = nelec: bracket [min - pad, max + pad],
    pad = max(10 sigma, 1e-3) doubled up to 12 times, 80 halvings, early exit at |dN| < 1e-8
    (fermi_level.zig:54-124)."""
    vals = [np.asarray(v, dtype=np.float64) for v in eigenvalues]
    lo_e = min(float(v.min()) for v in vals)
    hi_e = max(float(v.max()) for v in vals)

    def count(mu):
        return sum(2.0 * w * float(np.sum(smearing_occ(v, mu, sigma))) for v, w in zip(vals, kweights))

    pad = max(10.0 * sigma, 1e-3)
    low, high = lo_e - pad, hi_e + pad
    expand = 0
    while (count(low) > nelec or count(high) < nelec) and expand < 12:
        pad *= 2.0
        low, high = lo_e - pad, hi_e + pad
        expand += 1
    left, right = low, high
    for _ in range(80):
        mid = 0.5 * (left + right)
        c = count(mid)
        if abs(c - nelec) < 1e-8:
            return mid
        if c > nelec:
            right = mid
        else:
            left = mid
    return 0.5 * (left + right)


@dataclass
class ScfResult:
    energies: list = field(default_factory=list)  # total energy (Ry) per iteration
    drho: list = field(default_factory=list)
    eigenvalues: list = field(default_factory=list)  # per k, final
    rho: np.ndarray | None = None
    converged: bool = False


def model_dielectric(residual_r, grid, recip, diemac, dielng):
    """P(G) R(G) with P = (dielng^2 G^2 + 1/diemac) / (dielng^2 G^2 + 1), P(0) = 1/diemac
    (mixing.zig:309-338), applied to a real-space residual through the FFT."""
    n = residual_r.shape[0]
    _, g2 = g_squared_grid(grid, recip)
    x = dielng * dielng * g2
    p = np.where(g2 < 1e-12, 1.0 / diemac, (x + 1.0 / diemac) / (x + 1.0))
    return fftconv.inverse3d(fftconv.forward3d(residual_r.astype(np.complex128), grid) * p, grid).real


def run_scf(system, n_iter=40, beta=0.3, conv=1e-8, pulay_history=0, pulay_start=4, pulay_relative=True,
            diemac=1.0, dielng=1.0):
    """Fixed-occupation SCF with potential mixing (linear, or Pulay from iteration `pulay_start`
    when pulay_history > 0, scf.zig:164); dense eigh per k; system.xc ("lda_pz" default, "pbe")
    picks the functional."""
    xc = getattr(system, "xc", "lda_pz")
    grid, recip, vol = system.grid, system.recip, system.volume
    n = int(np.prod(grid))
    upfs = getattr(system, "species_upf", None)
    if upfs:
        if len(upfs) != 1:
            raise ValueError("the oracle SCF takes one species")
        vq = adapters.upf_local_q(upfs[0])
    else:
        vq = adapters.gth_local_q(system.z_ion, system.rloc, system.c1)
    v_ion = ionic_potential_r(grid, recip, vol, system.atom_pos, vq)
    rho = np.full(n, system.nelec / vol)
    vh, _ = hartree(rho, grid, recip, vol)
    _, vxc = xc_fields(rho, grid, recip, xc)
    v_in = vh + vxc
    res = ScfResult()
    mixer = PulayMixer(pulay_history, pulay_relative) if pulay_history > 0 else None
    sigma = float(getattr(system, "smearing", 0.0))
    for it in range(n_iter):
        local_r = v_ion + v_in
        rho_out = np.zeros(n)
        e_band = 0.0
        e_ts = 0.0
        eigs = []
        nb = system.nbands
        # Solve every k first (the smearing path needs all eigenvalues for mu, smearing.zig:107-169).
        solved = []
        for kp in system.kpoints:
            ctx = adapters.context_for_k(system, kp, local_r)
            ev, vec = np.linalg.eigh(hamiltonian.dense_hamiltonian(ctx))
            solved.append((ctx, ev[:nb], vec[:, :nb]))
        mu = find_fermi_level(system.nelec, sigma, [e for _, e, _ in solved], system.kweights) if sigma > 0 else None
        for (ctx, ev, vec), wk in zip(solved, system.kweights):
            if sigma > 0:
                occ = smearing_occ(ev, mu, sigma)
                wts = 2.0 * wk * occ
                e_ts += float(np.sum(2.0 * wk * sigma * smearing_entropy(occ)))
            else:
                wts = odens.band_weights(system.nelec, nb, wk)
            e_band += float(np.sum(wts * ev))
            rho_out = odens.accumulate_density(ctx.fft_idx, grid, vec.T, wts, 1.0 / vol, rho_out)
            eigs.append(ev)
        vh_out, e_h = hartree(rho_out, grid, recip, vol)
        f_xc, vxc_out = xc_fields(rho_out, grid, recip, xc)
        e_xc = float(np.sum(f_xc)) * vol / n
        # Kohn-Sham energy with the input potential: band - int V_Hxc^in rho_out + E_H + E_xc.
        e_tot = e_band - float(np.sum(v_in * rho_out)) * vol / n + e_h + e_xc - e_ts
        drho = math.sqrt(float(np.mean((rho_out - rho) ** 2)))
        res.energies.append(e_tot)
        res.drho.append(drho)
        res.eigenvalues = eigs
        rho = rho_out
        v_out = vh_out + vxc_out
        if drho < conv:
            res.converged = True
            break
        if mixer is not None and it >= pulay_start:
            resid = v_out - v_in
            if diemac > 1.0:
                resid = model_dielectric(resid, grid, recip, diemac, dielng)
            v_in = mixer.mix(v_in, resid, beta)
        else:
            v_in = v_in + beta * (v_out - v_in)
    res.rho = rho
    return res
