"""Band-by-band CG eigensolver restated in numpy FP64 (test infrastructure only).

Restates src/features/linalg/cg/cg.zig (paths relative to the dft-zig source tree):
  solve: defaults max_iter 20, tol 1e-6; random start from the reference LCG and MGS   :79-138
  solve_band: orthonormalise psi_b against lower bands, lambda = <psi|H psi>          :155-192
  prepare_cg_iteration: r = H psi - lambda psi, orthogonalise against bands < b;
      |r| < tol stops; p = K r (kinetic shift), orthogonalise p against bands <= b;
      d = p + max(0, zr/zr_prev) d (zr = Re<p, r>); orthogonalise d against bands <= b,
      |d| < 1e-14 stops; normalise; H d                                                 :194-248
  three-term Rayleigh-Ritz on {psi, d, psi_prev (orthogonalised, |.| > 1e-6)}, else
      the closed-form 2x2 on {psi, d}                                                   :250-397
  update psi = c1 psi_old + c2 d (+ c3 psi_prev), same for H psi, renormalise; stop when
      |lambda_new - lambda| < tol                                                       :399-448
  final subspace rotation: H projected on the band block, eigh, rotate                 :450-482

The 3x3 eigenvector's phase (LAPACK zheev's, not reproducible here) enters the next update of
the conjugate direction, so it is fixed by convention: the coefficient of psi is made real and
non-negative (the GPU solver uses the same convention).  The two-term step keeps the reference's
closed form (c1 = -a12, c2 = a11 - lambda), whose psi coefficient carries the phase of -a12.
"""
from __future__ import annotations

import numpy as np

from . import eigensolver as oeig


def _orth(v, basis, m):
    for j in range(m):
        v -= np.vdot(basis[j], v) * basis[j]
    return v


def solve(apply, diag, nbands, max_iter=0, tol=0.0, init=None):
    """apply(x [n]) -> H x.  Returns (values[nbands], vectors[nbands, n])."""
    n = diag.shape[0]
    max_iter = max_iter or 20
    tol = tol or 1e-6
    rng = oeig.Lcg()
    vecs = np.zeros((nbands, n), dtype=np.complex128)
    cols = 0
    if init is not None:
        cols = min(nbands, init.shape[0])
        vecs[:cols] = init[:cols]
    if cols < nbands:
        vecs[cols:] = oeig.init_random_vectors(n, nbands - cols, rng)
    oeig.orthonormalize_all(vecs, nbands, rng)
    values = np.zeros(nbands)
    for b in range(nbands):
        psi = vecs[b]
        if b > 0:
            psi = _orth(psi, vecs, b)
            psi /= np.linalg.norm(psi)
        hpsi = apply(psi)
        lam = float(np.vdot(psi, hpsi).real)
        d = np.zeros(n, dtype=np.complex128)
        prev_zr, has_prev = 0.0, False
        psi_old = hpsi_old = None
        for _ in range(max_iter):
            r = _orth(hpsi - lam * psi, vecs, b)
            if np.linalg.norm(r) < tol:
                break
            vecs[b] = psi
            p = _orth(oeig.precondition(diag, lam, r), vecs, b + 1)
            zr = float(np.vdot(p, r).real)
            d = p + max(0.0, zr / prev_zr) * d if (has_prev and prev_zr > 1e-30) else p.copy()
            prev_zr = zr
            d = _orth(d, vecs, b + 1)
            dn = np.linalg.norm(d)
            if dn < 1e-14:
                break
            d /= dn
            hd = apply(d)
            three = None
            if has_prev:
                pp = psi_old - np.vdot(psi, psi_old) * psi
                hp = hpsi_old - np.vdot(psi, psi_old) * hpsi
                ov = np.vdot(d, pp)
                pp = pp - ov * d
                hp = hp - ov * hd
                pn = np.linalg.norm(pp)
                if pn > 1e-6:
                    pp /= pn
                    hp /= pn
                    basis = np.stack([psi, d, pp])
                    hb = np.stack([hpsi, hd, hp])
                    h3 = basis.conj() @ hb.T
                    h3 = 0.5 * (h3 + h3.conj().T)
                    h3[0, 0] = lam
                    ev, evec = np.linalg.eigh(h3)
                    c = evec[:, 0]
                    if abs(c[0]) > 0.0:
                        c = c * (abs(c[0]) / c[0])
                    three = (ev[0], c, pp, hp)
            psi_old, hpsi_old = psi.copy(), hpsi.copy()
            has_prev = True
            if three is not None:
                lam_new, c, pp, hp = three
                psi = c[0] * psi_old + c[1] * d + c[2] * pp
                hpsi = c[0] * hpsi_old + c[1] * hd + c[2] * hp
            else:
                a11, a22, a12 = lam, float(np.vdot(d, hd).real), np.vdot(psi, hd)
                avg, diff = 0.5 * (a11 + a22), 0.5 * (a11 - a22)
                off = abs(a12) ** 2
                lam_new = avg - np.sqrt(diff * diff + off)
                if off > 1e-30:
                    c1, c2 = -a12, complex(a11 - lam_new)
                elif a11 <= a22:
                    c1, c2 = 1.0, 0.0
                else:
                    c1, c2 = 0.0, 1.0
                cn = np.sqrt(abs(c1) ** 2 + abs(c2) ** 2)
                if cn > 1e-30:
                    c1, c2 = c1 / cn, c2 / cn
                psi = c1 * psi_old + c2 * d
                hpsi = c1 * hpsi_old + c2 * hd
            nrm = np.linalg.norm(psi)
            if nrm > 1e-14:
                psi /= nrm
                hpsi /= nrm
            change = abs(lam_new - lam)
            lam = float(lam_new)
            if change < tol:
                break
        values[b] = lam
        vecs[b] = psi
        if b > 0:
            vecs[b] = _orth(vecs[b], vecs, b)
            vecs[b] /= np.linalg.norm(vecs[b])
    hv = np.stack([apply(v) for v in vecs])
    hsub = vecs.conj() @ hv.T
    hsub = 0.5 * (hsub + hsub.conj().T)
    ev, c = np.linalg.eigh(hsub)
    return ev, (c.T @ vecs)
