"""The device Hermitian eigensolver (dzhip_heev: Householder tridiagonalisation in one
grid-synchronised launch, multisection, inverse iteration, back-transformation) against numpy.linalg.eigh, which is
LAPACK zheevd: the reference's subspace eigenproblems call LAPACK zheev
(src/features/linalg/lobpcg/common.zig:440-537).

Bars: eigenvalues within 1e-12 of the spectral radius; Z^H Z = I within 1e-12; residuals
|A z - lambda z| within 1e-11 of the spectral radius.  Sizes 1, 2, 20, 164 and 648 (the block
Davidson's first and largest subspaces at 160 bands), exact and near degeneracies, and a
Rayleigh-Ritz matrix of the Si-8 Hamiltonian."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def random_hermitian(m, rng, spectrum=None):
    q, _ = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
    lam = np.sort(rng.uniform(-2.0, 30.0, m)) if spectrum is None else np.asarray(spectrum, dtype=float)
    a = (q * lam) @ q.conj().T
    return 0.5 * (a + a.conj().T), lam


def check(ctx, a, nev, tol_val=1e-12, tol_res=1e-11):
    w, z = ctx.heev(a, nev)
    ref = np.linalg.eigvalsh(a)[:nev]
    scale = max(1.0, float(np.max(np.abs(ref))), float(np.linalg.norm(a, 2)))
    dval = float(np.max(np.abs(w - ref)))
    orth = float(np.max(np.abs(z.conj().T @ z - np.eye(nev))))
    res = float(np.max(np.linalg.norm(a @ z - z * w, axis=0)))
    print(f"m={a.shape[0]} nev={nev} |dw|={dval:.2e} orth={orth:.2e} res={res:.2e} scale={scale:.1f}")
    assert dval <= tol_val * scale
    assert orth <= 1e-12
    assert res <= tol_res * scale
    return w, z


@pytest.mark.parametrize("m,nev", [(1, 1), (2, 2), (20, 20), (20, 7), (164, 160), (648, 160), (648, 648)])
def test_heev_random(ctx, m, nev):
    rng = np.random.default_rng(1000 + m + nev)
    a, _ = random_hermitian(m, rng)
    check(ctx, a, nev)


def test_heev_degenerate_and_clustered(ctx):
    """Exactly degenerate triplets and pairs (symmetry-degenerate bands), a cluster split by
    1e-13, and a zero off-diagonal block (a reducible tridiagonal)."""
    rng = np.random.default_rng(7)
    lam = np.sort(np.concatenate([np.repeat(rng.uniform(-1, 1, 20), 3), np.repeat(rng.uniform(1, 5, 30), 2),
                                  1.5 + 1e-13 * np.arange(6), rng.uniform(5, 20, 100)]))
    a, _ = random_hermitian(lam.size, rng, lam)
    check(ctx, a, 160)
    # Block diagonal: the reduction meets zero columns (tau = 0 reflectors).
    b1, _ = random_hermitian(40, rng)
    b = np.zeros((80, 80), dtype=complex)
    b[:40, :40] = b1
    b[40:, 40:] = b1
    check(ctx, b, 80)
    # Real diagonal matrix: every reflector is the identity.
    check(ctx, np.diag(np.linspace(-3, 3, 50)).astype(complex), 50)


def test_heev_errors(dz, ctx):
    with pytest.raises(dz.BadArgument):
        ctx.heev(np.eye(4, dtype=complex), 5)
    with pytest.raises(dz.BadArgument):
        ctx.heev(np.eye(1025, dtype=complex), 4)


def test_heev_rayleigh_ritz_matrix(dz, ctx):
    """T = V^H H V of the Si-8 Hamiltonian over an orthonormalised random block (the subspace
    matrix the Davidson solver diagonalises)."""
    from dzhip import workloads
    from oracle import adapters, hamiltonian

    w = workloads.si8(nbands=8)
    rng = np.random.default_rng(3)
    v = rng.standard_normal((w.npw, 200)) + 1j * rng.standard_normal((w.npw, 200))
    v, _ = np.linalg.qr(v)
    hv = hamiltonian.apply_hamiltonian_batched(adapters.context_from_workload(w), v.T).T
    t = v.conj().T @ hv
    t = 0.5 * (t + t.conj().T)
    check(ctx, t, 160)


def test_heev_concurrent_threads(dz, ctx):
    """Six host threads, each with its own context (own stream), solving at the same time.  The
    blocking eigensolver launches its tridiagonalisation as an ordinary dispatch admitted by the
    device's gate: at m = 648 a grid takes 14 CUs of each XCD, so at most two of these grids run
    at once and the rest wait for their share.  Every result must equal the one-thread result."""
    import threading

    rng = np.random.default_rng(11)
    sizes = (648, 492, 328, 648, 164, 492)
    mats = [random_hermitian(m, rng)[0] for m in sizes]
    ctxs = [dz.Context(0) for _ in sizes]
    alone = [ctxs[0].heev(a, 160) for a in mats]
    out, errs = [None] * len(sizes), []

    def work(i):
        try:
            res = []
            for _ in range(3):
                res.append(ctxs[i].heev(mats[i], 160))
            out[i] = res
        except Exception as e:  # reported on the main thread
            errs.append(e)

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(sizes))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for c in ctxs:
        c.close()
    assert not errs, errs
    for i, a in enumerate(mats):
        ref = np.linalg.eigvalsh(a)[:160]
        for w, z in out[i]:
            assert np.array_equal(w, alone[i][0]) and np.array_equal(z, alone[i][1])
        assert np.max(np.abs(alone[i][0] - ref)) <= 1e-12 * max(1.0, float(np.max(np.abs(ref))))


def test_heev_shared_launch_falls_back_to_cooperative():
    """The blocking eigensolver's admitted launch gives up when its barrier waits longer than its
    spin limit and the call is repeated cooperatively.  Forced from the host: a child process with
    the admitted launch's spin limit at one poll (its barrier gives up at once; every workgroup
    exits normally) must still return the exact spectrum."""
    import os
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    script = (
        "import sys, numpy as np; sys.path.insert(0, %r); import dzhip\n"
        "rng = np.random.default_rng(5); m = 648\n"
        "q, _ = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))\n"
        "lam = np.sort(rng.uniform(-2.0, 30.0, m)); a = (q * lam) @ q.conj().T; a = 0.5 * (a + a.conj().T)\n"
        "ctx = dzhip.Context(0); w, z = ctx.heev(a, 160)\n"
        "print(float(np.max(np.abs(w - lam[:160]))), float(np.max(np.abs(z.conj().T @ z - np.eye(160)))))\n"
    ) % str(root / "dft-zig_amd")
    env = dict(os.environ, DZHIP_HEEV_SHARED_SPIN="1")
    print("child: heev m=648 nev=160 with DZHIP_HEEV_SHARED_SPIN=1", flush=True)
    out = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    dval, orth = map(float, out.stdout.split()[-2:])
    assert dval < 1e-12 * 30.0 and orth < 1e-12
