"""The allocation-free eigensolver entry point (dzhip_heev_dev_noalloc: every buffer in the
library's static device array) against the allocating one (dzhip_heev_dev) on the same device
inputs, through the C ABI: two calls in a row, the second with a smaller matrix (the array keeps the
first call's contents), results identical bit for bit (the same kernels on the same data), and both
against numpy.linalg.eigh (eigenvalues within 1e-12 of the spectral radius)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hermitian(m, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
    lam = np.sort(rng.uniform(-3.0, 20.0, m))
    a = (q * lam) @ q.conj().T
    return 0.5 * (a + a.conj().T)


def test_noalloc_matches_allocating_entry(dz, ctx):
    import torch

    stream = torch.cuda.Stream()
    for m, nev, seed in ((240, 200, 11), (64, 20, 12)):
        a = _hermitian(m, seed)
        at = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
        w1 = torch.zeros(nev, dtype=torch.float64, device="cuda")
        z1 = torch.zeros((nev, m), dtype=torch.complex128, device="cuda")
        w2 = torch.full((nev,), 7.0, dtype=torch.float64, device="cuda")
        z2 = torch.full((nev, m), 3.0, dtype=torch.complex128, device="cuda")
        torch.cuda.synchronize()
        ctx.heev_dev(at.data_ptr(), m, m, nev, w1.data_ptr(), z1.data_ptr(), m, stream.cuda_stream)
        ctx.heev_check(stream.cuda_stream)
        ctx.heev_dev_noalloc(at.data_ptr(), m, m, nev, w2.data_ptr(), z2.data_ptr(), m, stream.cuda_stream)
        stream.synchronize()
        assert torch.equal(w1, w2) and torch.equal(z1, z2), (m, nev)
        ref = np.linalg.eigvalsh(a)[:nev]
        scale = max(1.0, float(np.max(np.abs(ref))))
        assert float(np.max(np.abs(w2.cpu().numpy() - ref))) <= 1e-12 * scale
        z = z2.cpu().numpy().T
        assert float(np.max(np.abs(z.conj().T @ z - np.eye(nev)))) <= 1e-12


def test_noalloc_rejects_large_matrices(dz, ctx):
    import torch

    m = 257
    a = torch.zeros((m, m), dtype=torch.complex128, device="cuda")
    w = torch.zeros(m, dtype=torch.float64, device="cuda")
    with pytest.raises(dz.BadArgument):
        ctx.heev_dev_noalloc(a.data_ptr(), m, m, 4, w.data_ptr(), a.data_ptr(), m)
