"""CPU model of the z-register plane kernels (dft-zig_amd/csrc/hpsi.hip, k_plane_inv_zr and
k_plane_fwd_zr): the plane's active z-frequency rows (ascending, row b of the plane holds
z-frequency zrows[b]), the three-lanes-per-y-line decomposition of the z-transform (nz = 3M,
lane t holds the residue z = 3u + t), its per-lane twiddles w^(+-t k0) and three-point factors,
the partner lanes (t + 1) mod 3 and (t + 2) mod 3 of the forward exchange, the uniform skip of
k0 whose three frequencies are all inactive, and writes to active rows only; checked against the
2-D transform written with numpy's FFT on random sphere-shaped supports.  The y-transforms of the
active rows are numpy's here (the kernels' in-LDS stages are pinned by tests/fft_model.py)."""
import numpy as np
import pytest

C3, S3 = -0.5, np.sqrt(3.0) / 2.0


def active_rows(zf):
    rows = sorted(set(int(z) for z in zf))
    zb = {z: b for b, z in enumerate(rows)}
    return rows, zb


def model_inverse(yf, zf, vals, ny, nz):
    """Sphere values at (yf, zf) -> real-space plane [z][y], as k_plane_inv_zr does it."""
    m = nz // 3
    rows, zb = active_rows(zf)
    buf = np.zeros((len(rows), ny), complex)
    for y, z, v in zip(yf, zf, vals):
        buf[zb[int(z)], int(y)] = v
    # Inverse along y on the active rows (unnormalised, e^{+i}).
    buf = np.fft.ifft(buf, axis=1) * ny
    zbt = [zb.get(k, -1) for k in range(nz)]
    out = np.zeros((nz, ny), complex)
    w = np.exp(2j * np.pi / nz)
    for yl in range(ny):
        for t in range(3):
            f = np.exp(2j * np.pi * t / 3)  # (fc, fs); Y[k0 + 2M] takes its conjugate
            c = np.zeros(m, complex)
            for k0 in range(m):
                b0, b1, b2 = zbt[k0], zbt[k0 + m], zbt[k0 + 2 * m]
                s = buf[b0, yl] if b0 >= 0 else 0.0
                if b1 >= 0:
                    s = s + buf[b1, yl] * f
                if b2 >= 0:
                    s = s + buf[b2, yl] * np.conj(f)
                c[k0] = s * w ** (t * k0)
            x = np.fft.ifft(c) * m  # inverse length-M transform
            for u in range(m):
                out[3 * u + t, yl] = x[u]
    return out


def model_forward(plane, yf, zf, ny, nz):
    """Real-space plane [z][y] -> its forward 2-D transform at the sphere points, as
    k_plane_fwd_zr does it (without the kinetic and accumulate terms)."""
    m = nz // 3
    rows, zb = active_rows(zf)
    zbt = [zb.get(k, -1) for k in range(nz)]
    buf = np.full((len(rows), ny), np.nan, complex)  # unwritten entries stay NaN
    w = np.exp(2j * np.pi / nz)
    # Lane factors of the three-point combination: own, next, after-next.
    own = {0: 1.0, 1: C3 - 1j * S3, 2: C3 - 1j * S3}
    nxt = {0: 1.0, 1: C3 + 1j * S3, 2: 1.0}
    aft = {0: 1.0, 1: 1.0, 2: C3 + 1j * S3}
    for yl in range(ny):
        h = np.zeros((3, m), complex)
        for t in range(3):
            f = np.fft.fft(plane[t::3, yl])  # forward length-M transform of the residue
            h[t] = f * w ** (-t * np.arange(m))
        for k0 in range(m):
            bs = (zbt[k0], zbt[k0 + m], zbt[k0 + 2 * m])
            if (bs[0] & bs[1] & bs[2]) < 0:
                continue
            for t in range(3):
                x = own[t] * h[t, k0] + nxt[t] * h[(t + 1) % 3, k0] + aft[t] * h[(t + 2) % 3, k0]
                if bs[t] >= 0:
                    buf[bs[t], yl] = x
    buf = np.fft.fft(buf, axis=1)
    return np.array([buf[zb[int(z)], int(y)] for y, z in zip(yf, zf)])


def sphere_support(ny, nz, radius, rng, shift=0):
    """A disc of frequencies (a plane of a G-sphere), wrapped into [0, n)."""
    pts = []
    for gy in range(-ny // 2, ny // 2 + 1):
        for gz in range(-nz // 2, nz // 2 + 1):
            if gy * gy + (gz - shift) ** 2 <= radius * radius:
                pts.append((gy % ny, gz % nz))
    pts = sorted(set(pts))
    rng.shuffle(pts)
    yf = np.array([p[0] for p in pts])
    zf = np.array([p[1] for p in pts])
    return yf, zf


@pytest.mark.parametrize("ny,nz,radius,shift", [(8, 12, 3.5, 0), (10, 15, 4.2, 1), (20, 24, 7.1, 0),
                                                (16, 45, 11.0, 3), (75, 75, 17.9, 0), (33, 90, 20.1, 2)])
def test_zreg_inverse_matches_2d_ifft(ny, nz, radius, shift):
    rng = np.random.default_rng(ny * 100 + nz)
    yf, zf = sphere_support(ny, nz, radius, rng, shift)
    vals = rng.standard_normal(len(yf)) + 1j * rng.standard_normal(len(yf))
    full = np.zeros((nz, ny), complex)
    full[zf, yf] = vals
    ref = np.fft.ifft2(full) * (ny * nz)
    got = model_inverse(yf, zf, vals, ny, nz)
    assert np.allclose(got, ref, atol=1e-10 * np.max(np.abs(ref)))
    # About half the z-frequencies are active for a disc of a grid of twice its diameter.
    assert len(set(zf)) < nz


@pytest.mark.parametrize("ny,nz,radius,shift", [(8, 12, 3.5, 0), (10, 15, 4.2, 1), (20, 24, 7.1, 0),
                                                (16, 45, 11.0, 3), (75, 75, 17.9, 0), (33, 90, 20.1, 2)])
def test_zreg_forward_matches_2d_fft(ny, nz, radius, shift):
    rng = np.random.default_rng(ny * 100 + nz + 1)
    yf, zf = sphere_support(ny, nz, radius, rng, shift)
    plane = rng.standard_normal((nz, ny)) + 1j * rng.standard_normal((nz, ny))
    ref = np.fft.fft2(plane)[zf, yf]
    got = model_forward(plane, yf, zf, ny, nz)
    assert np.all(np.isfinite(got))  # every sphere point's row was written
    assert np.allclose(got, ref, atol=1e-10 * np.max(np.abs(ref)))


def test_zreg_forward_with_a_wrong_factor_fails():
    """The model is sensitive to the lane factors: swapping the next / after-next factor of lane
    1 (a plausible slip) breaks the forward transform."""
    ny, nz = 10, 15
    rng = np.random.default_rng(3)
    yf, zf = sphere_support(ny, nz, 4.2, rng, 1)
    plane = rng.standard_normal((nz, ny)) + 1j * rng.standard_normal((nz, ny))
    ref = np.fft.fft2(plane)[zf, yf]
    m = nz // 3
    w = np.exp(2j * np.pi / nz)
    rows, zb = active_rows(zf)
    zbt = [zb.get(k, -1) for k in range(nz)]
    buf = np.zeros((len(rows), ny), complex)
    for yl in range(ny):
        h = np.array([np.fft.fft(plane[t::3, yl]) * w ** (-t * np.arange(m)) for t in range(3)])
        for k0 in range(m):
            for t in range(3):
                # Lane 1 with its next and after-next factors swapped.
                own = 1.0 if t == 0 else C3 - 1j * S3
                f1 = {0: 1.0, 1: 1.0, 2: 1.0}[t]
                f2 = {0: 1.0, 1: C3 + 1j * S3, 2: C3 + 1j * S3}[t]
                x = own * h[t, k0] + f1 * h[(t + 1) % 3, k0] + f2 * h[(t + 2) % 3, k0]
                if zbt[k0 + t * m] >= 0:
                    buf[zbt[k0 + t * m], yl] = x
    got = np.fft.fft(buf, axis=1)[[zb[int(z)] for z in zf], yf]
    assert not np.allclose(got, ref, atol=1e-6 * np.max(np.abs(ref)))
