"""Python model of the in-LDS FFT the HIP kernels run (dft-zig_amd/csrc/fft_lds.h, fft.hip):
the same radix factorisation, digit-reversal table and in-place DIT/DIF stage loops, checked
against numpy.  Used by tests only."""
import numpy as np


def factor(n):
    """factor_fft_len: register radices 9, 8, 10, 6, 4, 2, 3, 5 in that order."""
    out, m = [], n
    for r in (9, 8, 10, 6, 4, 2, 3, 5):
        while m % r == 0:
            out.append(r)
            m //= r
    assert m == 1, n
    return out


def digit_perm(n, rad):
    def pos(i, n, rad):
        if not rad:
            return 0
        r = rad[-1]
        m = n // r
        return (i % r) * m + pos(i // r, m, rad[:-1])

    return np.array([pos(i, n, rad) for i in range(n)])


def dit(buf, n, rad, sign):
    lp = 1
    for r in rad:
        l = lp * r
        for g in range(n // l):
            for k in range(lp):
                idx = [g * l + k + q * lp for q in range(r)]
                y = [buf[idx[q]] * np.exp(sign * 2j * np.pi * k * q / l) for q in range(r)]
                for q in range(r):
                    buf[idx[q]] = sum(y[s] * np.exp(sign * 2j * np.pi * s * q / r) for s in range(r))
        lp = l
    return buf


def dif(buf, n, rad, sign):
    l = n
    for r in reversed(rad):
        lq = l // r
        for g in range(n // l):
            for k in range(lq):
                idx = [g * l + k + q * lq for q in range(r)]
                x = [buf[i] for i in idx]
                for q in range(r):
                    buf[idx[q]] = sum(x[s] * np.exp(sign * 2j * np.pi * s * q / r) for s in range(r)) \
                        * np.exp(sign * 2j * np.pi * k * q / l)
        l = lq
    return buf


def check_sizes(sizes):
    rng = np.random.default_rng(0)
    for n in sizes:
        rad = factor(n)
        p = digit_perm(n, rad)
        assert sorted(p.tolist()) == list(range(n))
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        b = np.zeros(n, complex)
        b[p] = x
        assert np.allclose(dit(b.copy(), n, rad, -1.0), np.fft.fft(x))
        assert np.allclose(dit(b.copy(), n, rad, +1.0), np.fft.ifft(x) * n)
        assert np.allclose(dif(x.copy(), n, rad, -1.0)[p], np.fft.fft(x))
        assert np.allclose(dif(x.copy(), n, rad, +1.0)[p], np.fft.ifft(x) * n)
