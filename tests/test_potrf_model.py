"""CPU model of the one-workgroup Cholesky factorisation (dft-zig_amd/csrc/eig.hip, k_potrf_upper):
the same 16-row blocks, the diagonal block with lane l holding column l (row j broadcast from lane
i by shuffles), the row panel by forward substitution with the block's R, and the trailing update
over the upper-triangular enumeration of 4 x 4 tiles, checked against numpy.linalg.cholesky (which
is LAPACK zpotrf, the routine the eigensolver called before) and for LAPACK's info on a matrix
that is not positive definite."""
import math

import numpy as np
import pytest

NB = 16  # rows per block (kPotrfNb)


def tiles(nt):
    """(ti, tc) of tile index e = tc (tc + 1) / 2 + ti, ti <= tc, as the kernel inverts it."""
    out = []
    for e in range(nt * (nt + 1) // 2):
        tc = int((math.sqrt(8.0 * e + 1.0) - 1.0) * 0.5)
        while tc * (tc + 1) // 2 > e:
            tc -= 1
        while (tc + 1) * (tc + 2) // 2 <= e:
            tc += 1
        out.append((e - tc * (tc + 1) // 2, tc))
    return out


def model(a_in, k):
    """(matrix with R in its upper triangle, info) as k_potrf_upper leaves them."""
    a = a_in.copy()
    for jb in range(0, k, NB):
        nb = min(NB, k - jb)
        # 1. Diagonal block: lane l holds col[l][i] = A[jb + i][jb + l] for i <= l.
        col = np.zeros((64, NB), complex)
        for lane in range(nb):
            for i in range(lane + 1):
                col[lane, i] = a[jb + i, jb + lane]
        for j in range(nb):
            d = col[j, j].real  # shuffled from lane j
            if not d > 0:
                return a, jb + j + 1
            r = math.sqrt(d)
            col[j, j] = r
            col[j + 1:, j] /= r  # R[j][l] on lanes l > j
            for i in range(j + 1, NB):
                r_ji = col[i, j]  # shuffled from lane i
                for lane in range(i, nb):
                    col[lane, i] -= np.conj(r_ji) * col[lane, j]
        d_blk = np.zeros((NB, NB), complex)
        for lane in range(nb):
            for i in range(lane + 1):
                a[jb + i, jb + lane] = col[lane, i]
                d_blk[i, lane] = col[lane, i]
        c0 = jb + nb
        m = k - c0
        if m <= 0:
            break
        # 2. Rows of R right of the block: R_d^H y = A[jb:jb+nb, c], one thread per column.
        rows = np.zeros((NB, m), complex)
        for t_col in range(m):
            c = c0 + t_col
            y = np.zeros(NB, complex)
            for i in range(nb):
                v = a[jb + i, c] - sum(np.conj(d_blk[t, i]) * y[t] for t in range(i))
                y[i] = v / d_blk[i, i].real
                a[jb + i, c] = y[i]
                rows[i, t_col] = y[i]
        # 3. Trailing upper triangle in 4 x 4 tiles.
        for ti, tc in tiles((m + 3) // 4):
            for b in range(4):
                c = 4 * tc + b
                if c >= m:
                    continue
                for a_row in range(4):
                    i = 4 * ti + a_row
                    if i > c:
                        continue
                    a[c0 + i, c0 + c] -= sum(np.conj(rows[t, i]) * rows[t, c] for t in range(nb))
    return a, 0


@pytest.mark.parametrize("nt", [1, 2, 7, 41, 60])
def test_tile_enumeration_covers_upper_triangle_once(nt):
    assert sorted(tiles(nt)) == sorted((i, c) for c in range(nt) for i in range(c + 1))


@pytest.mark.parametrize("k", [1, 5, 16, 17, 40, 70])
def test_blocked_cholesky_model_equals_lapack(k):
    rng = np.random.default_rng(k)
    x = rng.standard_normal((k + 5, k)) + 1j * rng.standard_normal((k + 5, k))
    g = x.conj().T @ x
    a, info = model(g, k)
    assert info == 0
    np.testing.assert_allclose(np.triu(a), np.linalg.cholesky(g).conj().T, rtol=0, atol=1e-10 * np.abs(g).max())
    np.testing.assert_array_equal(np.tril(a, -1), np.tril(g, -1))  # the lower triangle is left alone


def test_blocked_cholesky_model_reports_first_bad_pivot():
    g = np.eye(20, dtype=complex)
    g[7, 7] = -1.0
    assert model(g, 20)[1] == 8
