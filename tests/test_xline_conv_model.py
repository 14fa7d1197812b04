"""CPU model of the convolution x-line kernel (dft-zig_amd/csrc/hpsi.hip, k_xline_conv): the same
12-row LDS chunks with their XOR-swizzled index, the same (line, 16-row block) jobs over eight
wavefronts, the MFMA 16x16x4 lane layouts (A[i = l % 16][k = l / 16], B[k = l / 16][j = l % 16],
accumulator r of lane l at row l / 16 + 4 r in FP64 and 4 (l / 16) + r in FP32), the FP64 tail
of n_p mod 16 <= 8 rows as (line, 4-row group) jobs on the 4x4x4 MFMA (four blocks q = band
groups; A[i][k] in lane i + 4 q + 16 k, B[k][j] in lane j + 4 q + 16 k, D[i][j] in lane
16 i + 4 q + j, as measured on the device), the four-product complex form (three products,
Re = t1 - t2 and Im = t3 - t1 - t2, in the FP64 16-row blocks when there are at most two) and the
chunked write-back, checked against the x-line step written with FFTs: inverse transform of the present
frequencies x_s + p, times V/N, forward transform (numpy)."""
import numpy as np
import pytest

KCONV_BANDS = 16


def conv_idx(p, b, li, lt):
    return (p * KCONV_BANDS + b) * lt + (li ^ (b & (lt - 1)))


CHUNK = 12


def mfma_4x4x4(a_lanes, b_lanes):
    """One 4x4x4 MFMA (four blocks) from per-lane operands: D lane 16 i + 4 q + j."""
    d = np.zeros(64)
    for q in range(4):
        for i in range(4):
            for j in range(4):
                d[16 * i + 4 * q + j] = sum(a_lanes[i + 4 * q + 16 * k] * b_lanes[j + 4 * q + 16 * k] for k in range(4))
    return d


def model_kernel(tile, chat, np_, nx, lt, fp64):
    """tile: [np][bands=16][lt] complex rows of one workgroup; chat: [lt][nx]."""
    nrem = np_ & 15
    tail = fp64 and 0 < nrem <= 8
    nmt = np_ >> 4 if tail else (np_ + 15) >> 4
    ntjobs = lt * ((nrem + 3) >> 2) if tail else 0
    three = fp64 and nmt <= 2
    jobs, tjobs, t3s = {}, {}, {}
    for wave in range(8):
        for jj in range(lt * 3 // 8):
            job = wave + 8 * jj
            if job < lt * nmt:
                jobs[job] = [np.zeros((16, 16)), np.zeros((16, 16))]
                t3s[job] = np.zeros((16, 16))
        for tt in range(lt * 2 // 8):
            job = wave + 8 * tt
            if job < ntjobs:
                tjobs[job] = [np.zeros(64), np.zeros(64)]  # per-lane accumulators
    for k0 in range(0, np_, CHUNK):
        xs = np.zeros(CHUNK * KCONV_BANDS * lt, complex)
        for r in range(CHUNK):
            for b in range(KCONV_BANDS):
                for li in range(lt):
                    p = k0 + r
                    xs[conv_idx(r, b, li, lt)] = tile[p, b, li] if p < np_ else 0.0
        for job, (re, im) in jobs.items():
            li, mt = job % lt, job // lt
            for ks in range(0, CHUNK, 4):
                a = np.zeros((16, 4), complex)
                bm = np.zeros((4, 16), complex)
                for lane in range(64):
                    l16, l4 = lane & 15, lane >> 4
                    q, p = mt * 16 + l16, k0 + ks + l4
                    if q < np_ and p < np_:
                        a[l16, l4] = chat[li, (q - p) % nx]
                    bm[l4, l16] = xs[conv_idx(ks + l4, l16, li, lt)]
                if three:  # re holds t1, im t2, t3s[job] t3
                    re += a.real @ bm.real
                    im += a.imag @ bm.imag
                    t3s[job] += (a.real + a.imag) @ (bm.real + bm.imag)
                else:
                    re += a.real @ bm.real - a.imag @ bm.imag
                    im += a.real @ bm.imag + a.imag @ bm.real
        for job, (tre, tim) in tjobs.items():
            li, g = job % lt, job // lt
            for ks in range(0, CHUNK, 4):
                a = np.zeros(64, complex)
                bl = np.zeros(64, complex)
                for lane in range(64):
                    q, p = 16 * nmt + 4 * g + (lane & 3), k0 + ks + (lane >> 4)
                    if q < np_ and p < np_:
                        a[lane] = chat[li, (q - p) % nx]
                    bl[lane] = xs[conv_idx(ks + (lane >> 4), lane & 15, li, lt)]
                tre += mfma_4x4x4(a.real, bl.real) - mfma_4x4x4(a.imag, bl.imag)
                tim += mfma_4x4x4(a.real, bl.imag) + mfma_4x4x4(a.imag, bl.real)
    out = np.zeros((np_, KCONV_BANDS, lt), complex)
    for r0 in range(0, np_, CHUNK):
        xs = np.zeros(CHUNK * KCONV_BANDS * lt, complex)
        for job, (re, im) in jobs.items():
            li, mt = job % lt, job // lt
            # Read the accumulators out through the lane layout, as the kernel does.
            for lane in range(64):
                l16, l4 = lane & 15, lane >> 4
                for r in range(4):
                    row = l4 + 4 * r if fp64 else 4 * l4 + r
                    q = mt * 16 + row
                    if r0 <= q < r0 + CHUNK and q < np_:
                        if three:
                            t1, t2, t3 = re[row, l16], im[row, l16], t3s[job][row, l16]
                            xs[conv_idx(q - r0, l16, li, lt)] = (t1 - t2) + 1j * (t3 - t1 - t2)
                        else:
                            xs[conv_idx(q - r0, l16, li, lt)] = re[row, l16] + 1j * im[row, l16]
        for job, (tre, tim) in tjobs.items():
            li, g = job % lt, job // lt
            for lane in range(64):
                q = 16 * nmt + 4 * g + (lane >> 4)
                if r0 <= q < r0 + CHUNK and q < np_:
                    xs[conv_idx(q - r0, lane & 15, li, lt)] = tre[lane] + 1j * tim[lane]
        for r in range(CHUNK):
            for b in range(KCONV_BANDS):
                for li in range(lt):
                    if r0 + r < np_:
                        out[r0 + r, b, li] = xs[conv_idx(r, b, li, lt)]
    return out


@pytest.mark.parametrize("nx,np_,xs,fp64", [(75, 36, 57, True), (75, 36, 57, False), (30, 13, 24, True),
                                             (90, 45, 68, False), (32, 17, 0, True), (48, 48, 0, True),
                                             (80, 40, 60, True), (64, 38, 50, True), (90, 45, 68, True),
                                             (40, 6, 37, True)])
def test_conv_kernel_model_equals_fft_form(nx, np_, xs, fp64):
    lt = 8 if fp64 else 16
    rng = np.random.default_rng(nx + np_)
    v = rng.standard_normal((lt, nx)) / nx  # V/N on the workgroup's lines
    tile = rng.standard_normal((np_, KCONV_BANDS, lt)) + 1j * rng.standard_normal((np_, KCONV_BANDS, lt))
    # chat[l][d] = sum_x v(x) e^{-2 pi i d x / nx} (k_line_dft)
    x = np.arange(nx)
    chat = np.array([[np.sum(v[l] * np.exp(-2j * np.pi * d * x / nx)) for d in range(nx)] for l in range(lt)])
    got = model_kernel(tile, chat, np_, nx, lt, fp64)
    freqs = (xs + np.arange(np_)) % nx
    for li in range(lt):
        for b in range(KCONV_BANDS):
            full = np.zeros(nx, complex)
            full[freqs] = tile[:, b, li]
            psi = np.fft.ifft(full) * nx  # inverse transform without 1/N (V holds it)
            ref = np.fft.fft(psi * v[li])[freqs]
            np.testing.assert_allclose(got[:, b, li], ref, rtol=0, atol=1e-11 * max(1.0, np.abs(ref).max()))


def test_conv_swizzle_is_a_permutation():
    for lt in (8, 16):
        idx = {conv_idx(p, b, li, lt) for p in range(CHUNK) for b in range(KCONV_BANDS) for li in range(lt)}
        assert idx == set(range(CHUNK * KCONV_BANDS * lt))
        # The 16 bands of one (p, li) land in distinct 128-byte bank rows' slots.
        slot = 16 if lt == 8 else 8  # bytes per element (FP64 / FP32)
        for li in range(lt):
            pos = [(conv_idx(1, b, li, lt) * slot) % 128 // slot for b in range(KCONV_BANDS)]
            assert len(set(pos)) == min(KCONV_BANDS, 128 // slot)
