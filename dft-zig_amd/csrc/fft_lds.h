// In-LDS batched mixed-radix (2/3/4/5) complex FFT on lines of a tile.
//
// Algorithm (own design): in-place Cooley-Tukey.  The decimation-in-time form
// takes its input at digit-reversed LDS positions (AxisTables::perm) and leaves
// natural order; the decimation-in-frequency form takes natural order and
// leaves frequency f at position perm[f].  Each butterfly reads and writes the
// same R slots, so any assignment of butterflies to threads is race-free within
// a stage and one LDS tile suffices (an FP64 72x72 plane is 83 KB, 90x90 is
// 130 KB, of the 160 KB a gfx950 workgroup may hold).
//
// Sign conventions follow the reference: forward e^{-i}, inverse e^{+i}, both
// unnormalised here (src/lib/fft/mixed_radix.zig:47-61); callers apply 1/N.
#pragma once

#include "dzhip_internal.h"

namespace dzhip {

template <typename T> struct RadixConst;
template <> struct RadixConst<double> {
    static constexpr double c3 = -0.5;
    static constexpr double s3 = 0.86602540378443864676;  // sin(2pi/3)
    static constexpr double c51 = 0.30901699437494742410; // cos(2pi/5)
    static constexpr double c52 = -0.80901699437494742410; // cos(4pi/5)
    static constexpr double s51 = 0.95105651629515357212; // sin(2pi/5)
    static constexpr double s52 = 0.58778525229247312917; // sin(4pi/5)
};
template <> struct RadixConst<float> {
    static constexpr float c3 = -0.5f;
    static constexpr float s3 = 0.86602540378443864676f;
    static constexpr float c51 = 0.30901699437494742410f;
    static constexpr float c52 = -0.80901699437494742410f;
    static constexpr float s51 = 0.95105651629515357212f;
    static constexpr float s52 = 0.58778525229247312917f;
};

// X_q = sum_r x_r exp(sign * 2 pi i r q / R), in registers.  sign is +1 or -1.
template <typename T, typename C>
__device__ __forceinline__ void dft2(C* v) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename T, typename C>
__device__ __forceinline__ void dft3(C* v, T sign) {
    using K = RadixConst<T>;
    C s = cadd(v[1], v[2]);
    C d = csub(v[1], v[2]);
    C t = C{v[0].x + K::c3 * s.x, v[0].y + K::c3 * s.y};
    T u = sign * K::s3;
    C iu = C{-u * d.y, u * d.x};  // i * u * d
    v[0] = cadd(v[0], s);
    v[1] = cadd(t, iu);
    v[2] = csub(t, iu);
}

template <typename T, typename C>
__device__ __forceinline__ void dft4(C* v, T sign) {
    C a0 = cadd(v[0], v[2]);
    C a1 = csub(v[0], v[2]);
    C b0 = cadd(v[1], v[3]);
    C b1 = csub(v[1], v[3]);
    C ib1 = C{-sign * b1.y, sign * b1.x};  // i * sign * (x1 - x3)
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, ib1);
    v[3] = csub(a1, ib1);
}

template <typename T, typename C>
__device__ __forceinline__ void dft5(C* v, T sign) {
    using K = RadixConst<T>;
    C a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    C a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    C x0 = v[0];
    C t1 = C{x0.x + K::c51 * a1.x + K::c52 * a2.x, x0.y + K::c51 * a1.y + K::c52 * a2.y};
    C t2 = C{x0.x + K::c52 * a1.x + K::c51 * a2.x, x0.y + K::c52 * a1.y + K::c51 * a2.y};
    T s1 = sign * K::s51, s2 = sign * K::s52;
    C u1 = C{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
    C u2 = C{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
    C iu1 = C{-u1.y, u1.x}, iu2 = C{-u2.y, u2.x};
    v[0] = C{x0.x + a1.x + a2.x, x0.y + a1.y + a2.y};
    v[1] = cadd(t1, iu1);
    v[4] = csub(t1, iu1);
    v[2] = cadd(t2, iu2);
    v[3] = csub(t2, iu2);
}

template <typename T, typename C, int R>
__device__ __forceinline__ void dft_r(C* v, T sign) {
    if constexpr (R == 2) dft2<T>(v);
    if constexpr (R == 3) dft3<T>(v, sign);
    if constexpr (R == 4) dft4<T>(v, sign);
    if constexpr (R == 5) dft5<T>(v, sign);
}

// w^t for the transform sign: twiddle table holds exp(-2 pi i t / n).
template <typename C, typename T>
__device__ __forceinline__ C tw_get(const C* __restrict__ tw, int t, T sign) {
    C w = tw[t];
    if (sign > T(0)) w.y = -w.y;
    return w;
}

// One DIT stage of radix R with sub-transform span Lp -> L = Lp*R on `nlines`
// lines; element e of line l lives at buf[l*line_stride + e*elem_stride].
template <typename T, typename C, int R>
__device__ __forceinline__ void dit_stage(C* buf, int n, int Lp, int nlines, int elem_stride,
                                          int line_stride, const C* __restrict__ tw, T sign,
                                          int tid, int nthr) {
    const int L = Lp * R;
    const int nbf = n / R;
    const int tstep = n / L;
    const int total = nlines * nbf;
    for (int w = tid; w < total; w += nthr) {
        const int line = w / nbf;
        const int j = w - line * nbf;
        const int g = j / Lp;
        const int k = j - g * Lp;
        C* p = buf + line * line_stride + (g * L + k) * elem_stride;
        C v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = p[r * Lp * elem_stride];
        if (k != 0) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw_get(tw, k * r * tstep, sign));
        }
        dft_r<T, C, R>(v, sign);
#pragma unroll
        for (int r = 0; r < R; ++r) p[r * Lp * elem_stride] = v[r];
    }
}

// One DIF stage of radix R with span L -> Lq = L/R.
template <typename T, typename C, int R>
__device__ __forceinline__ void dif_stage(C* buf, int n, int L, int nlines, int elem_stride,
                                          int line_stride, const C* __restrict__ tw, T sign,
                                          int tid, int nthr) {
    const int Lq = L / R;
    const int nbf = n / R;
    const int tstep = n / L;
    const int total = nlines * nbf;
    for (int w = tid; w < total; w += nthr) {
        const int line = w / nbf;
        const int j = w - line * nbf;
        const int g = j / Lq;
        const int k = j - g * Lq;
        C* p = buf + line * line_stride + (g * L + k) * elem_stride;
        C v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = p[r * Lq * elem_stride];
        dft_r<T, C, R>(v, sign);
        if (k != 0) {
#pragma unroll
            for (int q = 1; q < R; ++q) v[q] = cmul(v[q], tw_get(tw, k * q * tstep, sign));
        }
#pragma unroll
        for (int q = 0; q < R; ++q) p[q * Lq * elem_stride] = v[q];
    }
}

// Full in-place transform of `nlines` lines in LDS.  Ends with a barrier.
// DIT: input digit-reversed -> natural output.  DIF: natural input -> digit-reversed.
template <typename T, typename C, bool DIT>
__device__ void lds_fft_lines(C* buf, const FftAxis& ax, const C* __restrict__ tw, int nlines,
                              int elem_stride, int line_stride, T sign, int tid, int nthr) {
    const int n = ax.n;
    if (DIT) {
        int Lp = 1;
        for (int s = 0; s < ax.nstages; ++s) {
            const int R = ax.radix[s];
            switch (R) {
                case 2: dit_stage<T, C, 2>(buf, n, Lp, nlines, elem_stride, line_stride, tw, sign, tid, nthr); break;
                case 3: dit_stage<T, C, 3>(buf, n, Lp, nlines, elem_stride, line_stride, tw, sign, tid, nthr); break;
                case 4: dit_stage<T, C, 4>(buf, n, Lp, nlines, elem_stride, line_stride, tw, sign, tid, nthr); break;
                default: dit_stage<T, C, 5>(buf, n, Lp, nlines, elem_stride, line_stride, tw, sign, tid, nthr); break;
            }
            Lp *= R;
            __syncthreads();
        }
    } else {
        int L = n;
        for (int s = ax.nstages - 1; s >= 0; --s) {
            const int R = ax.radix[s];
            switch (R) {
                case 2: dif_stage<T, C, 2>(buf, n, L, nlines, elem_stride, line_stride, tw, sign, tid, nthr); break;
                case 3: dif_stage<T, C, 3>(buf, n, L, nlines, elem_stride, line_stride, tw, sign, tid, nthr); break;
                case 4: dif_stage<T, C, 4>(buf, n, L, nlines, elem_stride, line_stride, tw, sign, tid, nthr); break;
                default: dif_stage<T, C, 5>(buf, n, L, nlines, elem_stride, line_stride, tw, sign, tid, nthr); break;
            }
            L /= R;
            __syncthreads();
        }
    }
}

template <typename T> struct TwSel;
template <> struct TwSel<double> {
    static __host__ __device__ const double2* get(const double2* t64, const float2*) { return t64; }
};
template <> struct TwSel<float> {
    static __host__ __device__ const float2* get(const double2*, const float2* t32) { return t32; }
};

}  // namespace dzhip
