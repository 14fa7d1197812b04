// In-LDS batched mixed-radix complex FFT on the lines of a tile.
//
// Algorithm (own design): in-place Cooley-Tukey with register butterflies of radix
// 2, 3, 4, 5, 6, 8, 9 or 10 (6 and 10 as two-level register DFTs).  The decimation-in-time form
// takes its input at digit-reversed positions (AxisTables::perm) and leaves natural order; the decimation-in-frequency
// form takes natural order and leaves frequency f at position perm[f].  Each butterfly
// reads and writes the same R slots, so any assignment of butterflies to threads is
// race-free within a stage and one LDS tile suffices (an FP64 72x73 tile is 84 KB,
// 90x91 is 131 KB, of the 160 KB a gfx950 workgroup may hold).
//
// Thread mapping: consecutive threads take consecutive lines of the same butterfly, so
// a wavefront touches addresses line_stride apart; tiles are laid out with an odd row
// pitch so those accesses spread over the LDS banks.
//
// Sign conventions follow the reference: forward e^{-i}, inverse e^{+i}, both
// unnormalised here (src/lib/fft/mixed_radix.zig:47-61); callers apply 1/N.
#pragma once

#include "dzhip_internal.h"
#include "fft_reg.h"

#if defined(__HIP_DEVICE_COMPILE__)
#define DZ_BARRIER() __syncthreads()
#else
#define DZ_BARRIER() ((void)0)
#endif

namespace dzhip {

// q = w / d for 0 <= w < 2^17, 1 <= d <= 1024 (m = 0xFFFFFFFF / d + 1).
__host__ __device__ __forceinline__ unsigned div_magic(int d) { return d == 1 ? 0u : 0xFFFFFFFFu / (unsigned)d + 1u; }
__host__ __device__ __forceinline__ int fdiv(int w, int d, unsigned m) {
    return d == 1 ? w : (int)(((unsigned long long)(unsigned)w * m) >> 32);
}

template <typename T> struct RadixConst;
template <> struct RadixConst<double> {
    static constexpr double c3 = -0.5;
    static constexpr double s3 = 0.86602540378443864676;   // sin(2pi/3)
    static constexpr double c51 = 0.30901699437494742410;  // cos(2pi/5)
    static constexpr double c52 = -0.80901699437494742410; // cos(4pi/5)
    static constexpr double s51 = 0.95105651629515357212;  // sin(2pi/5)
    static constexpr double s52 = 0.58778525229247312917;  // sin(4pi/5)
    static constexpr double r2 = 0.70710678118654752440;   // 1/sqrt(2)
    static constexpr double c91 = 0.76604444311897803520;  // cos(2pi/9)
    static constexpr double s91 = 0.64278760968653932632;
    static constexpr double c92 = 0.17364817766693034885;  // cos(4pi/9)
    static constexpr double s92 = 0.98480775301220805936;
    static constexpr double c94 = -0.93969262078590838405; // cos(8pi/9)
    static constexpr double s94 = 0.34202014332566873304;
};
template <> struct RadixConst<float> {
    static constexpr float c3 = -0.5f;
    static constexpr float s3 = 0.86602540378443864676f;
    static constexpr float c51 = 0.30901699437494742410f;
    static constexpr float c52 = -0.80901699437494742410f;
    static constexpr float s51 = 0.95105651629515357212f;
    static constexpr float s52 = 0.58778525229247312917f;
    static constexpr float r2 = 0.70710678118654752440f;
    static constexpr float c91 = 0.76604444311897803520f;
    static constexpr float s91 = 0.64278760968653932632f;
    static constexpr float c92 = 0.17364817766693034885f;
    static constexpr float s92 = 0.98480775301220805936f;
    static constexpr float c94 = -0.93969262078590838405f;
    static constexpr float s94 = 0.34202014332566873304f;
};

// X_q = sum_r x_r exp(sign * 2 pi i r q / R), in registers.  sign is +1 or -1.
template <typename T, typename C>
__host__ __device__ __forceinline__ void dft2(C& a, C& b) {
    const C t = a;
    a = C{t.x + b.x, t.y + b.y};
    b = C{t.x - b.x, t.y - b.y};
}

template <typename T, typename C>
__host__ __device__ __forceinline__ void dft3(C& x0, C& x1, C& x2, T sign) {
    using K = RadixConst<T>;
    const C s = C{x1.x + x2.x, x1.y + x2.y};
    const C d = C{x1.x - x2.x, x1.y - x2.y};
    const C t = C{x0.x + K::c3 * s.x, x0.y + K::c3 * s.y};
    const T u = sign * K::s3;
    const C iu = C{-u * d.y, u * d.x};
    x0 = C{x0.x + s.x, x0.y + s.y};
    x1 = C{t.x + iu.x, t.y + iu.y};
    x2 = C{t.x - iu.x, t.y - iu.y};
}

template <typename T, typename C>
__host__ __device__ __forceinline__ void dft4(C& x0, C& x1, C& x2, C& x3, T sign) {
    const C a0 = C{x0.x + x2.x, x0.y + x2.y};
    const C a1 = C{x0.x - x2.x, x0.y - x2.y};
    const C b0 = C{x1.x + x3.x, x1.y + x3.y};
    const C b1 = C{x1.x - x3.x, x1.y - x3.y};
    const C ib1 = C{-sign * b1.y, sign * b1.x};
    x0 = C{a0.x + b0.x, a0.y + b0.y};
    x2 = C{a0.x - b0.x, a0.y - b0.y};
    x1 = C{a1.x + ib1.x, a1.y + ib1.y};
    x3 = C{a1.x - ib1.x, a1.y - ib1.y};
}

template <typename T, typename C>
__host__ __device__ __forceinline__ void dft5(C* v, T sign) {
    using K = RadixConst<T>;
    const C a1 = C{v[1].x + v[4].x, v[1].y + v[4].y}, b1 = C{v[1].x - v[4].x, v[1].y - v[4].y};
    const C a2 = C{v[2].x + v[3].x, v[2].y + v[3].y}, b2 = C{v[2].x - v[3].x, v[2].y - v[3].y};
    const C x0 = v[0];
    const C t1 = C{x0.x + K::c51 * a1.x + K::c52 * a2.x, x0.y + K::c51 * a1.y + K::c52 * a2.y};
    const C t2 = C{x0.x + K::c52 * a1.x + K::c51 * a2.x, x0.y + K::c52 * a1.y + K::c51 * a2.y};
    const T s1 = sign * K::s51, s2 = sign * K::s52;
    const C u1 = C{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
    const C u2 = C{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
    v[0] = C{x0.x + a1.x + a2.x, x0.y + a1.y + a2.y};
    v[1] = C{t1.x - u1.y, t1.y + u1.x};
    v[4] = C{t1.x + u1.y, t1.y - u1.x};
    v[2] = C{t2.x - u2.y, t2.y + u2.x};
    v[3] = C{t2.x + u2.y, t2.y - u2.x};
}

// 8 = 2 x 4: X[k] = E[k] + w8^k O[k], X[k+4] = E[k] - w8^k O[k].
template <typename T, typename C>
__host__ __device__ __forceinline__ void dft8(C* v, T sign) {
    using K = RadixConst<T>;
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<T>(e0, e1, e2, e3, sign);
    dft4<T>(o0, o1, o2, o3, sign);
    // w8^1 = (1 + s i)/sqrt2, w8^2 = s i, w8^3 = (-1 + s i)/sqrt2
    const C w1o = C{K::r2 * (o1.x - sign * o1.y), K::r2 * (o1.y + sign * o1.x)};
    const C w2o = C{-sign * o2.y, sign * o2.x};
    const C w3o = C{K::r2 * (-o3.x - sign * o3.y), K::r2 * (-o3.y + sign * o3.x)};
    v[0] = C{e0.x + o0.x, e0.y + o0.y};
    v[4] = C{e0.x - o0.x, e0.y - o0.y};
    v[1] = C{e1.x + w1o.x, e1.y + w1o.y};
    v[5] = C{e1.x - w1o.x, e1.y - w1o.y};
    v[2] = C{e2.x + w2o.x, e2.y + w2o.y};
    v[6] = C{e2.x - w2o.x, e2.y - w2o.y};
    v[3] = C{e3.x + w3o.x, e3.y + w3o.y};
    v[7] = C{e3.x - w3o.x, e3.y - w3o.y};
}

// 9 = 3 x 3: inputs j = 3 j1 + j2; 3-point DFTs over j1, twiddle w9^(j2 k1), 3-point over j2.
template <typename T, typename C>
__host__ __device__ __forceinline__ void dft9(C* v, T sign) {
    using K = RadixConst<T>;
    dft3<T>(v[0], v[3], v[6], sign);  // -> Y[k1][0] at v[0 + 3 k1]
    dft3<T>(v[1], v[4], v[7], sign);  // Y[k1][1] at v[1 + 3 k1]
    dft3<T>(v[2], v[5], v[8], sign);  // Y[k1][2] at v[2 + 3 k1]
    auto tw = [sign](C a, T c, T s) { const T ss = sign * s; return C{a.x * c - a.y * ss, a.x * ss + a.y * c}; };
    v[4] = tw(v[4], K::c91, K::s91);  // j2=1,k1=1: w^1
    v[7] = tw(v[7], K::c92, K::s92);  // j2=1,k1=2: w^2
    v[5] = tw(v[5], K::c92, K::s92);  // j2=2,k1=1: w^2
    v[8] = tw(v[8], K::c94, K::s94);  // j2=2,k1=2: w^4
    // X[k1 + 3 k2] = sum_j2 Y[k1][j2] w3^(j2 k2)
    C a[3] = {v[0], v[1], v[2]}, b[3] = {v[3], v[4], v[5]}, c[3] = {v[6], v[7], v[8]};
    dft3<T>(a[0], a[1], a[2], sign);
    dft3<T>(b[0], b[1], b[2], sign);
    dft3<T>(c[0], c[1], c[2], sign);
    v[0] = a[0]; v[3] = a[1]; v[6] = a[2];
    v[1] = b[0]; v[4] = b[1]; v[7] = b[2];
    v[2] = c[0]; v[5] = c[1]; v[8] = c[2];
}

template <typename T, typename C, int R>
__host__ __device__ __forceinline__ void dft_r(C* v, T sign);

// Composite radix N = N1 N2 in registers (Cooley-Tukey): inputs j = N2 j1 + j2, outputs
// k = k1 + N1 k2; N1-point DFTs over j1, twiddles w_N^(j2 k1) (compile-time constants), N2-point
// DFTs over j2.  Every index is a compile-time constant, so the final reordering is a renaming.
template <typename T, typename C, int N1, int N2>
__host__ __device__ __forceinline__ void dft_ct(C* v, T sign) {
    constexpr int N = N1 * N2;
#pragma unroll
    for (int j2 = 0; j2 < N2; ++j2) {
        C t[N1];
#pragma unroll
        for (int j1 = 0; j1 < N1; ++j1) t[j1] = v[N2 * j1 + j2];
        dft_r<T, C, N1>(t, sign);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) v[N2 * k1 + j2] = t[k1];
    }
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1)
#pragma unroll
        for (int j2 = 1; j2 < N2; ++j2) {
            const freg::CosSin w = freg::unit_root((long)j2 * k1, N);
            const T c = T(w.c), sn = sign * T(w.s);
            C& a = v[N2 * k1 + j2];
            a = C{a.x * c - a.y * sn, a.x * sn + a.y * c};
        }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) dft_r<T, C, N2>(v + N2 * k1, sign);
    C out[N];
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) out[k1 + N1 * k2] = v[N2 * k1 + k2];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = out[k];
}

template <typename T, typename C, int R>
__host__ __device__ __forceinline__ void dft_r(C* v, T sign) {
    if constexpr (R == 2) dft2<T>(v[0], v[1]);
    if constexpr (R == 3) dft3<T>(v[0], v[1], v[2], sign);
    if constexpr (R == 4) dft4<T>(v[0], v[1], v[2], v[3], sign);
    if constexpr (R == 5) dft5<T>(v, sign);
    if constexpr (R == 6) dft_ct<T, C, 2, 3>(v, sign);
    if constexpr (R == 8) dft8<T>(v, sign);
    if constexpr (R == 9) dft9<T>(v, sign);
    if constexpr (R == 10) dft_ct<T, C, 2, 5>(v, sign);
}

// w^t for the transform sign S (+1 inverse, -1 forward): the table holds exp(-2 pi i t / n).
template <int S, typename C>
__host__ __device__ __forceinline__ C tw_get(const C* tw, int t) {
    C w = tw[t];
    if (S > 0) w.y = -w.y;
    return w;
}

template <typename C>
__host__ __device__ __forceinline__ C cmul_hd(C a, C b) {
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Lines: line l starts at buf + base(l), element e at + e*es, where base(l) = l*ls, or
// lines[l] when an explicit list of line offsets is given (pruned transforms).
struct LineSet {
    int nlines;
    int es;
    int ls;
    const short* lines;  // nullptr: uniform stride
    unsigned mlines;     // div_magic(nlines), computed once by the constructor
    __host__ __device__ __forceinline__ LineSet(int n, int e, int l, const short* ln)
        : nlines(n), es(e), ls(l), lines(ln), mlines(div_magic(n)) {}
    __host__ __device__ __forceinline__ int base(int l) const { return lines ? (int)lines[l] : l * ls; }
};

// One DIT stage (span Lp -> L = Lp*R) or DIF stage (span L -> L/R).
template <typename T, typename C, int R, bool DIT, int S>
__host__ __device__ __forceinline__ void fft_stage(C* buf, int n, int span, const LineSet& ls,
                                                   const C* tw, unsigned m_sub, int tid, int nthr) {
    const T sign = T(S);
    const int sub = DIT ? span : span / R;  // distance between the R inputs of a butterfly (in elements)
    const int L = DIT ? span * R : span;
    const int tstep = n / L;
    const int nbf = n / R;
    const int total = ls.nlines * nbf;
    const unsigned m_lines = ls.mlines;
    for (int w = tid; w < total; w += nthr) {
        const int j = fdiv(w, ls.nlines, m_lines);
        const int line = w - j * ls.nlines;
        const int g = fdiv(j, sub, m_sub);
        const int k = j - g * sub;
        C* p = buf + ls.base(line) + (g * L + k) * ls.es;
        const int st = sub * ls.es;
        C v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = p[r * st];
        if (DIT && k != 0) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul_hd(v[r], tw_get<S>(tw, k * r * tstep));
        }
        dft_r<T, C, R>(v, sign);
        if (!DIT && k != 0) {
#pragma unroll
            for (int q = 1; q < R; ++q) v[q] = cmul_hd(v[q], tw_get<S>(tw, k * q * tstep));
        }
#pragma unroll
        for (int r = 0; r < R; ++r) p[r * st] = v[r];
    }
}

template <typename T, typename C, bool DIT, int S>
__host__ __device__ __forceinline__ void fft_stage_dispatch(int R, C* buf, int n, int span,
                                                            const LineSet& ls, const C* tw, unsigned m_sub,
                                                            int tid, int nthr) {
    switch (R) {
        case 2: fft_stage<T, C, 2, DIT, S>(buf, n, span, ls, tw, m_sub, tid, nthr); break;
        case 3: fft_stage<T, C, 3, DIT, S>(buf, n, span, ls, tw, m_sub, tid, nthr); break;
        case 4: fft_stage<T, C, 4, DIT, S>(buf, n, span, ls, tw, m_sub, tid, nthr); break;
        case 5: fft_stage<T, C, 5, DIT, S>(buf, n, span, ls, tw, m_sub, tid, nthr); break;
        case 6: fft_stage<T, C, 6, DIT, S>(buf, n, span, ls, tw, m_sub, tid, nthr); break;
        case 8: fft_stage<T, C, 8, DIT, S>(buf, n, span, ls, tw, m_sub, tid, nthr); break;
        case 10: fft_stage<T, C, 10, DIT, S>(buf, n, span, ls, tw, m_sub, tid, nthr); break;
        default: fft_stage<T, C, 9, DIT, S>(buf, n, span, ls, tw, m_sub, tid, nthr); break;
    }
}

// Full in-place transform of the lines, sign S (+1 inverse, -1 forward).  Ends with a barrier.
// DIT: input digit-reversed -> natural output.  DIF: natural input -> digit-reversed.
template <typename T, typename C, bool DIT, int S>
__host__ __device__ void lds_fft_lines(C* buf, const FftAxis& ax, const C* tw, const LineSet& ls,
                                       int tid, int nthr) {
    const int n = ax.n;
    if (DIT) {
        int Lp = 1;
        for (int s = 0; s < ax.nstages; ++s) {
            const int R = ax.radix[s];
            fft_stage_dispatch<T, C, true, S>(R, buf, n, Lp, ls, tw, ax.sub_magic[s], tid, nthr);
            Lp *= R;
            DZ_BARRIER();
        }
    } else {
        int L = n;
        for (int s = ax.nstages - 1; s >= 0; --s) {
            const int R = ax.radix[s];
            fft_stage_dispatch<T, C, false, S>(R, buf, n, L, ls, tw, ax.sub_magic[s], tid, nthr);
            L /= R;
            DZ_BARRIER();
        }
    }
}

// Copy an axis' twiddle table into LDS.
template <typename C>
__device__ __forceinline__ void load_twiddles(C* dst, const C* __restrict__ src, int n, int tid, int nthr) {
    for (int i = tid; i < n; i += nthr) dst[i] = src[i];
}

template <typename T> struct TwSel;
template <> struct TwSel<double> {
    static __host__ __device__ const double2* get(const double2* t64, const float2*) { return t64; }
};
template <> struct TwSel<float> {
    static __host__ __device__ const float2* get(const double2*, const float2* t32) { return t32; }
};

// Threads per workgroup for the tile kernels: FP64 tiles are twice as large, so fewer fit
// per CU; more threads per workgroup keep the CU occupied.
template <typename T> struct TileThreads { static constexpr int value = 512; };
template <> struct TileThreads<double> { static constexpr int value = 1024; };
// Global loads each thread issues before its LDS stores in the tile kernels' load loops.  FP32
// tiles run 512 threads and gain from 8 in flight; FP64 tiles run 1024 threads at a 128-VGPR
// budget, where 8 double2 loads per thread raise register pressure more than they hide latency.
template <typename T> struct LoadBatch { static constexpr int value = 8; };
template <> struct LoadBatch<double> { static constexpr int value = 2; };

}  // namespace dzhip
