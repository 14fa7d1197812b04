// Register-resident FFT of one line per thread, for line lengths known at compile time.
//
// Mixed-radix Stockham autosort (decimation in time): stage s with radix R and span Ns (the
// product of the earlier radices) reads a[j + r N/R], multiplies by w^(r (j mod Ns) N/(Ns R)),
// does a radix-R butterfly and writes b[(j / Ns) Ns R + (j mod Ns) + r Ns]; natural order in
// and out.  Radices are taken as 8s, then 9s, 4, 2, 3, 5.  Every index is a compile-time
// constant after unrolling, so the line lives in VGPRs and no LDS or barrier is involved; the
// twiddles w^k = e^{-2 pi i k / N} come from a table with compile-time offsets (wave-uniform
// loads).  Sign S = -1 is the forward transform e^{-i...}, S = +1 the unnormalised inverse.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace dzhip {
namespace freg {

constexpr int next_radix(int n) {
    return n % 8 == 0 ? 8 : (n % 9 == 0 ? 9 : (n % 4 == 0 ? 4 : (n % 2 == 0 ? 2 : (n % 3 == 0 ? 3 : 5))));
}

constexpr bool supported(int n) {
    if (n < 2) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    return n == 1;
}

constexpr int num_stages(int n) {
    int c = 0;
    while (n > 1) {
        n /= next_radix(n);
        ++c;
    }
    return c;
}

constexpr int radix_at(int n, int k) {
    for (int i = 0; i < k; ++i) n /= next_radix(n);
    return next_radix(n);
}

// cos and sin of 2 pi k / n as compile-time constants: the angle is reduced exactly in integers to
// a quadrant and to [0, pi/4] (by the complement), where the Taylor series of sin and cos, summed
// in Horner form to x^29, are exact to a few units in the last place of a double.
struct CosSin {
    double c, s;
};

constexpr double sin_series(double x) {
    double r = 1.0;
    for (int n = 14; n >= 1; --n) r = 1.0 - x * x / ((2.0 * n) * (2.0 * n + 1.0)) * r;
    return x * r;
}

constexpr double cos_series(double x) {
    double r = 1.0;
    for (int n = 14; n >= 1; --n) r = 1.0 - x * x / ((2.0 * n - 1.0) * (2.0 * n)) * r;
    return r;
}

constexpr CosSin unit_root(long k, long n) {
    k = ((k % n) + n) % n;
    const long quad = 4 * k / n, rem = 4 * k - quad * n;  // 2 pi k / n = (pi / 2) (quad + rem / n)
    const bool comp = 2 * rem > n;
    const double x = 1.57079632679489661923 * (double)(comp ? n - rem : rem) / (double)n;
    double s = sin_series(x), c = cos_series(x);
    if (comp) {
        const double t = s;
        s = c;
        c = t;
    }
    switch (quad) {
        case 0: return CosSin{c, s};
        case 1: return CosSin{-s, c};
        case 2: return CosSin{-c, -s};
        default: return CosSin{s, -c};
    }
}

template <int N>
struct RootTable {
    double c[N], s[N];
    constexpr RootTable() : c{}, s{} {
        for (int k = 0; k < N; ++k) {
            const CosSin v = unit_root(k, N);
            c[k] = v.c;
            s[k] = v.s;
        }
    }
};

// cos / sin of 2 pi k / N, k < N; indexed with indices that are constant after unrolling, the
// loads fold into literal operands.
template <int N>
inline constexpr RootTable<N> kRoots{};

template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C, typename T>
__device__ __forceinline__ C cscale(C a, T s) { return C{a.x * s, a.y * s}; }
// i^(S) * a with S = +-1: +1 -> i a, -1 -> -i a
template <int S, typename C>
__device__ __forceinline__ C cmuli(C a) { return S > 0 ? C{-a.y, a.x} : C{a.y, -a.x}; }
template <typename C>
__device__ __forceinline__ C cmul(C a, C w) { return C{a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x}; }

// In-place DFT of size R with kernel e^{S 2 pi i q s / R}.
template <int R, int S, typename C>
__device__ __forceinline__ void butterfly(C (&v)[R]) {
    using T = decltype(v[0].x);
    if constexpr (R == 2) {
        const C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const C s13 = cadd(v[1], v[3]), d13 = cmuli<S>(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    } else if constexpr (R == 3) {
        constexpr T c = T(-0.5);
        constexpr T sn = T(0.86602540378443864676);  // sin(2 pi / 3)
        const C t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        const C m = C{v[0].x + c * t1.x, v[0].y + c * t1.y};
        const C n = cscale(cmuli<S>(t2), sn);
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, n);
        v[2] = csub(m, n);
    } else if constexpr (R == 8) {
        // Two radix-4 transforms of the even and odd samples, odd ones turned by w8^(S k).
        C e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        butterfly<4, S>(e);
        butterfly<4, S>(o);
        constexpr T h = T(0.70710678118654752440);  // 1 / sqrt(2)
        const C o1 = C{h * (o[1].x - S * o[1].y), h * (S * o[1].x + o[1].y)};
        const C o2 = cmuli<S>(o[2]);
        const C o3 = C{h * (-o[3].x - S * o[3].y), h * (S * o[3].x - o[3].y)};
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    } else if constexpr (R == 9) {
        // 3 x 3: radix-3 over n1 for each n2 (samples n2 + 3 n1), twiddle w9^(S n2 k1), radix-3
        // over n2; output k1 + 3 k2.
        C t[3][3];
#pragma unroll
        for (int n2 = 0; n2 < 3; ++n2) {
            C u[3] = {v[n2], v[n2 + 3], v[n2 + 6]};
            butterfly<3, S>(u);
            t[n2][0] = u[0];
            t[n2][1] = u[1];
            t[n2][2] = u[2];
        }
        constexpr T c1 = T(0.76604444311897803520), s1 = T(0.64278760968653932632);   // 40 deg
        constexpr T c2 = T(0.17364817766693034885), s2 = T(0.98480775301220805936);   // 80 deg
        constexpr T c4 = T(-0.93969262078590838405), s4 = T(0.34202014332566873304);  // 160 deg
        t[1][1] = cmul(t[1][1], C{c1, S * s1});
        t[1][2] = cmul(t[1][2], C{c2, S * s2});
        t[2][1] = cmul(t[2][1], C{c2, S * s2});
        t[2][2] = cmul(t[2][2], C{c4, S * s4});
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) {
            C u[3] = {t[0][k1], t[1][k1], t[2][k1]};
            butterfly<3, S>(u);
            v[k1] = u[0];
            v[k1 + 3] = u[1];
            v[k1 + 6] = u[2];
        }
    } else {
        static_assert(R == 5, "radix");
        constexpr T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        constexpr T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
        const C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        const C t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        const C m1 = C{v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
        const C m2 = C{v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
        const C n1 = cmuli<S>(C{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y});
        const C n2 = cmuli<S>(C{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y});
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
}

template <typename C, int N, int R, int Ns, int S, int TS>
__device__ __forceinline__ void stage(const C (&in)[N], C (&out)[N], const C* __restrict__ tw) {
#pragma unroll
    for (int j = 0; j < N / R; ++j) {
        C v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = in[j + r * (N / R)];
        const int k = j % Ns;
        if (k != 0) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const int e = (r * k * (N / (Ns * R))) % N;
                C w;
                if constexpr (TS == 0) {
                    using T = decltype(w.x);
                    w = C{T(kRoots<N>.c[e]), T(S * kRoots<N>.s[e])};
                } else {
                    w = tw[e * TS];
                    if (S > 0) w.y = -w.y;
                }
                v[r] = cmul(v[r], w);
            }
        }
        butterfly<R, S>(v);
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) out[base + r * Ns] = v[r];
    }
}

template <typename C, int N, int S, int K, int Ns, int TS>
__device__ __forceinline__ void stages(C (&a)[N], const C* __restrict__ tw) {
    if constexpr (K < num_stages(N)) {
        constexpr int R = radix_at(N, K);
        C b[N];
        stage<C, N, R, Ns, S, TS>(a, b, tw);
#pragma unroll
        for (int i = 0; i < N; ++i) a[i] = b[i];
        stages<C, N, S, K + 1, Ns * R, TS>(a, tw);
    }
}

// a <- DFT_S(a); tw[k * TS] = e^{-2 pi i k / N}, k < N (TS > 1 reads the table of a multiple
// length N TS).
template <int N, int S, int TS = 1, typename C>
__device__ __forceinline__ void fft(C (&a)[N], const C* __restrict__ tw) {
    static_assert(supported(N), "line length must be 2^a 3^b 5^c");
    stages<C, N, S, 0, 1, TS>(a, tw);
}

// a <- DFT_S(a) with the twiddles as compile-time constants (kRoots).
template <int N, int S, typename C>
__device__ __forceinline__ void fft(C (&a)[N]) {
    static_assert(supported(N), "line length must be 2^a 3^b 5^c");
    stages<C, N, S, 0, 1, 0>(a, nullptr);
}

}  // namespace freg
}  // namespace dzhip
