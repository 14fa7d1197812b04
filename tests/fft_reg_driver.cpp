// Host build of the register FFT (dft-zig_amd/csrc/fft_reg.h) for the CPU tests: reads n, sign
// and n complex values (double) on stdin, writes the transform in double and in float with the
// twiddle table, then in double and in float with the compile-time twiddles; "n 0" writes the
// compile-time table cos, sin of 2 pi k / n instead.
#include <cmath>
#include <cstdio>
#include <vector>

#define __device__
#define __forceinline__ inline
struct float2 {
    float x, y;
};
struct double2 {
    double x, y;
};
#include "fft_reg.h"

template <int N, typename C, typename T>
static void run(int sign, const std::vector<double>& in, const std::vector<C>& tw) {
    C a[N];
    for (int i = 0; i < N; ++i) a[i] = C{(T)in[2 * i], (T)in[2 * i + 1]};
    if (sign > 0)
        dzhip::freg::fft<N, 1>(a, tw.data());
    else
        dzhip::freg::fft<N, -1>(a, tw.data());
    for (int i = 0; i < N; ++i) std::printf("%.17g %.17g\n", (double)a[i].x, (double)a[i].y);
}

template <int N, typename C, typename T>
static void run_const(int sign, const std::vector<double>& in) {
    C a[N];
    for (int i = 0; i < N; ++i) a[i] = C{(T)in[2 * i], (T)in[2 * i + 1]};
    if (sign > 0)
        dzhip::freg::fft<N, 1>(a);
    else
        dzhip::freg::fft<N, -1>(a);
    for (int i = 0; i < N; ++i) std::printf("%.17g %.17g\n", (double)a[i].x, (double)a[i].y);
}

template <int N>
static void both(int sign, const std::vector<double>& in) {
    std::vector<double2> tw64(N);
    std::vector<float2> tw32(N);
    for (int k = 0; k < N; ++k) {
        const double ang = -2.0 * M_PI * k / N;
        tw64[k] = double2{std::cos(ang), std::sin(ang)};
        tw32[k] = float2{(float)std::cos(ang), (float)std::sin(ang)};
    }
    run<N, double2, double>(sign, in, tw64);
    run<N, float2, float>(sign, in, tw32);
    run_const<N, double2, double>(sign, in);
    run_const<N, float2, float>(sign, in);
}

template <int N>
static void roots() {
    for (int k = 0; k < N; ++k)
        std::printf("%.17g %.17g\n", dzhip::freg::kRoots<N>.c[k], dzhip::freg::kRoots<N>.s[k]);
}

int main() {
    int n, sign;
    while (std::scanf("%d %d", &n, &sign) == 2) {
        std::vector<double> in(sign == 0 ? 0 : 2 * n);
        for (auto& v : in)
            if (std::scanf("%lf", &v) != 1) return 1;
        switch (n) {
#define DZ_CASE(N)            \
    case N:                   \
        if (sign == 0)        \
            roots<N>();       \
        else                  \
            both<N>(sign, in); \
        break;
            DZ_CASE(8) DZ_CASE(9) DZ_CASE(12) DZ_CASE(15) DZ_CASE(16) DZ_CASE(18) DZ_CASE(24) DZ_CASE(25)
            DZ_CASE(27) DZ_CASE(30) DZ_CASE(32) DZ_CASE(36) DZ_CASE(45) DZ_CASE(64) DZ_CASE(72) DZ_CASE(75)
            DZ_CASE(81) DZ_CASE(90) DZ_CASE(96) DZ_CASE(128)
#undef DZ_CASE
            default:
                return 2;
        }
    }
    return 0;
}
