// Host-side run of the in-LDS FFT core (dft-zig_amd/csrc/fft_lds.h): the same stage code the
// kernels execute, run sequentially on the CPU over a strided multi-line tile, checked against
// a textbook DFT for DIT (digit-reversed in) and DIF (digit-reversed out), both signs.
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "../../dft-zig_amd/csrc/fft_lds.h"

namespace dzhip {
bool factor_fft_len(int n, std::vector<int>& radices);
std::vector<int> digit_reverse_perm(int n, const std::vector<int>& radices);
}  // namespace dzhip

using namespace dzhip;

template <typename T>
static double run(int n, bool dit, double sign, int nlines, int es, int ls, bool use_list) {
    using C = typename C2T<T>::type;
    std::vector<int> rad;
    if (!factor_fft_len(n, rad)) return 1e9;
    const FftAxis ax = make_fft_axis(n, rad);
    std::vector<int> perm = digit_reverse_perm(n, rad);
    std::vector<C> tw(n);
    for (int t = 0; t < n; ++t) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * t / n;
        tw[t] = C{(T)cosl(a), (T)sinl(a)};
    }
    const int size = (nlines - 1) * ls + (n - 1) * es + 1 + 7;
    std::vector<C> buf(size, C{T(99), T(99)});
    std::vector<short> lines(nlines);
    for (int l = 0; l < nlines; ++l) lines[l] = (short)((nlines - 1 - l) * ls);  // reversed order
    std::vector<std::vector<std::complex<double>>> in(nlines, std::vector<std::complex<double>>(n));
    unsigned long long st = 12345 + n;
    auto rnd = [&]() { st = st * 6364136223846793005ULL + 1; return ((st >> 11) * (1.0 / 9007199254740992.0)) - 0.5; };
    for (int l = 0; l < nlines; ++l) {
        const int b = use_list ? lines[l] : l * ls;
        for (int i = 0; i < n; ++i) {
            in[l][i] = {rnd(), rnd()};
            const int pos = dit ? perm[i] : i;
            buf[b + pos * es] = C{(T)in[l][i].real(), (T)in[l][i].imag()};
        }
    }
    LineSet lsd{nlines, es, ls, use_list ? lines.data() : nullptr};
    if (dit && sign > 0) lds_fft_lines<T, C, true, 1>(buf.data(), ax, tw.data(), lsd, 0, 1);
    else if (dit) lds_fft_lines<T, C, true, -1>(buf.data(), ax, tw.data(), lsd, 0, 1);
    else if (sign > 0) lds_fft_lines<T, C, false, 1>(buf.data(), ax, tw.data(), lsd, 0, 1);
    else lds_fft_lines<T, C, false, -1>(buf.data(), ax, tw.data(), lsd, 0, 1);
    double err = 0, mx = 0;
    for (int l = 0; l < nlines; ++l) {
        const int b = use_list ? lines[l] : l * ls;
        for (int k = 0; k < n; ++k) {
            std::complex<double> ref = 0;
            for (int j = 0; j < n; ++j) ref += in[l][j] * std::polar(1.0, sign * 2 * M_PI * (double)j * k / n);
            const int pos = dit ? k : perm[k];
            const C v = buf[b + pos * es];
            err = std::max(err, std::abs(std::complex<double>(v.x, v.y) - ref));
            mx = std::max(mx, std::abs(ref));
        }
    }
    return err / mx;
}

int main() {
    const int sizes[] = {1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 16, 18, 24, 27, 30, 32, 36, 40, 45, 48, 60, 64,
                         72, 75, 80, 81, 90, 96, 100, 108, 120, 125, 128, 144, 150, 160, 180, 192, 200, 216, 243, 256};
    int fails = 0;
    for (int n : sizes)
        for (int dit = 0; dit < 2; ++dit)
            for (double sign : {-1.0, 1.0}) {
                const int nl = 5, es = 7, ls = 1;  // lines interleaved, elements strided (es > nl: no overlap)
                double e64 = run<double>(n, dit, sign, nl, es, ls, false);
                double e64b = run<double>(n, dit, sign, 3, 1, n + 1, true);
                double e32 = run<float>(n, dit, sign, nl, es, ls, false);
                const bool ok = e64 < 1e-13 && e64b < 1e-13 && e32 < 2e-6;
                if (!ok) {
                    ++fails;
                    printf("FAIL n=%d dit=%d sign=%g err64=%.3e err64list=%.3e err32=%.3e\n", n, dit, sign, e64, e64b, e32);
                }
            }
    printf("%s (%d failures)\n", fails ? "FFT core FAILED" : "FFT core ok", fails);
    return fails ? 1 : 0;
}
