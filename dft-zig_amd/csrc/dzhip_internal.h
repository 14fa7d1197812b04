// Internal definitions shared by the HIP translation units of libdzhip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dzhip.h"

namespace dzhip {

// ----------------------------------------------------------------- errors
void set_error(const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

#define DZ_HIP(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            throw ::dzhip::Error{DZHIP_ERR_HIP, std::string(#call) + ": " +          \
                                                    hipGetErrorString(e_)};          \
    } while (0)

#define DZ_REQUIRE(cond, code, msg)                                                  \
    do {                                                                             \
        if (!(cond)) throw ::dzhip::Error{(code), (msg)};                            \
    } while (0)

// ----------------------------------------------------------------- complex
template <typename T> struct C2T;
template <> struct C2T<float> { using type = float2; };
template <> struct C2T<double> { using type = double2; };

template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename C>
__device__ __forceinline__ C cmulc(C a, C b) {
    return C{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename C, typename T>
__device__ __forceinline__ C cscale(C a, T s) { return C{a.x * s, a.y * s}; }
// i * s * a  (s = +-1 folded by caller)
template <typename C>
__device__ __forceinline__ C cmul_i(C a) { return C{-a.y, a.x}; }

// ----------------------------------------------------------------- diagonal preconditioners
// Kinetic shift 1 / (|k+G|^2 - lambda) with |denominator| clamped at 1e-8, sign kept
// (src/features/linalg/lobpcg/common.zig:214-226).
__device__ __forceinline__ double precond_kinetic_shift(double diag, double lam) {
    double den = diag - lam;
    if (fabs(den) < 1e-8) den = den >= 0.0 ? 1e-8 : -1e-8;
    return 1.0 / den;
}
// Teter-Payne-Allan K(x) = (27 + 18x + 12x^2 + 8x^3) / (27 + 18x + 12x^2 + 8x^3 + 16x^4),
// x = |k+G|^2 / <T>_b (PRB 40, 12255); <T>_b is floored at 1e-30 so that a band of zero kinetic
// energy gives K(0) = 1 at G = 0 and K -> 0 elsewhere instead of 0 / 0.
__device__ __forceinline__ double precond_tpa(double diag, double ekin) {
    const double x = diag / fmax(ekin, 1e-30);
    const double p = 27.0 + x * (18.0 + x * (12.0 + x * 8.0));
    return p / (p + 16.0 * x * x * x * x);
}

// ----------------------------------------------------------------- FFT axis
// One transform length n = prod radix[s] (radices 2/3/4/5/8/9), with device tables:
//   twiddle[t] = exp(-2 pi i t / n), t in [0, n)   (inverse uses the conjugate)
//   perm[i]    = LDS position of input index i for the in-place DIT, which is also
//                the position where the in-place DIF leaves output frequency i.
constexpr int kMaxStages = 12;
constexpr int kMaxFftLen = 256;

struct FftAxis {
    int n;
    int nstages;
    int radix[kMaxStages];
    int tw_off;   // element offset of this axis' twiddles in the table
    int perm_off; // element offset of this axis' permutation in the table
    // Division constant (div_magic in fft_lds.h) of each stage's butterfly distance
    // prod_{i < s} radix[i], so the kernels divide by it without computing it.
    unsigned sub_magic[kMaxStages];
};

// Host-side factorisation and tables (fft_core.cpp-like helpers in fft.hip).
bool factor_fft_len(int n, std::vector<int>& radices);
std::vector<int> digit_reverse_perm(int n, const std::vector<int>& radices);
// Axis of length n with the given stage radices: fills n, nstages, radix[] and sub_magic[]
// (table offsets left 0).
FftAxis make_fft_axis(int n, const std::vector<int>& radices);

// Device-resident set of axis tables for a grid (up to 3 distinct lengths).
struct AxisTables {
    FftAxis ax[3];
    double2* tw64 = nullptr;
    float2* tw32 = nullptr;
    int* perm = nullptr;
    size_t bytes = 0;
    void build(int nx, int ny, int nz);
    void release();
};

// ----------------------------------------------------------------- context
struct HeevWork;

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    HeevWork* heev = nullptr;  // dense eigensolver workspace (dzhip_heev*), created on first use
};

__host__ __device__ inline size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Device buffer that only grows.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    void ensure(size_t need) {
        if (need <= bytes) return;
        if (ptr) DZ_HIP(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&ptr, need);
        if (e != hipSuccess) {
            ptr = nullptr;
            throw Error{DZHIP_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(need) +
                                              "): " + hipGetErrorString(e)};
        }
        bytes = need;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T> T* as() const { return reinterpret_cast<T*>(ptr); }
};

void set_max_dynamic_lds(const void* func, size_t bytes);
size_t fft_xy_lds_bytes(int nx, int ny, size_t esz);
size_t fft_z_lds_bytes(int nx, int nz, size_t esz);
constexpr size_t kMaxLdsBytes = 160 * 1024;

}  // namespace dzhip
