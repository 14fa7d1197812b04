// Standalone batched 3D FFT (the `hip` FFT backend) and the shared axis tables.
//
// Replaces Fft3dPlan.forward/inverse for the fftw backend
// (src/features/fft/fft.zig:224-262, src/lib/fft/fftw_fft.zig:141-166):
// in place, forward unnormalised, inverse scaled by 1/N.
//
// Two passes over a grid stored x + nx*(y + ny*z):
//   pass 1: one workgroup per z-plane (contiguous nx*ny block): x- and y-lines in LDS;
//   pass 2: one workgroup per y (nz rows of nx contiguous values): z-lines in LDS.
#include <cmath>
#include <cstring>
#include <type_traits>

#include "fft_lds.h"

namespace dzhip {

bool factor_fft_len(int n, std::vector<int>& radices) {
    radices.clear();
    if (n < 1 || n > kMaxFftLen) return false;
    int m = n;
    // Largest register radices first: fewer LDS round trips per line (90 = 9 x 10).  No radix 15:
    // its 15 complex values do not fit beside the FP64 tile kernels' other live state at their
    // 128-VGPR budget (the forward plane kernel then spills).
    for (int r : {9, 8, 10, 6, 4, 2, 3, 5}) {
        while (m % r == 0) { radices.push_back(r); m /= r; }
    }
    if (m != 1) return false;
    return (int)radices.size() <= kMaxStages;
}

// Position of input index i for the in-place DIT whose stage s uses radices[s]:
// the last stage splits the input by i mod R_S into blocks of n/R_S, recursively.
static int dr_pos(int i, int n, const int* rad, int nrad) {
    if (nrad == 0) return 0;
    const int R = rad[nrad - 1];
    const int m = n / R;
    return (i % R) * m + dr_pos(i / R, m, rad, nrad - 1);
}

FftAxis make_fft_axis(int n, const std::vector<int>& radices) {
    FftAxis ax{};
    ax.n = n;
    ax.nstages = (int)radices.size();
    for (int s = 0; s < kMaxStages; ++s) ax.radix[s] = s < ax.nstages ? radices[s] : 1;
    for (int s = 0, sub = 1; s < kMaxStages; sub *= ax.radix[s], ++s) ax.sub_magic[s] = div_magic(sub);
    return ax;
}

std::vector<int> digit_reverse_perm(int n, const std::vector<int>& radices) {
    std::vector<int> p(n);
    for (int i = 0; i < n; ++i) p[i] = dr_pos(i, n, radices.data(), (int)radices.size());
    return p;
}

void AxisTables::build(int nx, int ny, int nz) {
    const int lens[3] = {nx, ny, nz};
    size_t tw_total = 0, perm_total = 0;
    std::vector<std::vector<int>> rads(3);
    for (int d = 0; d < 3; ++d) {
        DZ_REQUIRE(factor_fft_len(lens[d], rads[d]), DZHIP_ERR_ARG,
                   "FFT length " + std::to_string(lens[d]) + " is not 2/3/5-smooth or exceeds 256");
        ax[d] = make_fft_axis(lens[d], rads[d]);
        ax[d].tw_off = (int)tw_total;
        ax[d].perm_off = (int)perm_total;
        tw_total += lens[d];
        perm_total += lens[d];
    }
    std::vector<double2> h64(tw_total);
    std::vector<float2> h32(tw_total);
    std::vector<int> hp(perm_total);
    for (int d = 0; d < 3; ++d) {
        const int n = lens[d];
        for (int t = 0; t < n; ++t) {
            // exp(-2 pi i t/n), evaluated in long double and rounded once.
            long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / n;
            double c = (double)cosl(ang), s = (double)sinl(ang);
            h64[ax[d].tw_off + t] = make_double2(c, s);
            h32[ax[d].tw_off + t] = make_float2((float)c, (float)s);
        }
        std::vector<int> p = digit_reverse_perm(n, rads[d]);
        std::memcpy(&hp[ax[d].perm_off], p.data(), n * sizeof(int));
    }
    DZ_HIP(hipMalloc(&tw64, tw_total * sizeof(double2)));
    DZ_HIP(hipMalloc(&tw32, tw_total * sizeof(float2)));
    DZ_HIP(hipMalloc(&perm, perm_total * sizeof(int)));
    DZ_HIP(hipMemcpy(tw64, h64.data(), tw_total * sizeof(double2), hipMemcpyHostToDevice));
    DZ_HIP(hipMemcpy(tw32, h32.data(), tw_total * sizeof(float2), hipMemcpyHostToDevice));
    DZ_HIP(hipMemcpy(perm, hp.data(), perm_total * sizeof(int), hipMemcpyHostToDevice));
    bytes = tw_total * (sizeof(double2) + sizeof(float2)) + perm_total * sizeof(int);
}

void AxisTables::release() {
    if (tw64) (void)hipFree(tw64);
    if (tw32) (void)hipFree(tw32);
    if (perm) (void)hipFree(perm);
    tw64 = nullptr;
    tw32 = nullptr;
    perm = nullptr;
}

size_t fft_xy_lds_bytes(int nx, int ny, size_t esz) { return ((size_t)(nx + 1) * ny + nx + ny) * esz; }
size_t fft_z_lds_bytes(int nx, int nz, size_t esz) { return ((size_t)(nx + 1) * nz + nz) * esz; }

void set_max_dynamic_lds(const void* func, size_t bytes) {
    if (bytes <= 65536) return;
    DZ_HIP(hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// ----------------------------------------------------------------- kernels
// Pass 1: z-plane = contiguous [y][x] block; LDS tile rows of pitch nx+1; transform along
// x, then y.
template <typename T, int S>
__global__ __launch_bounds__(TileThreads<T>::value) void k_fft_xy(
    typename C2T<T>::type* __restrict__ data, FftAxis axx, FftAxis axy,
    const typename C2T<T>::type* __restrict__ tw, const int* __restrict__ perm) {
    using C = typename C2T<T>::type;
    extern __shared__ unsigned char smem_raw[];
    const int nx = axx.n, ny = axy.n, pitch = nx + 1;
    C* buf = reinterpret_cast<C*>(smem_raw);
    C* twx = buf + pitch * ny;
    C* twy = twx + nx;
    const int tid = threadIdx.x, nthr = blockDim.x;
    load_twiddles(twx, tw + axx.tw_off, nx, tid, nthr);
    load_twiddles(twy, tw + axy.tw_off, ny, tid, nthr);
    C* plane = data + (size_t)blockIdx.x * nx * ny;
    const int* px = perm + axx.perm_off;
    const int* py = perm + axy.perm_off;
    const unsigned mx = div_magic(nx);
    for (int i = tid; i < nx * ny; i += nthr) {
        const int y = fdiv(i, nx, mx), x = i - y * nx;
        buf[px[x] + pitch * py[y]] = plane[i];
    }
    __syncthreads();
    lds_fft_lines<T, C, true, S>(buf, axx, twx, LineSet{ny, 1, pitch, nullptr}, tid, nthr);
    lds_fft_lines<T, C, true, S>(buf, axy, twy, LineSet{nx, pitch, 1, nullptr}, tid, nthr);
    for (int i = tid; i < nx * ny; i += nthr) {
        const int y = fdiv(i, nx, mx), x = i - y * nx;
        plane[i] = buf[x + pitch * y];
    }
}

// Pass 2: fixed y, LDS tile [z][x] of pitch nx+1; transform along z; optional scale.
template <typename T, int S>
__global__ __launch_bounds__(TileThreads<T>::value) void k_fft_z(
    typename C2T<T>::type* __restrict__ data, int nx, int ny, FftAxis axz,
    const typename C2T<T>::type* __restrict__ tw, const int* __restrict__ perm, T scale) {
    using C = typename C2T<T>::type;
    extern __shared__ unsigned char smem_raw[];
    const int nz = axz.n, pitch = nx + 1;
    C* buf = reinterpret_cast<C*>(smem_raw);
    C* twz = buf + pitch * nz;
    const int y = blockIdx.x;
    const int tid = threadIdx.x, nthr = blockDim.x;
    load_twiddles(twz, tw + axz.tw_off, nz, tid, nthr);
    const int* pz = perm + axz.perm_off;
    const unsigned mx = div_magic(nx);
    for (int i = tid; i < nx * nz; i += nthr) {
        const int z = fdiv(i, nx, mx), x = i - z * nx;
        buf[x + pitch * pz[z]] = data[x + (size_t)nx * (y + (size_t)ny * z)];
    }
    __syncthreads();
    lds_fft_lines<T, C, true, S>(buf, axz, twz, LineSet{nx, pitch, 1, nullptr}, tid, nthr);
    for (int i = tid; i < nx * nz; i += nthr) {
        const int z = fdiv(i, nx, mx), x = i - z * nx;
        const C v = buf[x + pitch * z];
        data[x + (size_t)nx * (y + (size_t)ny * z)] = C{v.x * scale, v.y * scale};
    }
}

template <typename T>
__global__ void k_to_float(const double2* __restrict__ in, float2* __restrict__ out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2((float)in[i].x, (float)in[i].y);
}
__global__ void k_to_double(const float2* __restrict__ in, double2* __restrict__ out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_double2((double)in[i].x, (double)in[i].y);
}

void convert_to_float(const double2* in, float2* out, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_to_float<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
    DZ_HIP(hipGetLastError());
}
void convert_to_double(const float2* in, double2* out, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_to_double, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
    DZ_HIP(hipGetLastError());
}

template <typename T>
void fft3d_device(typename C2T<T>::type* data, const AxisTables& tab, bool inverse,
                  hipStream_t s) {
    using C = typename C2T<T>::type;
    const int nx = tab.ax[0].n, ny = tab.ax[1].n, nz = tab.ax[2].n;
    const C* tw = TwSel<T>::get(tab.tw64, tab.tw32);
    const T scale = inverse ? T(1.0 / ((double)nx * ny * nz)) : T(1);
    const size_t lds1 = fft_xy_lds_bytes(nx, ny, sizeof(C));
    const size_t lds2 = fft_z_lds_bytes(nx, nz, sizeof(C));
    constexpr int nt = TileThreads<T>::value;
    auto run = [&](auto sign_tag) {
        constexpr int S = decltype(sign_tag)::value;
        set_max_dynamic_lds((const void*)k_fft_xy<T, S>, lds1);
        set_max_dynamic_lds((const void*)k_fft_z<T, S>, lds2);
        hipLaunchKernelGGL((k_fft_xy<T, S>), dim3(nz), dim3(nt), lds1, s, data, tab.ax[0], tab.ax[1], tw, tab.perm);
        DZ_HIP(hipGetLastError());
        hipLaunchKernelGGL((k_fft_z<T, S>), dim3(ny), dim3(nt), lds2, s, data, nx, ny, tab.ax[2], tw, tab.perm, scale);
        DZ_HIP(hipGetLastError());
    };
    if (inverse)
        run(std::integral_constant<int, 1>{});
    else
        run(std::integral_constant<int, -1>{});
}

template void fft3d_device<double>(double2*, const AxisTables&, bool, hipStream_t);
template void fft3d_device<float>(float2*, const AxisTables&, bool, hipStream_t);

}  // namespace dzhip
