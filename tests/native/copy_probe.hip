// Plain device copy of a known size (256 MiB read + 256 MiB written per launch, 3 launches):
// the calibration that turns rocprofv3's FETCH_SIZE / WRITE_SIZE counters into bytes for the
// H.psi kernels' traffic figures.  Built by tests/native/Makefile as tests/native/copy_probe.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy_kernel(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

int main() {
    const size_t bytes = 256ull << 20;
    const size_t n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 0, bytes) != hipSuccess) return 1;
    for (int it = 0; it < 3; ++it)
        hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 1;
    std::printf("copied %zu bytes x3\n", bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
