// Learned diagonal preconditioner for the Davidson residuals: a 3 -> 32 -> 32 -> 1 SiLU network
// (FP32) that rescales a base diagonal preconditioner per coefficient,
//     s(G, b) = B(G, b) * exp(z),   z = MLP(|k+G|^2/2 / e_s, lambda_b/2 / e_s, log10|r_b| / 4),
// with e_s the model's feature energy scale (Ha) and B either the reference's kinetic shift
// 1/(|k+G|^2 - lambda_b) (src/features/linalg/lobpcg/common.zig:214-226, 1e-8 clamp) or the
// Teter-Payne-Allan K(|k+G|^2 / T_b), T_b = <x_b|T|x_b>.  The reference has no learned
// preconditioner.
//
// Kernels:
//   k_precondition_mlp  inference, one thread per coefficient of one selected band; the weights
//                       and the band's first-layer constants sit in LDS and are read as broadcasts
//   k_sample_band       per (snapshot, band) reductions for the training targets
//   k_sample_write      training samples (features, target, weight) from residual/correction pairs
//   k_mlp_grad          forward + backward of one minibatch; per-sample activations are staged in
//                       LDS (sample-major, odd pitch) and each thread then sums the gradient of
//                       ~10 parameters over the block's samples: one partial gradient per block
//   k_mlp_reduce        fixed-order sum of the per-block partials (deterministic), FP64
//   k_adam              Adam update of the 1217 parameters
#include <cmath>

#include "ham.h"

namespace dzhip {

namespace {

constexpr int H = kMlpHidden;
// Parameter offsets (layout W1[32][3], b1[32], W2[32][32], b2[32], w3[32], b3).
constexpr int OW1 = 0, OB1 = OW1 + 3 * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;
static_assert(OB3 + 1 == kMlpParams, "parameter count");

__device__ __forceinline__ float silu(float u) { return u / (1.0f + __expf(-u)); }
__device__ __forceinline__ float dsilu(float u) {
    const float sg = 1.0f / (1.0f + __expf(-u));
    return sg * (1.0f + u * (1.0f - sg));
}
// Base factor B(G, b) for diag = |k+G|^2: kinetic shift (base DZHIP_PRECOND_KINETIC) or TPA.
__device__ __forceinline__ double base_factor(int base, double diag, double lam, double ekin) {
    return base == DZHIP_PRECOND_KINETIC ? precond_kinetic_shift(diag, lam) : precond_tpa(diag, ekin);
}

constexpr int kInferThreads = 256;

// Q[:, j] = s(G, sel_j) R[:, sel_j]
__global__ __launch_bounds__(kInferThreads) void k_precondition_mlp(
    const double2* __restrict__ r, const int* __restrict__ sel, const double* __restrict__ lam,
    const double* __restrict__ ekin, const double* __restrict__ rnorm, const double* __restrict__ diag,
    const float* __restrict__ prm, float inv_es, int base, double2* __restrict__ q, size_t ld, int n) {
    __shared__ float w2[H * H];
    __shared__ float b2[H], w3[H], w1x[H], c1[H];
    const int j = blockIdx.y;
    const int b = sel[j];
    const float x2 = (float)(0.5 * lam[b]) * inv_es;
    const float x3 = (float)(log10(fmax(rnorm[b], 1e-30)) * 0.25);
    for (int t = threadIdx.x; t < H * H; t += blockDim.x) w2[t] = prm[OW2 + t];
    if (threadIdx.x < H) {
        const int k = threadIdx.x;
        b2[k] = prm[OB2 + k];
        w3[k] = prm[OW3 + k];
        w1x[k] = prm[OW1 + 3 * k];
        c1[k] = prm[OW1 + 3 * k + 1] * x2 + prm[OW1 + 3 * k + 2] * x3 + prm[OB1 + k];
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = diag[i];
    const float x1 = (float)(0.5 * d) * inv_es;
    float h1[H];
#pragma unroll
    for (int k = 0; k < H; ++k) h1[k] = silu(fmaf(w1x[k], x1, c1[k]));
    float z = prm[OB3];
#pragma unroll 4
    for (int o = 0; o < H; ++o) {
        float u = b2[o];
#pragma unroll
        for (int k = 0; k < H; ++k) u = fmaf(w2[o * H + k], h1[k], u);
        z = fmaf(w3[o], silu(u), z);
    }
    const double f = base_factor(base, d, lam[b], ekin[b]) * (double)__expf(fminf(z, 30.0f));
    const double2 v = r[(size_t)b * ld + i];
    q[(size_t)j * ld + i] = make_double2(v.x * f, v.y * f);
}

// Per (snapshot i, band b), with e the correction and r the residual of that band:
// out[0] = <K r, e> (real part), out[1] = |K r|^2, out[2] = |e|^2; K = B(G, b), the base factor.
__global__ void k_sample_band(const double2* __restrict__ e, const double2* __restrict__ r,
                              const double* __restrict__ lam, const double* __restrict__ ekin,
                              const double* __restrict__ diag, size_t ld, int n, int base, double* __restrict__ out) {
    __shared__ double red[3][256];
    const int col = blockIdx.x;
    const double2* ec = e + (size_t)col * ld;
    const double2* rc = r + (size_t)col * ld;
    const double ek = ekin[col], lb = lam[col];
    double a = 0.0, kk = 0.0, ee = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double k = base_factor(base, diag[i], lb, ek);
        const double2 rv = rc[i], ev = ec[i];
        a += k * (rv.x * ev.x + rv.y * ev.y);
        kk += k * k * (rv.x * rv.x + rv.y * rv.y);
        ee += ev.x * ev.x + ev.y * ev.y;
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = kk;
    red[2][threadIdx.x] = ee;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[3 * col + threadIdx.x] = red[threadIdx.x][0];
}

// Samples for coefficients i = 0, stride, 2 stride, ... of each (snapshot, band) column:
// features x1..x3, the base factor B(G, b), target t = Re(e conj r) / (alpha |r|^2) and weight
// w = alpha^2 |r|^2 / |e|^2 with alpha = <K r, e> / |K r|^2 (the best scale of the base
// preconditioner for the band); columns with alpha <= 0 or vanishing norms get weight 0.
// Output SoA [6][cap].
__global__ void k_sample_write(const double2* __restrict__ e, const double2* __restrict__ r,
                               const double* __restrict__ lam, const double* __restrict__ ekin,
                               const double* __restrict__ rnorm, const double* __restrict__ bandsum,
                               const double* __restrict__ diag, size_t ld, int n, int stride, int per_col,
                               double inv_es, int pbase, float* __restrict__ out, size_t cap, size_t base) {
    const int col = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per_col) return;
    const int i = t * stride;
    const size_t o = base + (size_t)col * per_col + t;
    if (i >= n || o >= cap) return;
    const double a_num = bandsum[3 * col], kk = bandsum[3 * col + 1], ee = bandsum[3 * col + 2];
    const double alpha = kk > 0.0 ? a_num / kk : 0.0;
    const double2 rv = r[(size_t)col * ld + i], ev = e[(size_t)col * ld + i];
    const double r2 = rv.x * rv.x + rv.y * rv.y;
    const bool ok = alpha > 0.0 && ee > 1e-24 && r2 > 1e-300 && rnorm[col] > 0.0;
    const double tgt = ok ? (ev.x * rv.x + ev.y * rv.y) / (alpha * r2) : 0.0;
    const double wgt = ok ? alpha * alpha * r2 / ee : 0.0;
    out[0 * cap + o] = (float)(0.5 * diag[i] * inv_es);
    out[1 * cap + o] = (float)(0.5 * lam[col] * inv_es);
    out[2 * cap + o] = (float)(log10(fmax(rnorm[col], 1e-30)) * 0.25);
    out[3 * cap + o] = (float)base_factor(pbase, diag[i], lam[col], ekin[col]);
    out[4 * cap + o] = (float)tgt;
    out[5 * cap + o] = (float)wgt;
}

constexpr int kGradThreads = 128;  // samples per block = threads
constexpr int kP = H + 1;          // LDS pitch (odd)

// One minibatch: samples idx(j) = (mul * (start + j) + add) mod n for j < count.  Loss term
// w (s - t)^2 / count; partial gradients per block into part[block][kMlpParams], partial loss
// into part[block][kMlpParams].
__global__ __launch_bounds__(kGradThreads) void k_mlp_grad(
    const float* __restrict__ smp, size_t cap, size_t n, unsigned long long mul, unsigned long long add,
    size_t start, size_t count, const float* __restrict__ prm, float inv_count, float* __restrict__ part) {
    extern __shared__ float lds[];
    float* sh1 = lds;                        // [S][kP] h1
    float* sd1 = sh1 + kGradThreads * kP;    // [S][kP] dL/du1
    float* sh2 = sd1 + kGradThreads * kP;    // [S][kP] h2
    float* sd2 = sh2 + kGradThreads * kP;    // [S][kP] dL/du2
    float* sx = sd2 + kGradThreads * kP;     // [S][4] x1 x2 x3 dz
    float* sw2 = sx + kGradThreads * 4;      // W2
    float* sloss = sw2 + H * H;              // [S]
    const int tid = threadIdx.x;
    for (int t = tid; t < H * H; t += blockDim.x) sw2[t] = prm[OW2 + t];
    __syncthreads();
    const size_t j = (size_t)blockIdx.x * kGradThreads + tid;
    float x1 = 0.f, x2 = 0.f, x3 = 0.f, dz = 0.f, loss = 0.f;
    // Each thread works in its own LDS rows (pitch 33: conflict-free across a wave); sd1/sd2
    // first hold silu'(u1)/silu'(u2) and are turned into dL/du1, dL/du2 in place.
    float* h1r = sh1 + tid * kP;
    float* d1r = sd1 + tid * kP;
    float* h2r = sh2 + tid * kP;
    float* d2r = sd2 + tid * kP;
    if (j < count) {
        const size_t s = (size_t)((mul * (unsigned long long)(start + j) + add) % (unsigned long long)n);
        x1 = smp[s];
        x2 = smp[cap + s];
        x3 = smp[2 * cap + s];
        const float base = smp[3 * cap + s], tg = smp[4 * cap + s], wg = smp[5 * cap + s];
        float h1[H];
#pragma unroll
        for (int k = 0; k < H; ++k) {
            const float u = fmaf(prm[OW1 + 3 * k], x1, fmaf(prm[OW1 + 3 * k + 1], x2, fmaf(prm[OW1 + 3 * k + 2], x3, prm[OB1 + k])));
            h1[k] = silu(u);
            h1r[k] = h1[k];
            d1r[k] = dsilu(u);
        }
        float z = prm[OB3];
#pragma unroll 2
        for (int o = 0; o < H; ++o) {
            float u = prm[OB2 + o];
#pragma unroll
            for (int k = 0; k < H; ++k) u = fmaf(sw2[o * H + k], h1[k], u);
            const float h2 = silu(u);
            h2r[o] = h2;
            d2r[o] = dsilu(u);
            z = fmaf(prm[OW3 + o], h2, z);
        }
        const float sv = base * __expf(fminf(z, 30.0f));
        const float diff = sv - tg;
        loss = wg * diff * diff * inv_count;
        dz = 2.0f * wg * diff * sv * inv_count;
        for (int o = 0; o < H; ++o) d2r[o] *= dz * prm[OW3 + o];
        for (int k = 0; k < H; ++k) {
            float acc = 0.f;
#pragma unroll 8
            for (int o = 0; o < H; ++o) acc = fmaf(d2r[o], sw2[o * H + k], acc);
            d1r[k] *= acc;
        }
    } else {
        for (int k = 0; k < H; ++k) {
            h1r[k] = 0.f;
            d1r[k] = 0.f;
            h2r[k] = 0.f;
            d2r[k] = 0.f;
        }
    }
    sx[tid * 4 + 0] = x1;
    sx[tid * 4 + 1] = x2;
    sx[tid * 4 + 2] = x3;
    sx[tid * 4 + 3] = dz;
    sloss[tid] = loss;
    __syncthreads();
    float* outp = part + (size_t)blockIdx.x * (kMlpParams + 1);
    for (int pi = tid; pi <= kMlpParams; pi += blockDim.x) {
        float acc = 0.f;
        if (pi < OB1) {  // W1[k][c] = sum d1[k] x[c]
            const int k = pi / 3, c = pi % 3;
            for (int s = 0; s < kGradThreads; ++s) acc = fmaf(sd1[s * kP + k], sx[s * 4 + c], acc);
        } else if (pi < OW2) {
            const int k = pi - OB1;
            for (int s = 0; s < kGradThreads; ++s) acc += sd1[s * kP + k];
        } else if (pi < OB2) {  // W2[o][k] = sum d2[o] h1[k]
            const int o = (pi - OW2) / H, k = (pi - OW2) % H;
            for (int s = 0; s < kGradThreads; ++s) acc = fmaf(sd2[s * kP + o], sh1[s * kP + k], acc);
        } else if (pi < OW3) {
            const int o = pi - OB2;
            for (int s = 0; s < kGradThreads; ++s) acc += sd2[s * kP + o];
        } else if (pi < OB3) {  // w3[o] = sum dz h2[o]
            const int o = pi - OW3;
            for (int s = 0; s < kGradThreads; ++s) acc = fmaf(sx[s * 4 + 3], sh2[s * kP + o], acc);
        } else if (pi == OB3) {
            for (int s = 0; s < kGradThreads; ++s) acc += sx[s * 4 + 3];
        } else {  // loss
            for (int s = 0; s < kGradThreads; ++s) acc += sloss[s];
        }
        outp[pi] = acc;
    }
}

// g[p] = sum over blocks of part[block][p] (p <= kMlpParams; the last entry is the loss).
__global__ void k_mlp_reduce(const float* __restrict__ part, int nblocks, double* __restrict__ g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    double acc = 0.0;
    for (int b = 0; b < nblocks; ++b) acc += (double)part[(size_t)b * (kMlpParams + 1) + p];
    g[p] = acc;
}

__global__ void k_adam(float* __restrict__ prm, const double* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float gi = (float)g[p];
    const float mi = 0.9f * m[p] + 0.1f * gi;
    const float vi = 0.999f * v[p] + 0.001f * gi * gi;
    m[p] = mi;
    v[p] = vi;
    prm[p] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
}

size_t grad_lds_bytes() { return sizeof(float) * (4 * kGradThreads * kP + 4 * kGradThreads + H * H + kGradThreads); }

}  // namespace

void precondition_mlp(const double2* r, const int* sel, int nsel, const double* lam, const double* ekin,
                      const double* rnorm, const double* diag, const float* prm, double energy_scale, int base,
                      double2* q, size_t ld, int n, hipStream_t s) {
    if (nsel <= 0) return;
    hipLaunchKernelGGL(k_precondition_mlp, dim3((unsigned)((n + kInferThreads - 1) / kInferThreads), (unsigned)nsel),
                       dim3(kInferThreads), 0, s, r, sel, lam, ekin, rnorm, diag, prm, (float)(1.0 / energy_scale), base,
                       q, ld, n);
    DZ_HIP(hipGetLastError());
}

size_t mlp_write_samples(const double2* e, const double2* r, const double* lam, const double* ekin,
                         const double* rnorm, const double* diag, size_t ld, int n, int cols, int stride,
                         double energy_scale, int pbase, double* bandsum, float* out, size_t cap, size_t base,
                         hipStream_t s) {
    if (cols <= 0) return 0;
    hipLaunchKernelGGL(k_sample_band, dim3((unsigned)cols), dim3(256), 0, s, e, r, lam, ekin, diag, ld, n, pbase,
                       bandsum);
    DZ_HIP(hipGetLastError());
    const int per_col = (n + stride - 1) / stride;
    hipLaunchKernelGGL(k_sample_write, dim3((unsigned)((per_col + 255) / 256), (unsigned)cols), dim3(256), 0, s, e, r,
                       lam, ekin, rnorm, bandsum, diag, ld, n, stride, per_col, 1.0 / energy_scale, pbase, out, cap,
                       base);
    DZ_HIP(hipGetLastError());
    return std::min<size_t>((size_t)cols * per_col, cap > base ? cap - base : 0);
}

MlpTrainer::MlpTrainer(size_t cap_) : cap(cap_) {
    samples.ensure(6 * cap * sizeof(float));
    prm.ensure(kMlpParams * sizeof(float));
    m.ensure(kMlpParams * sizeof(float));
    v.ensure(kMlpParams * sizeof(float));
    g.ensure((kMlpParams + 1) * sizeof(double));
    set_max_dynamic_lds(reinterpret_cast<const void*>(&k_mlp_grad), grad_lds_bytes());
}

MlpTrainer::~MlpTrainer() {
    DevBuf* bufs[] = {&samples, &prm, &m, &v, &g, &part};
    for (DevBuf* b : bufs) b->release();
}

double MlpTrainer::loss_grad(size_t n, unsigned long long mul, unsigned long long add, size_t start, size_t count,
                             hipStream_t s) {
    const int nblocks = (int)((count + kGradThreads - 1) / kGradThreads);
    part.ensure((size_t)nblocks * (kMlpParams + 1) * sizeof(float));
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)nblocks), dim3(kGradThreads), grad_lds_bytes(), s,
                       samples.as<float>(), cap, n, mul, add, start, count, prm.as<float>(), 1.0f / (float)count,
                       part.as<float>());
    DZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, s, part.as<float>(), nblocks,
                       g.as<double>());
    DZ_HIP(hipGetLastError());
    return 0.0;
}

void MlpTrainer::adam(int step, double lr, hipStream_t s) {
    const float bc1 = 1.0f - powf(0.9f, (float)step), bc2 = 1.0f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, prm.as<float>(), g.as<double>(),
                       m.as<float>(), v.as<float>(), (float)lr, bc1, bc2);
    DZ_HIP(hipGetLastError());
}

}  // namespace dzhip
