// Device-resident SCF grid state: density, Hartree + LDA-PZ or PBE potential, energies and
// potential mixing (linear, then Pulay/DIIS), so an SCF iteration keeps rho(r) and V(r) in HBM.
//
// Follows (paths relative to the dft-zig source tree):
//   Hartree V_H(G) = 8 pi rho(G) / G^2, G = 0 dropped              src/features/scf/potential.zig:34-44
//   LDA-PZ energy density and potential (Ry)                       src/features/xc/xc.zig:195-210, 344-384
//   PBE (Ry): exchange, correlation H(t^2, A) on PW92               src/features/xc/xc.zig:213-342
//   GGA fields: grad n = IFFT(i G n(G)), b = df/d|grad n|^2 grad n,
//   v_xc = df/dn - 2 div b, div by i G . b(G), real parts kept     src/features/scf/xc_fields.zig:87-140, 306-375
//   local_r = V_ion + V_Hxc on the FFT-order grid                   src/features/scf/potential.zig:320-334
//   linear potential mixing V <- V + beta (V_out - V)              src/features/scf/mixing.zig:10-18
//   Pulay on (V_in, R = V_out - V_in), history h, from iteration s  src/features/scf/mixing.zig:183-234,
//                                                                   src/features/scf/scf.zig:147-188
//   RMS density change                                              src/features/scf/util.zig:17-24
// The Pulay overlaps are taken in real space divided by N (the reference's G-space norm by
// Parseval) and the small system is scaled by its largest diagonal before elimination; a
// singular system keeps only the newest pair and takes a linear step (the reference would
// average the whole history).
#include <cmath>
#include <vector>

#include "potential.h"

namespace dzhip {

namespace {

constexpr int kRedThreads = 256;

// |G|^2 on the FFT-order grid (index x + nx (y + ny z), wrapped frequencies); b rows = b1, b2, b3.
__global__ void k_g2(double* __restrict__ g2, int nx, int ny, int nz, double3 b1, double3 b2, double3 b3) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (i >= n) return;
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((size_t)nx * ny));
    const int fx = x <= (nx - 1) / 2 ? x : x - nx;
    const int fy = y <= (ny - 1) / 2 ? y : y - ny;
    const int fz = z <= (nz - 1) / 2 ? z : z - nz;
    const double gx = fx * b1.x + fy * b2.x + fz * b3.x;
    const double gy = fx * b1.y + fy * b2.y + fz * b3.y;
    const double gz = fx * b1.z + fy * b2.z + fz * b3.z;
    g2[i] = gx * gx + gy * gy + gz * gz;
}

__global__ void k_real_to_complex(const double* __restrict__ in, double2* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_double2(in[i], 0.0);
}

// w(G) *= 8 pi / G^2 (G = 0 -> 0): unnormalised forward transform in, so the normalised inverse
// gives V_H(r) = sum_G V_H(G) e^{iG.r}.
__global__ void k_hartree_kernel(double2* __restrict__ w, const double* __restrict__ g2, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double g = g2[i];
    const double f = g > 1e-12 ? 8.0 * M_PI / g : 0.0;
    w[i] = make_double2(w[i].x * f, w[i].y * f);
}

__device__ inline void lda_pz(double n, double* f, double* v) {
    if (!(n > 1e-12)) {
        *f = 0.0;
        *v = 0.0;
        return;
    }
    const double cbrt_n = cbrt(n);
    const double cx = -0.75 * cbrt(3.0 / M_PI);
    const double rs = cbrt(3.0 / (4.0 * M_PI * n));
    double ec, vc;
    if (rs < 1.0) {
        const double ln = log(rs);
        ec = 0.0311 * ln - 0.048 + 0.0020 * rs * ln - 0.0116 * rs;
        vc = ec - (rs / 3.0) * (0.0311 / rs + 0.0020 * ln + 0.0020 - 0.0116);
    } else {
        const double sq = sqrt(rs);
        const double den = 1.0 + 1.0529 * sq + 0.3334 * rs;
        ec = -0.1423 / den;
        vc = ec - (rs / 3.0) * (0.1423 * (0.5 * 1.0529 / sq + 0.3334) / (den * den));
    }
    *f = 2.0 * n * (cx * cbrt_n + ec);
    *v = 2.0 * (4.0 / 3.0 * cx * cbrt_n + vc);
}

// PW92 unpolarised correlation (Ha): eps, d eps / d rs, v_c = eps - rs/3 d eps/d rs.
__device__ inline void pw92(double rs, double* eps, double* deps, double* vc) {
    constexpr double a = 0.031091, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
    const double sr = sqrt(rs);
    const double f = b1 * sr + b2 * rs + b3 * rs * sr + b4 * rs * rs;
    const double g = 1.0 + 1.0 / (2.0 * a * f);
    const double lg = log(g);
    const double df = 0.5 * b1 / sr + b2 + 1.5 * b3 * sr + 2.0 * b4 * rs;
    const double dlg = -(1.0 / (2.0 * a * f * f * g)) * df;
    *eps = -2.0 * a * (1.0 + a1 * rs) * lg;
    *deps = -2.0 * a * (a1 * lg + (1.0 + a1 * rs) * dlg);
    *vc = *eps - (rs / 3.0) * *deps;
}

// PBE at density n and g2 = |grad n|^2: f, df/dn, df/dg2 (Ry); zero below n = 1e-12.
__device__ inline void pbe(double n, double g2, double* f, double* fn, double* fg) {
    if (!(n > 1e-12)) {
        *f = *fn = *fg = 0.0;
        return;
    }
    // Exchange: F_x = 1 + kappa - kappa / (1 + mu s^2 / kappa).
    constexpr double mu = 0.2195149727645171, kappa = 0.804;
    const double cx = -0.75 * cbrt(3.0 / M_PI);
    const double n13 = cbrt(n);
    const double ex0 = n * cx * n13;
    const double kf = cbrt(3.0 * M_PI * M_PI * n);
    const double den_x = 4.0 * kf * kf * n * n;
    const double s2 = g2 > 0.0 ? g2 / den_x : 0.0;
    const double t = 1.0 + mu * s2 / kappa;
    const double fx = 1.0 + kappa - kappa / t;
    const double dfx = mu / (t * t);
    const double fx_n = 4.0 / 3.0 * cx * n13 * fx + ex0 * dfx * (-8.0 / 3.0 * s2 / n);
    const double fx_g = ex0 * dfx / den_x;
    // Correlation: H = gamma ln(1 + (beta/gamma) t^2 (1 + A t^2) / (1 + A t^2 + A^2 t^4)).
    constexpr double beta = 0.06672455060314922, gamma = 0.031090690869654895, bg = beta / gamma;
    const double rs = cbrt(3.0 / (4.0 * M_PI * n));
    double ec, deps, vc;
    pw92(rs, &ec, &deps, &vc);
    const double deps_dn = deps * (-rs / (3.0 * n));
    const double den_c = 4.0 * (4.0 * kf / M_PI) * n * n;
    const double t2 = g2 > 0.0 ? g2 / den_c : 0.0;
    const double x = bg * t2;
    const double u = exp(-ec / gamma);
    const double um = u - 1.0;
    const bool big = fabs(um) > 1e-12;
    const double A = big ? bg / um : 0.0;
    const double y = A * t2;
    const double d2 = 1.0 + y + y * y;
    const double q = (1.0 + y) / d2;
    const double arg = 1.0 + x * q;
    const bool pos = arg > 0.0;
    const double H = pos ? gamma * log(arg) : 0.0;
    const double dq = -A * y * (2.0 + y) / (d2 * d2);
    const double dH_dt2 = pos ? gamma * (bg * q + x * dq) / arg : 0.0;
    const double dH_dA = pos ? -gamma * x * t2 * y * (2.0 + y) / (arg * d2 * d2) : 0.0;
    const double dA_dn = big ? bg * u / (gamma * um * um) * deps_dn : 0.0;
    const double dH_dn = dH_dt2 * (-7.0 / 3.0 * t2 / n) + dH_dA * dA_dn;
    *f = 2.0 * (ex0 * fx + n * (ec + H));
    *fn = 2.0 * (fx_n + vc + H + n * dH_dn);
    *fg = 2.0 * (fx_g + n * dH_dt2 / den_c);
}

// Cartesian G of FFT-order point i (wrapped frequencies); b = reciprocal rows.
__device__ inline double3 g_at(size_t i, int nx, int ny, int nz, double3 b1, double3 b2, double3 b3) {
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((size_t)nx * ny));
    const int fx = x <= (nx - 1) / 2 ? x : x - nx;
    const int fy = y <= (ny - 1) / 2 ? y : y - ny;
    const int fz = z <= (nz - 1) / 2 ? z : z - nz;
    return make_double3(fx * b1.x + fy * b2.x + fz * b3.x, fx * b1.y + fy * b2.y + fz * b3.y,
                        fx * b1.z + fy * b2.z + fz * b3.z);
}

__device__ inline double comp(double3 v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }

// out(G) = i G_c w(G)
__global__ void k_grad_comp(const double2* __restrict__ w, double2* __restrict__ out, int c, int nx, int ny, int nz,
                            double3 b1, double3 b2, double3 b3) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (i >= n) return;
    const double g = comp(g_at(i, nx, ny, nz, b1, b2, b3), c);
    out[i] = make_double2(-g * w[i].y, g * w[i].x);
}

// acc(G) (+)= i G_c w(G)
__global__ void k_div_acc(const double2* __restrict__ w, double2* __restrict__ acc, int c, int first, int nx, int ny,
                          int nz, double3 b1, double3 b2, double3 b3) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (i >= n) return;
    const double g = comp(g_at(i, nx, ny, nz, b1, b2, b3), c);
    const double2 t = make_double2(-g * w[i].y, g * w[i].x);
    acc[i] = first ? t : make_double2(acc[i].x + t.x, acc[i].y + t.y);
}

// f, df/dn and b_c = df/dg2 * (grad n)_c (overwriting the gradient components).
__global__ void k_pbe_point(const double* __restrict__ rho, double* __restrict__ gx, double* __restrict__ gy,
                            double* __restrict__ gz, double* __restrict__ fxc, double* __restrict__ vn, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ax = gx[i], ay = gy[i], az = gz[i];
    double f, fn, fg;
    pbe(rho[i], ax * ax + ay * ay + az * az, &f, &fn, &fg);
    fxc[i] = f;
    vn[i] = fn;
    gx[i] = fg * ax;
    gy[i] = fg * ay;
    gz[i] = fg * az;
}

// vn -= 2 Re(div)
__global__ void k_gga_combine(double* __restrict__ vn, const double2* __restrict__ div, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) vn[i] -= 2.0 * div[i].x;
}

// V_out = Re(V_H) + v_xc; per-block partial sums of [V_H rho, f_xc, v_in rho, (rho - rho_in)^2].
// With gga_f/gga_v (PBE fields computed before) those replace the point-wise LDA-PZ.
__global__ void k_hxc(const double2* __restrict__ vh, const double* __restrict__ rho, const double* __restrict__ rho_in,
                      const double* __restrict__ vin, const double* __restrict__ gga_f, const double* __restrict__ gga_v,
                      double* __restrict__ vout, double* __restrict__ part, size_t n) {
    __shared__ double red[4][kRedThreads];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double r = rho[i];
        double f, v;
        if (gga_f) {
            f = gga_f[i];
            v = gga_v[i];
        } else {
            lda_pz(r, &f, &v);
        }
        const double h = vh[i].x;
        vout[i] = h + v;
        s0 += h * r;
        s1 += f;
        s2 += vin[i] * r;
        const double d = r - rho_in[i];
        s3 += d * d;
    }
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    red[2][threadIdx.x] = s2;
    red[3][threadIdx.x] = s3;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// part[b] = sum over block b of a[i] * c[i] (fixed order, reduced on the host).
__global__ void k_dot(const double* __restrict__ a, const double* __restrict__ c, double* __restrict__ part,
                      size_t n) {
    __shared__ double red[kRedThreads];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += a[i] * c[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_fill(double* __restrict__ a, double v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// out = a + s b (out may alias a)
__global__ void k_axpy_to(const double* a, const double* __restrict__ b, double s, double* out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] + s * b[i];
}

// r = vout - vin
__global__ void k_sub(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

// w(G) *= P(G), P = (l^2 G^2 + 1/eps) / (l^2 G^2 + 1), P(0) = 1/eps (mixing.zig:309-338); the
// forward transform is unnormalised and the inverse carries the 1/N, so only P is applied.
__global__ void k_dielectric(double2* __restrict__ w, const double* __restrict__ g2, double inv_eps, double l2,
                             size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double g = g2[i];
    const double x = l2 * g;
    const double p = g < 1e-12 ? inv_eps : (x + inv_eps) / (x + 1.0);
    w[i] = make_double2(w[i].x * p, w[i].y * p);
}

// out = Re(w)
__global__ void k_real_part(const double2* __restrict__ w, double* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = w[i].x;
}

// v = sum_j c_j (hv[slot_j] + beta hr[slot_j])
__global__ void k_pulay_combine(const double* __restrict__ hv, const double* __restrict__ hr, const int* __restrict__ slot,
                                const double* __restrict__ c, int m, double beta, double* __restrict__ v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int j = 0; j < m; ++j) {
        const size_t o = (size_t)slot[j] * n + i;
        acc += c[j] * (hv[o] + beta * hr[o]);
    }
    v[i] = acc;
}

unsigned grid_blocks(size_t n) { return (unsigned)((n + 255) / 256); }
constexpr int kRedBlocks = 512;

}  // namespace

double ScfGrid::reduce(const double* part_dev, int nblocks, int stride, int comp) {
    std::vector<double> h((size_t)nblocks * stride);
    DZ_HIP(hipMemcpyAsync(h.data(), part_dev, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream()));
    DZ_HIP(hipStreamSynchronize(stream()));
    double s = 0.0;
    for (int i = 0; i < nblocks; ++i) s += h[(size_t)i * stride + comp];
    return s;
}

double ScfGrid::dot(const double* a, const double* c) {
    part.ensure(kRedBlocks * sizeof(double));
    hipLaunchKernelGGL(k_dot, dim3(kRedBlocks), dim3(kRedThreads), 0, stream(), a, c, part.as<double>(), n);
    DZ_HIP(hipGetLastError());
    return reduce(part.as<double>(), kRedBlocks, 1, 0);
}

void ScfGrid::gga(const double* rho) {
    hipStream_t s = stream();
    const unsigned nb = grid_blocks(n);
    double* gc[3] = {grad.as<double>(), grad.as<double>() + n, grad.as<double>() + 2 * n};
    double2* w = work.as<double2>();
    double2* w2 = work2.as<double2>();
    // grad n = IFFT(i G FFT(n)), real part kept.
    hipLaunchKernelGGL(k_real_to_complex, dim3(nb), dim3(256), 0, s, rho, w, n);
    DZ_HIP(hipGetLastError());
    fft3d_device<double>(w, tab, false, s);
    for (int c = 0; c < 3; ++c) {
        hipLaunchKernelGGL(k_grad_comp, dim3(nb), dim3(256), 0, s, w, w2, c, nx, ny, nz, recip_rows[0], recip_rows[1],
                           recip_rows[2]);
        DZ_HIP(hipGetLastError());
        fft3d_device<double>(w2, tab, true, s);
        hipLaunchKernelGGL(k_real_part, dim3(nb), dim3(256), 0, s, w2, gc[c], n);
        DZ_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_pbe_point, dim3(nb), dim3(256), 0, s, rho, gc[0], gc[1], gc[2], gga_f.as<double>(),
                       gga_v.as<double>(), n);
    DZ_HIP(hipGetLastError());
    // div b = IFFT(sum_c i G_c FFT(b_c)), then v_xc = df/dn - 2 Re(div b).
    for (int c = 0; c < 3; ++c) {
        hipLaunchKernelGGL(k_real_to_complex, dim3(nb), dim3(256), 0, s, gc[c], w, n);
        DZ_HIP(hipGetLastError());
        fft3d_device<double>(w, tab, false, s);
        hipLaunchKernelGGL(k_div_acc, dim3(nb), dim3(256), 0, s, w, w2, c, c == 0 ? 1 : 0, nx, ny, nz,
                           recip_rows[0], recip_rows[1], recip_rows[2]);
        DZ_HIP(hipGetLastError());
    }
    fft3d_device<double>(w2, tab, true, s);
    hipLaunchKernelGGL(k_gga_combine, dim3(nb), dim3(256), 0, s, gga_v.as<double>(), w2, n);
    DZ_HIP(hipGetLastError());
}

void ScfGrid::hxc(const double* rho, const double* rho_ref, const double* vin, double* vout, double out[4]) {
    hipStream_t s = stream();
    if (xc == DZHIP_XC_PBE) gga(rho);
    hipLaunchKernelGGL(k_real_to_complex, dim3(grid_blocks(n)), dim3(256), 0, s, rho, work.as<double2>(), n);
    DZ_HIP(hipGetLastError());
    fft3d_device<double>(work.as<double2>(), tab, false, s);
    hipLaunchKernelGGL(k_hartree_kernel, dim3(grid_blocks(n)), dim3(256), 0, s, work.as<double2>(), g2.as<double>(), n);
    DZ_HIP(hipGetLastError());
    fft3d_device<double>(work.as<double2>(), tab, true, s);
    part.ensure(4 * kRedBlocks * sizeof(double));
    const bool is_gga = xc == DZHIP_XC_PBE;
    hipLaunchKernelGGL(k_hxc, dim3(kRedBlocks), dim3(kRedThreads), 0, s, work.as<double2>(), rho, rho_ref, vin,
                       is_gga ? gga_f.as<double>() : nullptr, is_gga ? gga_v.as<double>() : nullptr, vout,
                       part.as<double>(), n);
    DZ_HIP(hipGetLastError());
    std::vector<double> h((size_t)kRedBlocks * 4);
    DZ_HIP(hipMemcpyAsync(h.data(), part.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    DZ_HIP(hipStreamSynchronize(s));
    double t[4] = {0, 0, 0, 0};
    for (int i = 0; i < kRedBlocks; ++i)
        for (int c = 0; c < 4; ++c) t[c] += h[4 * i + c];
    const double dv = volume / (double)n;
    out[0] = 0.5 * t[0] * dv;
    out[1] = t[1] * dv;
    out[2] = t[2] * dv;
    out[3] = std::sqrt(t[3] / (double)n);
}

ScfGrid* scf_grid_create(Context* ctx, int nx, int ny, int nz, const double* recip, double volume, int history) {
    auto* g = new ScfGrid();
    try {
        g->ctx = ctx;
        g->nx = nx;
        g->ny = ny;
        g->nz = nz;
        g->n = (size_t)nx * ny * nz;
        g->volume = volume;
        g->hist_cap = history;
        for (int r = 0; r < 3; ++r) g->recip_rows[r] = make_double3(recip[3 * r], recip[3 * r + 1], recip[3 * r + 2]);
        g->tab.build(nx, ny, nz);
        const size_t n = g->n;
        g->g2.ensure(n * sizeof(double));
        g->work.ensure(n * sizeof(double2));
        DevBuf* bufs[] = {&g->rho_in, &g->rho_out, &g->v_in, &g->v_out, &g->v_ion, &g->local};
        for (DevBuf* b : bufs) b->ensure(n * sizeof(double));
        if (history > 0) {
            g->hv.ensure((size_t)history * n * sizeof(double));
            g->hr.ensure((size_t)history * n * sizeof(double));
            g->slots.ensure(history * sizeof(int));
            g->coef.ensure(history * sizeof(double));
        }
        g->b.assign(std::max(history, 1), std::vector<double>(std::max(history, 1), 0.0));
        hipStream_t s = g->stream();
        hipLaunchKernelGGL(k_g2, dim3(grid_blocks(n)), dim3(256), 0, s, g->g2.as<double>(), nx, ny, nz,
                           make_double3(recip[0], recip[1], recip[2]), make_double3(recip[3], recip[4], recip[5]),
                           make_double3(recip[6], recip[7], recip[8]));
        DZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_fill, dim3(grid_blocks(n)), dim3(256), 0, s, g->v_ion.as<double>(), 0.0, n);
        DZ_HIP(hipGetLastError());
        DZ_HIP(hipStreamSynchronize(s));
    } catch (...) {
        scf_grid_destroy(g);
        throw;
    }
    return g;
}

void scf_grid_destroy(ScfGrid* g) {
    if (!g) return;
    DevBuf* bufs[] = {&g->g2, &g->work, &g->rho_in, &g->rho_out, &g->v_in, &g->v_out, &g->v_ion, &g->local,
                      &g->part, &g->hv, &g->hr, &g->slots, &g->coef, &g->work2, &g->grad, &g->gga_f, &g->gga_v};
    for (DevBuf* b : bufs) b->release();
    g->tab.release();
    delete g;
}

void scf_grid_init(ScfGrid& g, double nelec, const double* rho_host) {
    hipStream_t s = g.stream();
    if (rho_host) {
        DZ_HIP(hipMemcpyAsync(g.rho_in.ptr, rho_host, g.n * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        hipLaunchKernelGGL(k_fill, dim3(grid_blocks(g.n)), dim3(256), 0, s, g.rho_in.as<double>(), nelec / g.volume, g.n);
        DZ_HIP(hipGetLastError());
    }
    double e[4];
    // v_in = V_Hxc[rho_in]; the reference arguments for rho_ref/vin only feed the unused sums.
    g.hxc(g.rho_in.as<double>(), g.rho_in.as<double>(), g.rho_in.as<double>(), g.v_in.as<double>(), e);
    g.order.clear();
    hipLaunchKernelGGL(k_fill, dim3(grid_blocks(g.n)), dim3(256), 0, s, g.rho_out.as<double>(), 0.0, g.n);
    DZ_HIP(hipGetLastError());
    DZ_HIP(hipStreamSynchronize(s));
}

void scf_grid_add(ScfGrid& g, const double* rho_k) {
    hipLaunchKernelGGL(k_axpy_to, dim3(grid_blocks(g.n)), dim3(256), 0, g.stream(), g.rho_out.as<double>(), rho_k, 1.0,
                       g.rho_out.as<double>(), g.n);
    DZ_HIP(hipGetLastError());
}

void scf_grid_local(ScfGrid& g) {
    hipLaunchKernelGGL(k_axpy_to, dim3(grid_blocks(g.n)), dim3(256), 0, g.stream(), g.v_ion.as<double>(),
                       g.v_in.as<double>(), 1.0, g.local.as<double>(), g.n);
    DZ_HIP(hipGetLastError());
}

// After the k-point densities are in rho_out: V_out, energies, drho; rho_in <- rho_out, rho_out <- 0.
void scf_grid_hxc(ScfGrid& g, double out[4]) {
    g.hxc(g.rho_out.as<double>(), g.rho_in.as<double>(), g.v_in.as<double>(), g.v_out.as<double>(), out);
    hipStream_t s = g.stream();
    DZ_HIP(hipMemcpyAsync(g.rho_in.ptr, g.rho_out.ptr, g.n * sizeof(double), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_fill, dim3(grid_blocks(g.n)), dim3(256), 0, s, g.rho_out.as<double>(), 0.0, g.n);
    DZ_HIP(hipGetLastError());
}

void scf_grid_mix(ScfGrid& g, double beta, int pulay) {
    hipStream_t s = g.stream();
    const size_t n = g.n;
    if (!pulay || g.hist_cap <= 0) {
        // v_in += beta (v_out - v_in)
        hipLaunchKernelGGL(k_sub, dim3(grid_blocks(n)), dim3(256), 0, s, g.v_out.as<double>(), g.v_in.as<double>(),
                           g.work.as<double>(), n);
        DZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_axpy_to, dim3(grid_blocks(n)), dim3(256), 0, s, g.v_in.as<double>(), g.work.as<double>(),
                           beta, g.v_in.as<double>(), n);
        DZ_HIP(hipGetLastError());
        return;
    }
    // Append (v_in, R) into a free or the oldest slot.
    int slot;
    if ((int)g.order.size() < g.hist_cap) {
        slot = (int)g.order.size();
    } else {
        slot = g.order.front();
        g.order.erase(g.order.begin());
    }
    g.order.push_back(slot);
    double* hv = g.hv.as<double>() + (size_t)slot * n;
    double* hr = g.hr.as<double>() + (size_t)slot * n;
    DZ_HIP(hipMemcpyAsync(hv, g.v_in.ptr, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_sub, dim3(grid_blocks(n)), dim3(256), 0, s, g.v_out.as<double>(), g.v_in.as<double>(), hr, n);
    DZ_HIP(hipGetLastError());
    if (g.diemac > 1.0) {
        // R <- P R through the FFT (scf.zig:164-178 applies it in G space before the Pulay step).
        hipLaunchKernelGGL(k_real_to_complex, dim3(grid_blocks(n)), dim3(256), 0, s, hr, g.work.as<double2>(), n);
        DZ_HIP(hipGetLastError());
        fft3d_device<double>(g.work.as<double2>(), g.tab, false, s);
        hipLaunchKernelGGL(k_dielectric, dim3(grid_blocks(n)), dim3(256), 0, s, g.work.as<double2>(), g.g2.as<double>(),
                           1.0 / g.diemac, g.dielng * g.dielng, n);
        DZ_HIP(hipGetLastError());
        fft3d_device<double>(g.work.as<double2>(), g.tab, true, s);
        hipLaunchKernelGGL(k_real_part, dim3(grid_blocks(n)), dim3(256), 0, s, g.work.as<double2>(), hr, n);
        DZ_HIP(hipGetLastError());
    }
    for (int j : g.order) {
        const double d = g.dot(hr, g.hr.as<double>() + (size_t)j * n) / (double)n;
        g.b[slot][j] = d;
        g.b[j][slot] = d;
    }
    const int m = (int)g.order.size();
    auto linear = [&]() {
        hipLaunchKernelGGL(k_axpy_to, dim3(grid_blocks(n)), dim3(256), 0, s, g.v_in.as<double>(), hr, beta,
                           g.v_in.as<double>(), n);
        DZ_HIP(hipGetLastError());
    };
    if (m < 2) {
        linear();
        return;
    }
    // [B -1; -1 0] c = [0; -1], B scaled by its largest diagonal; partial-pivot elimination.
    const int sz = m + 1;
    std::vector<double> a((size_t)sz * sz, 0.0), rhs(sz, 0.0);
    double dmax = 0.0;
    for (int i = 0; i < m; ++i) dmax = std::max(dmax, g.b[g.order[i]][g.order[i]]);
    dmax = std::max(dmax, 1e-300);
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) a[(size_t)i * sz + j] = g.b[g.order[i]][g.order[j]] / dmax;
        a[(size_t)i * sz + m] = -1.0;
        a[(size_t)m * sz + i] = -1.0;
    }
    rhs[m] = -1.0;
    bool singular = false;
    for (int col = 0; col < sz && !singular; ++col) {
        int p = col;
        for (int r = col + 1; r < sz; ++r)
            if (std::fabs(a[(size_t)r * sz + col]) > std::fabs(a[(size_t)p * sz + col])) p = r;
        if (p != col) {
            for (int j = 0; j < sz; ++j) std::swap(a[(size_t)col * sz + j], a[(size_t)p * sz + j]);
            std::swap(rhs[col], rhs[p]);
        }
        const double piv = a[(size_t)col * sz + col];
        if (std::fabs(piv) < 1e-12) {
            singular = true;
            break;
        }
        for (int r = col + 1; r < sz; ++r) {
            const double f = a[(size_t)r * sz + col] / piv;
            for (int j = col; j < sz; ++j) a[(size_t)r * sz + j] -= f * a[(size_t)col * sz + j];
            rhs[r] -= f * rhs[col];
        }
    }
    if (singular) {
        g.order.assign(1, slot);
        linear();
        return;
    }
    std::vector<double> x(sz, 0.0);
    for (int i = sz - 1; i >= 0; --i) {
        double acc = rhs[i];
        for (int j = i + 1; j < sz; ++j) acc -= a[(size_t)i * sz + j] * x[j];
        x[i] = acc / a[(size_t)i * sz + i];
    }
    DZ_HIP(hipMemcpyAsync(g.slots.ptr, g.order.data(), m * sizeof(int), hipMemcpyHostToDevice, s));
    DZ_HIP(hipMemcpyAsync(g.coef.ptr, x.data(), m * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pulay_combine, dim3(grid_blocks(n)), dim3(256), 0, s, g.hv.as<double>(), g.hr.as<double>(),
                       g.slots.as<int>(), g.coef.as<double>(), m, beta, g.v_in.as<double>(), n);
    DZ_HIP(hipGetLastError());
    DZ_HIP(hipStreamSynchronize(s));  // slots/coef are host vectors
}

void scf_grid_set_xc(ScfGrid& g, int functional) {
    if (functional == DZHIP_XC_PBE) {
        g.work2.ensure(g.n * sizeof(double2));
        g.grad.ensure(3 * g.n * sizeof(double));
        g.gga_f.ensure(g.n * sizeof(double));
        g.gga_v.ensure(g.n * sizeof(double));
    }
    g.xc = functional;
}

}  // namespace dzhip
