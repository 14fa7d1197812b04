// Lowest eigenpairs of a dense Hermitian matrix on the device (the Rayleigh-Ritz step of the
// block Davidson and CG solvers; src/features/linalg/lobpcg/common.zig:440-537 calls LAPACK
// zheev there).  Own design for one GPU; see heev.h for the four stages.
#include "heev.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>

namespace dzhip {

namespace {

constexpr int kTrdThreads = 256;
constexpr int kTrdColsPerGroup = 6;   // trailing-matrix columns held in each workgroup's LDS
constexpr int kTrdMaxCols = 16;
constexpr unsigned kSpinLimit = 1u << 22;  // barrier polls (each after s_sleep) before giving up

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum of n values over the workgroup (NT threads); every thread gets the n sums in out[].
// red holds (NT / 64) * n doubles.
template <int NT, int N>
__device__ __forceinline__ void block_sums(double (&vals)[N], double* red, double (&out)[N]) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < N; ++k) vals[k] = wave_sum(vals[k]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) red[wv * N + k] = vals[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) s += red[i * N + k];
        out[k] = s;
    }
}

// Barrier over the whole (co-resident) grid.  Each workgroup owns a flag on its own 64-byte line
// and stores the barrier's epoch there (no read-modify-write on a shared counter, whose atomics
// serialise: about 70 ns per arriving workgroup); wavefront 0 of every workgroup then polls all
// flags (64 per load instruction) until each has reached the epoch.  The global writes before
// the barrier are released at device scope and acquired after it.  A wait longer than
// kSpinLimit polls (or an error raised elsewhere) sets *err; every thread returns false then and
// the kernel winds down.
__device__ bool grid_sync(unsigned* flags, unsigned epoch, int* err, int* s_ok) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = (int)threadIdx.x;
        const int G = (int)gridDim.x;
        if (lane == 0)
            __hip_atomic_store(flags + (size_t)blockIdx.x * 16, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        int ok = 1;
        while (true) {
            int arrived = 1;
            for (int q = lane; q < G; q += 64)
                arrived &= __hip_atomic_load(flags + (size_t)q * 16, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= epoch;
            if (__all(arrived)) break;
            const int bad = lane == 0 ? __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            if (__any(bad) || ++spins > kSpinLimit) {
                if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (lane == 0) *s_ok = ok;
        __threadfence();
    }
    __syncthreads();
    __threadfence();
    return *s_ok != 0;
}

__device__ __forceinline__ double2 cm(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// conj(a) * b
__device__ __forceinline__ double2 cjm(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

struct TrdArgs {
    const double2* A;
    int lda, m, ncl;
    double2* vh;   // m x m: column j holds reflector v_j in rows j+1..m-1 (v_j[j+1] = 1)
    double2* tau;  // m
    double* d;     // m
    double* e;     // m - 1
    double2* pbuf; // 2 x m: tau_j A v_j, double-buffered by step parity
    double2* nbuf; // 2 x m: the next column after the previous rank-2 update
    unsigned* flags;  // one per workgroup, 16 apart
    int* err;
};

// Householder tridiagonalisation A = Q T Q^H, Q = H_0 ... H_{m-2}, H_j = I - tau_j v_j v_j^H, with
// zhetd2's reflectors (zlarfg: H_j^H x = beta e_1, beta real).  Workgroup g holds the trailing
// columns i = g + c G (c < ncl) in LDS for the whole reduction.  Step j:
//   A. every workgroup forms v_j, tau_j from the current column j (in LDS);
//   B. fused pass over its own columns i > j: the rank-2 update of step j-1
//      a_i -= v_{j-1} conj(w_{j-1}[i]) + w_{j-1} conj(v_{j-1}[i]), then p_j[i] = tau_j a_i^H v_j
//      (A is Hermitian, so (A v)_i = a_i^H v); the owner of column j+1 publishes it (updated);
//   C. grid barrier;
//   D. w_j = p_j - (tau_j / 2) (p_j^H v_j) v_j, from the published p_j;
//   E. column j+1 = published column - v_j conj(w_j[j+1]) - w_j conj(v_j[j+1]).
// The owner never updates column j+1 in place at step j, so the published copy stays valid while
// the others read it; pbuf and nbuf alternate by step parity.
// NCL (columns per workgroup) is a template parameter so the per-column partial sums stay in
// registers.
template <int NCL>
__global__ __launch_bounds__(kTrdThreads) void k_hetrd(TrdArgs a) {
    extern __shared__ double2 lds[];
    __shared__ double red[(kTrdThreads / 64) * 2 * kTrdMaxCols];
    __shared__ int s_ok;
    constexpr int ncl = NCL;
    const int m = a.m, G = (int)gridDim.x, g = (int)blockIdx.x, t = (int)threadIdx.x;
    constexpr int NT = kTrdThreads;
    double2* cols = lds;
    // Roles, rotated every step: vc = current column (becomes v_j in place), wn = w_j,
    // vp / wp = v_{j-1}, w_{j-1}.
    double2* vc_b = cols + (size_t)ncl * m;
    double2* wn_b = cols + (size_t)(ncl + 1) * m;
    double2* vp_b = cols + (size_t)(ncl + 2) * m;
    double2* wp_b = cols + (size_t)(ncl + 3) * m;
    // Only the lower triangle of A is read (the upper one is its mirror; the diagonal is real), as
    // LAPACK's zheev with uplo = 'L'.
    for (int c = 0; c < ncl; ++c) {
        const int i = g + c * G;
        if (i >= m) break;
        for (int r = t; r < m; r += NT) {
            double2 x;
            if (r > i) x = a.A[(size_t)i * a.lda + r];
            else if (r == i) x = make_double2(a.A[(size_t)i * a.lda + r].x, 0.0);
            else {
                const double2 y = a.A[(size_t)r * a.lda + i];
                x = make_double2(y.x, -y.y);
            }
            cols[(size_t)c * m + r] = x;
        }
    }
    for (int r = t; r < m; r += NT) vc_b[r] = r == 0 ? make_double2(a.A[0].x, 0.0) : a.A[r];
    __syncthreads();
    for (int j = 0; j + 1 < m; ++j) {
        double2* vc = vc_b;
        // A. reflector of column j (rows j+1..m-1).
        double xs[1] = {0.0};
        for (int r = j + 2 + t; r < m; r += NT) xs[0] += vc[r].x * vc[r].x + vc[r].y * vc[r].y;
        double xo[1];
        block_sums<NT>(xs, red, xo);
        const double xn2 = xo[0];
        const double2 alpha = vc[j + 1];
        double beta;
        double2 tau, scale;
        if (xn2 == 0.0 && alpha.y == 0.0) {
            tau = make_double2(0.0, 0.0);
            beta = alpha.x;
            scale = make_double2(0.0, 0.0);
        } else {
            const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
            beta = alpha.x >= 0.0 ? -nrm : nrm;
            tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
            const double2 den = make_double2(alpha.x - beta, alpha.y);
            const double dd = den.x * den.x + den.y * den.y;
            scale = make_double2(den.x / dd, -den.y / dd);
        }
        if (g == 0 && t == 0) {
            a.d[j] = vc[j].x;
            a.e[j] = beta;
            a.tau[j] = tau;
        }
        __syncthreads();  // every thread has read alpha before vc[j + 1] becomes 1
        for (int r = j + 2 + t; r < m; r += NT) vc[r] = cm(vc[r], scale);
        if (t == 0) vc[j + 1] = make_double2(1.0, 0.0);
        __syncthreads();
        if (g == 0)
            for (int r = j + 1 + t; r < m; r += NT) a.vh[(size_t)j * m + r] = vc[r];
        // B. fused update (step j-1) and product (step j) over the own columns i >= j+1.
        const double2* vp = vp_b;
        const double2* wp = wp_b;
        double part[2 * ncl];
#pragma unroll
        for (int c = 0; c < 2 * ncl; ++c) part[c] = 0.0;
        double2* nb_out = a.nbuf + (size_t)((j + 1) & 1) * m;
#pragma unroll
        for (int c = 0; c < ncl; ++c) {
            const int i = g + c * G;
            if (i <= j || i >= m) continue;
            double2* col = cols + (size_t)c * m;
            const double2 wi = j > 0 ? wp[i] : make_double2(0.0, 0.0);
            const double2 vi = j > 0 ? vp[i] : make_double2(0.0, 0.0);
            double2 acc = make_double2(0.0, 0.0);
            for (int r = j + 1 + t; r < m; r += NT) {
                double2 x = col[r];
                if (j > 0) {
                    // x -= vp[r] conj(wi) + wp[r] conj(vi)
                    const double2 u1 = cjm(wi, vp[r]);
                    const double2 u2 = cjm(vi, wp[r]);
                    x.x -= u1.x + u2.x;
                    x.y -= u1.y + u2.y;
                    col[r] = x;
                }
                const double2 pr = cjm(x, vc[r]);
                acc.x += pr.x;
                acc.y += pr.y;
                if (i == j + 1) nb_out[r] = x;
            }
            part[2 * c] = acc.x;
            part[2 * c + 1] = acc.y;
        }
        double sums[2 * ncl];
        block_sums<NT>(part, red, sums);
        double2* pb = a.pbuf + (size_t)(j & 1) * m;
#pragma unroll
        for (int c = 0; c < ncl; ++c) {
            const int i = g + c * G;
            if (t == c && i > j && i < m) pb[i] = cm(tau, make_double2(sums[2 * c], sums[2 * c + 1]));
        }
        // C.
        if (!grid_sync(a.flags, (unsigned)(j + 1), a.err, &s_ok)) return;
        // D. w_j = p_j - (tau_j / 2) (p_j^H v_j) v_j.
        double2* wn = wn_b;
        double ps[2] = {0.0, 0.0};
        for (int r = j + 1 + t; r < m; r += NT) {
            const double2 p = pb[r];
            wn[r] = p;
            const double2 pr = cjm(p, vc[r]);
            ps[0] += pr.x;
            ps[1] += pr.y;
        }
        double s2[2];
        block_sums<NT>(ps, red, s2);
        const double2 al = cm(make_double2(-0.5 * tau.x, -0.5 * tau.y), make_double2(s2[0], s2[1]));
        for (int r = j + 1 + t; r < m; r += NT) {
            const double2 u = cm(al, vc[r]);
            wn[r].x += u.x;
            wn[r].y += u.y;
        }
        __syncthreads();
        // E. column j+1 after the step-j update, into the buffer of v_{j-1} (no longer needed).
        double2* vnext = vp_b;
        const double2* nb_in = a.nbuf + (size_t)((j + 1) & 1) * m;
        const double2 wj1 = wn[j + 1], vj1 = vc[j + 1];
        for (int r = j + 1 + t; r < m; r += NT) {
            const double2 x = nb_in[r];
            const double2 u1 = cjm(wj1, vc[r]);
            const double2 u2 = cjm(vj1, wn[r]);
            vnext[r] = make_double2(x.x - u1.x - u2.x, x.y - u1.y - u2.y);
        }
        __syncthreads();
        if (j + 2 == m && g == 0 && t == 0) a.d[m - 1] = vnext[m - 1].x;
        // Rotate: v_{j-1} <- v_j, w_{j-1} <- w_j, current column <- column j+1.
        double2* const old_vp = vp_b;
        double2* const old_wp = wp_b;
        vp_b = vc_b;
        wp_b = wn_b;
        vc_b = old_vp;
        wn_b = old_wp;
    }
    if (m == 1 && g == 0 && t == 0) a.d[0] = a.A[0].x;
}

__global__ void k_zero_u32(unsigned* p, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0u;
}

// Sturm count: number of eigenvalues of T(d, e) below x (e2 = e^2), with tiny pivots replaced
// by -pivmin (LAPACK dlaebz's rule).
__device__ __forceinline__ int sturm_count(const double* d, const double* e2, int m, double x, double pivmin) {
    double q = d[0] - x;
    int cnt = 0;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++cnt;
    for (int k = 1; k < m; ++k) {
        q = d[k] - x - e2[k - 1] / q;
        if (fabs(q) < pivmin) q = -pivmin;
        if (q < 0.0) ++cnt;
    }
    return cnt;
}

// Eigenvalue idx (ascending) of T by multisection: each round the 64 lanes of a wavefront count
// at 64 interior points of [lo, hi] and keep the sub-interval that holds eigenvalue idx.
// Also writes the Gershgorin scale of T to scal[0] (block 0).
__global__ __launch_bounds__(256) void k_stebz(const double* __restrict__ d_in, const double* __restrict__ e_in, int m,
                                               int nev, double* __restrict__ w, double* __restrict__ scal,
                                               const int* __restrict__ err) {
    extern __shared__ double sh[];
    double* d = sh;
    double* e2 = sh + m;
    __shared__ double red[4 * 3];
    const int t = (int)threadIdx.x;
    double lo_l = 1e300, hi_l = -1e300, e2max = 0.0;
    for (int k = t; k < m; k += 256) {
        const double dk = d_in[k];
        const double el = k > 0 ? fabs(e_in[k - 1]) : 0.0;
        const double er = k + 1 < m ? fabs(e_in[k]) : 0.0;
        d[k] = dk;
        e2[k] = k + 1 < m ? e_in[k] * e_in[k] : 0.0;
        lo_l = fmin(lo_l, dk - el - er);
        hi_l = fmax(hi_l, dk + el + er);
        e2max = fmax(e2max, er * er);
    }
    // Block min / max / max via negation into a sum-free reduction.
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo_l = fmin(lo_l, __shfl_xor(lo_l, o, 64));
        hi_l = fmax(hi_l, __shfl_xor(hi_l, o, 64));
        e2max = fmax(e2max, __shfl_xor(e2max, o, 64));
    }
    if ((t & 63) == 0) {
        red[(t >> 6) * 3] = lo_l;
        red[(t >> 6) * 3 + 1] = hi_l;
        red[(t >> 6) * 3 + 2] = e2max;
    }
    __syncthreads();
    double gl = red[0], gu = red[1], em = red[2];
    for (int i = 1; i < 4; ++i) {
        gl = fmin(gl, red[i * 3]);
        gu = fmax(gu, red[i * 3 + 1]);
        em = fmax(em, red[i * 3 + 2]);
    }
    const double bnorm = fmax(fabs(gl), fabs(gu));
    const double pivmin = 2.2250738585072014e-308 * fmax(1.0, em) / 2.220446049250313e-16;
    // Widen so that the ends are strictly outside the spectrum.
    const double wid = 2.220446049250313e-16 * 4.0 * bnorm + 2.0 * pivmin + 1e-300;
    gl -= wid + 1e-14 * (gu - gl);
    gu += wid + 1e-14 * (gu - gl);
    if (blockIdx.x == 0 && t == 0) scal[0] = bnorm;
    const int lane = t & 63;
    const int idx = (int)blockIdx.x * 4 + (t >> 6);
    if (idx >= nev) return;
    // A tridiagonalisation whose grid barrier did not complete leaves NaN eigenvalues (the signal
    // a caller without a synchronising check sees).
    if (*err != 0) {
        if (lane == 0) w[idx] = __builtin_nan("");
        return;
    }
    double lo = gl, hi = gu;
    for (int round = 0; round < 14; ++round) {
        const double width = hi - lo;
        if (width <= 2.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi)) + 2.0 * pivmin) break;
        const double x = lo + width * (double)(lane + 1) / 65.0;
        const int cnt = sturm_count(d, e2, m, x, pivmin);
        const unsigned long long below = __ballot(cnt <= idx);
        // Highest lane whose point still has eigenvalue idx at or above it.
        double nlo = lo, nhi = hi;
        if (below == 0ull) {
            nhi = lo + width * 1.0 / 65.0;
        } else {
            const int L = 63 - __clzll(below);
            nlo = lo + width * (double)(L + 1) / 65.0;
            if (L < 63) nhi = lo + width * (double)(L + 2) / 65.0;
        }
        lo = nlo;
        hi = nhi;
    }
    if (lane == 0) w[idx] = 0.5 * (lo + hi);
}

// Eigenvectors of T by inverse iteration, one thread per cluster of eigenvalues closer than
// tau_c (consecutive gaps).  Members are iterated in order, each (T - lambda I) solve followed by
// Gram-Schmidt against the cluster's earlier members; three iterations from a seeded start.
// fac: 5 m doubles per eigenvalue (LU of T - lambda I with partial pivoting, as dlagtf).
__global__ __launch_bounds__(64) void k_stein(const double* __restrict__ d, const double* __restrict__ e, int m,
                                              int nev, const double* __restrict__ w, const double* __restrict__ scal,
                                              double* __restrict__ Z, int ldz, double* __restrict__ fac) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= nev) return;
    const double bnorm = fmax(scal[0], 1e-300);
    const double tau_c = 1e-12 * fmax(bnorm, 1.0);
    if (i > 0 && w[i] - w[i - 1] <= tau_c) return;
    int last = i;
    while (last + 1 < nev && w[last + 1] - w[last] <= tau_c) ++last;
    double* u0 = fac + (size_t)i * 5 * m;
    double* u1 = u0 + m;
    double* u2 = u1 + m;
    double* lm = u2 + m;
    double* pv = lm + m;
    const double pert = 2.220446049250313e-16 * bnorm + 1e-300;
    for (int k = i; k <= last; ++k) {
        const double lam = w[k];
        // Factor T - lam I = P L U (dlagtf): U has two superdiagonals.
        double ca = d[0] - lam, cb = m > 1 ? e[0] : 0.0;
        for (int r = 0; r + 1 < m; ++r) {
            const double c = e[r];
            const double na = d[r + 1] - lam;
            const double nb = r + 2 < m ? e[r + 1] : 0.0;
            if (fabs(ca) >= fabs(c)) {
                const double l = ca != 0.0 ? c / ca : 0.0;
                u0[r] = ca;
                u1[r] = cb;
                u2[r] = 0.0;
                lm[r] = l;
                pv[r] = 0.0;
                ca = na - l * cb;
                cb = nb;
            } else {
                const double l = ca / c;
                u0[r] = c;
                u1[r] = na;
                u2[r] = nb;
                lm[r] = l;
                pv[r] = 1.0;
                ca = cb - l * na;
                cb = -l * nb;
            }
        }
        u0[m - 1] = ca;
        for (int r = 0; r < m; ++r)
            if (fabs(u0[r]) < pert) u0[r] = u0[r] < 0.0 ? -pert : pert;
        double* x = Z + (size_t)k * ldz;
        unsigned long long s = 0x9E3779B97F4A7C15ull * (unsigned long long)(k + 1) + 12345ull;
        for (int r = 0; r < m; ++r) {
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            x[r] = (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        }
        for (int it = 0; it < 3; ++it) {
            // Solve P L U y = x in place.
            for (int r = 0; r + 1 < m; ++r) {
                if (pv[r] != 0.0) {
                    const double tmp = x[r];
                    x[r] = x[r + 1];
                    x[r + 1] = tmp;
                }
                x[r + 1] -= lm[r] * x[r];
            }
            x[m - 1] /= u0[m - 1];
            if (m > 1) x[m - 2] = (x[m - 2] - u1[m - 2] * x[m - 1]) / u0[m - 2];
            for (int r = m - 3; r >= 0; --r) x[r] = (x[r] - u1[r] * x[r + 1] - u2[r] * x[r + 2]) / u0[r];
            // Scale down before Gram-Schmidt (the solve amplifies by up to 1 / pert).
            double mx = 0.0;
            for (int r = 0; r < m; ++r) mx = fmax(mx, fabs(x[r]));
            const double inv_mx = mx > 0.0 ? 1.0 / mx : 1.0;
            for (int r = 0; r < m; ++r) x[r] *= inv_mx;
            for (int p = i; p < k; ++p) {
                const double* zp = Z + (size_t)p * ldz;
                double dot = 0.0;
                for (int r = 0; r < m; ++r) dot += zp[r] * x[r];
                for (int r = 0; r < m; ++r) x[r] -= dot * zp[r];
            }
            double nrm = 0.0;
            for (int r = 0; r < m; ++r) nrm += x[r] * x[r];
            nrm = sqrt(nrm);
            const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
            for (int r = 0; r < m; ++r) x[r] *= inv;
        }
    }
}

// C[0:n, 0:n] = Z^T Z (Z: m x n, leading dimension ldz), 16 x 16 tiles.
__global__ __launch_bounds__(256) void k_gram_real(const double* __restrict__ Z, int ldz, int m, int n,
                                                   double* __restrict__ C) {
    __shared__ double za[32][17], zb[32][17];
    const int tx = (int)threadIdx.x & 15, ty = (int)threadIdx.x >> 4;
    const int i0 = (int)blockIdx.x * 16, j0 = (int)blockIdx.y * 16;
    double acc = 0.0;
    for (int r0 = 0; r0 < m; r0 += 32) {
        for (int q = (int)threadIdx.x; q < 32 * 16; q += 256) {
            const int rr = q & 31, cc = q >> 5;
            const int r = r0 + rr;
            za[rr][cc] = (r < m && i0 + cc < n) ? Z[(size_t)(i0 + cc) * ldz + r] : 0.0;
            zb[rr][cc] = (r < m && j0 + cc < n) ? Z[(size_t)(j0 + cc) * ldz + r] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int rr = 0; rr < 32; ++rr) acc += za[rr][ty] * zb[rr][tx];
        __syncthreads();
    }
    const int i = i0 + ty, j = j0 + tx;
    if (i < n && j < n) C[(size_t)j * n + i] = acc;
}

// Out = Z (1.5 I - 0.5 C): one Newton-Schulz step towards orthonormal columns.
__global__ __launch_bounds__(256) void k_ns_update(const double* __restrict__ Z, int ldz, int m, int n,
                                                   const double* __restrict__ C, double* __restrict__ out) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = (int)blockIdx.y;
    if (r >= m) return;
    double acc = 1.5 * Z[(size_t)j * ldz + r];
    const double* cj = C + (size_t)j * n;
    for (int k = 0; k < n; ++k) acc -= 0.5 * cj[k] * Z[(size_t)k * ldz + r];
    out[(size_t)j * ldz + r] = acc;
}

// Y[:, col] = Q Z[:, col] = H_0 (H_1 (... H_{m-2} z)), one wavefront per column, the column held in
// registers (lane l holds rows l + 64 q).
constexpr int kUnmtrRegs = kHeevMaxM / 64;

__global__ __launch_bounds__(256) void k_unmtr(const double2* __restrict__ vh, const double2* __restrict__ tau, int m,
                                               int nev, const double* __restrict__ Z, int ldz, double2* __restrict__ Y,
                                               int ldy) {
    const int col = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = (int)(threadIdx.x & 63);
    if (col >= nev) return;
    double2 y[kUnmtrRegs];
#pragma unroll
    for (int q = 0; q < kUnmtrRegs; ++q) {
        const int r = lane + 64 * q;
        y[q] = make_double2(r < m ? Z[(size_t)col * ldz + r] : 0.0, 0.0);
    }
    const int nq = (m + 63) / 64;
    for (int k = m - 2; k >= 0; --k) {
        const double2 tk = tau[k];
        if (tk.x == 0.0 && tk.y == 0.0) continue;
        const double2* v = vh + (size_t)k * m;
        double sr = 0.0, si = 0.0;
        double2 vr[kUnmtrRegs];
#pragma unroll
        for (int q = 0; q < kUnmtrRegs; ++q) {
            const int r = lane + 64 * q;
            vr[q] = make_double2(0.0, 0.0);
            if (q < nq && r > k && r < m) {
                vr[q] = r == k + 1 ? make_double2(1.0, 0.0) : v[r];
                const double2 p = cjm(vr[q], y[q]);
                sr += p.x;
                si += p.y;
            }
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        const double2 ts = cm(tk, make_double2(sr, si));
#pragma unroll
        for (int q = 0; q < kUnmtrRegs; ++q) {
            const double2 u = cm(vr[q], ts);
            y[q].x -= u.x;
            y[q].y -= u.y;
        }
    }
#pragma unroll
    for (int q = 0; q < kUnmtrRegs; ++q) {
        const int r = lane + 64 * q;
        if (r < m) Y[(size_t)col * ldy + r] = y[q];
    }
}

size_t trd_lds_bytes(int m, int ncl) { return (size_t)(ncl + 4) * (size_t)m * sizeof(double2); }

// The tridiagonalisation kernel for ncl columns per workgroup (1..kTrdMaxCols).
const void* trd_kernel(int ncl) {
    static const void* const table[kTrdMaxCols] = {
        (const void*)k_hetrd<1>,  (const void*)k_hetrd<2>,  (const void*)k_hetrd<3>,  (const void*)k_hetrd<4>,
        (const void*)k_hetrd<5>,  (const void*)k_hetrd<6>,  (const void*)k_hetrd<7>,  (const void*)k_hetrd<8>,
        (const void*)k_hetrd<9>,  (const void*)k_hetrd<10>, (const void*)k_hetrd<11>, (const void*)k_hetrd<12>,
        (const void*)k_hetrd<13>, (const void*)k_hetrd<14>, (const void*)k_hetrd<15>, (const void*)k_hetrd<16>};
    return table[std::max(1, std::min(ncl, kTrdMaxCols)) - 1];
}

// Co-resident workgroups per device for the tridiagonalisation (one per CU at its LDS size),
// queried once per device.
struct CoopInfo {
    int groups = 0;
    int cols_per_group = 0;
};
std::mutex g_coop_mu;
CoopInfo g_coop[64];

CoopInfo coop_info(int dev) {
    DZ_REQUIRE(dev >= 0 && dev < 64, DZHIP_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_coop_mu);
    CoopInfo& c = g_coop[dev];
    if (c.groups == 0) {
        int cus = 0, coop = 0;
        DZ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        DZ_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
        DZ_REQUIRE(coop != 0, DZHIP_ERR_HIP, "device does not support cooperative launches");
        for (int c = 1; c <= kTrdMaxCols; ++c) set_max_dynamic_lds(trd_kernel(c), kMaxLdsBytes - 2048);
        // Launch-shape override for sweeps: trailing columns per workgroup.
        if (const char* env = std::getenv("DZHIP_HEEV_COLS")) c.cols_per_group = std::max(0, std::atoi(env));
        c.groups = cus;
    }
    return c;
}

// Launch shape of the tridiagonalisation: G workgroups holding ncl columns each.
void trd_shape(int m, int* G_out, int* ncl_out) {
    int dev = 0;
    DZ_HIP(hipGetDevice(&dev));
    const CoopInfo ci = coop_info(dev);
    const int per = ci.cols_per_group > 0 ? ci.cols_per_group : kTrdColsPerGroup;
    int G = std::max(1, std::min((m + per - 1) / per, ci.groups));
    int ncl = (m + G - 1) / G;
    while (trd_lds_bytes(m, ncl) > kMaxLdsBytes - 2048 && G < ci.groups) {
        ++G;
        ncl = (m + G - 1) / G;
    }
    DZ_REQUIRE(ncl <= kTrdMaxCols && trd_lds_bytes(m, ncl) <= kMaxLdsBytes - 2048, DZHIP_ERR_ARG,
               "Hermitian eigensolver: matrix too large for the tridiagonalisation's LDS");
    // Every workgroup must be co-resident (one per CU at this LDS size).
    int per_cu = 0;
    DZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trd_kernel(ncl), kTrdThreads,
                                                        trd_lds_bytes(m, ncl)));
    DZ_REQUIRE(per_cu >= 1 && G <= per_cu * ci.groups, DZHIP_ERR_HIP,
               "Hermitian eigensolver: grid does not fit on the device");
    *G_out = G;
    *ncl_out = ncl;
}

// Device buffers of one eigensolver call (from a HeevWork, or carved from the static scratch array).
struct HeevBufs {
    double2 *vh, *tau, *pbuf, *nbuf;
    double *d, *e, *z, *zt, *gram, *fac;
    unsigned* sync;  // (G + 1) * 16 words: barrier flags, then the error flag
    int err_offset;
};

void run_tridiag(const HeevBufs& b, int G, int ncl, const double2* A, int lda, int m, hipStream_t s) {
    // Barrier flags (one 64-byte line per workgroup) and the error flag after them, cleared by a
    // kernel (not a memset node) on the same stream.
    const int nsync = (G + 1) * 16;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((nsync + 255) / 256)), dim3(256), 0, s, b.sync, nsync);
    DZ_HIP(hipGetLastError());
    TrdArgs args{A, lda, m, ncl, b.vh, b.tau, b.d, b.e, b.pbuf, b.nbuf, b.sync,
                 reinterpret_cast<int*>(b.sync) + b.err_offset};
    void* kargs[] = {&args};
    DZ_HIP(hipLaunchCooperativeKernel(trd_kernel(ncl), dim3((unsigned)G), dim3(kTrdThreads), kargs,
                                      (unsigned)trd_lds_bytes(m, ncl), s));
}

void run_eig(const HeevBufs& b, int m, int nev, double* w_dev, double2* Y, int ldy, hipStream_t s) {
    double* scal = b.gram + (size_t)nev * nev;
    hipLaunchKernelGGL(k_stebz, dim3((unsigned)((nev + 3) / 4)), dim3(256), 2 * m * sizeof(double), s, b.d, b.e, m,
                       nev, w_dev, scal, reinterpret_cast<const int*>(b.sync) + b.err_offset);
    DZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_stein, dim3((unsigned)((nev + 63) / 64)), dim3(64), 0, s, b.d, b.e, m, nev, w_dev, scal, b.z,
                       m, b.fac);
    DZ_HIP(hipGetLastError());
    double* zc = b.z;
    double* zn = b.zt;
    for (int it = 0; it < 2; ++it) {
        hipLaunchKernelGGL(k_gram_real, dim3((unsigned)((nev + 15) / 16), (unsigned)((nev + 15) / 16)), dim3(256), 0, s,
                           zc, m, m, nev, b.gram);
        DZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ns_update, dim3((unsigned)((m + 255) / 256), (unsigned)nev), dim3(256), 0, s, zc, m, m,
                           nev, b.gram, zn);
        DZ_HIP(hipGetLastError());
        std::swap(zc, zn);
    }
    hipLaunchKernelGGL(k_unmtr, dim3((unsigned)((nev + 3) / 4)), dim3(256), 0, s, b.vh, b.tau, m, nev, zc, m, Y, ldy);
    DZ_HIP(hipGetLastError());
}

HeevBufs work_bufs(HeevWork& w, int m, int nev, int G) {
    w.vh.ensure((size_t)m * m * sizeof(double2));
    w.tau.ensure((size_t)m * sizeof(double2));
    w.d.ensure((size_t)m * sizeof(double));
    w.e.ensure((size_t)m * sizeof(double));
    w.pbuf.ensure((size_t)2 * m * sizeof(double2));
    w.nbuf.ensure((size_t)2 * m * sizeof(double2));
    w.sync.ensure((size_t)(G + 1) * 16 * sizeof(unsigned));
    if (nev > 0) {
        w.z.ensure((size_t)m * nev * sizeof(double));
        w.zt.ensure((size_t)m * nev * sizeof(double));
        w.gram.ensure((size_t)nev * nev * sizeof(double) + 64);
        w.fac.ensure((size_t)nev * 5 * m * sizeof(double));
    }
    w.err_offset = G * 16;
    return HeevBufs{w.vh.as<double2>(), w.tau.as<double2>(), w.pbuf.as<double2>(), w.nbuf.as<double2>(),
                    w.d.as<double>(),   w.e.as<double>(),     w.z.as<double>(),     w.zt.as<double>(),
                    w.gram.as<double>(), w.fac.as<double>(),  w.sync.as<unsigned>(), G * 16};
}

// ----------------------------------------------------------------- allocation-free variant
// A file-scope device array holds every buffer of an eigensolver call of m <= kHeevNoAllocMaxM, for
// callers that cannot allocate.  Its device address is taken once per device at set-up
// (heev_scratch_init, from dzhip_ctx_create with the device current) and only that address is used
// on the host.  Calls that overlap on one device share the array and race.
constexpr size_t kHeevNoAllocBytes = 6u << 20;
static __device__ __align__(256) unsigned char scratch[kHeevNoAllocBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

size_t carve(size_t& off, size_t bytes) {
    const size_t at = off;
    off = round_up(off + bytes, 256);
    return at;
}

// Offsets of the buffers in the scratch for (m, nev, G); returns the bytes needed.
size_t scratch_layout(int m, int nev, int G, size_t out[11]) {
    size_t off = 0;
    out[0] = carve(off, (size_t)m * m * sizeof(double2));       // vh
    out[1] = carve(off, (size_t)m * sizeof(double2));           // tau
    out[2] = carve(off, (size_t)2 * m * sizeof(double2));       // pbuf
    out[3] = carve(off, (size_t)2 * m * sizeof(double2));       // nbuf
    out[4] = carve(off, (size_t)m * sizeof(double));            // d
    out[5] = carve(off, (size_t)m * sizeof(double));            // e
    out[6] = carve(off, (size_t)m * nev * sizeof(double));      // z
    out[7] = carve(off, (size_t)m * nev * sizeof(double));      // zt
    out[8] = carve(off, (size_t)nev * nev * sizeof(double) + 64);  // gram + scale
    out[9] = carve(off, (size_t)nev * 5 * m * sizeof(double));  // fac
    out[10] = carve(off, (size_t)(G + 1) * 16 * sizeof(unsigned));  // sync
    return off;
}

}  // namespace

void heev_scratch_init(int device) {
    DZ_REQUIRE(device >= 0 && device < 64, DZHIP_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    DZ_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    DZ_REQUIRE(ptr != nullptr, DZHIP_ERR_HIP, "no device address for the eigensolver scratch");
    g_scratch_ptr[device] = ptr;
}

void heev_lowest_noalloc(int device, const double2* A, int lda, int m, int nev, double* w_dev, double2* Y, int ldy,
                         hipStream_t s) {
    DZ_REQUIRE(m >= 1 && m <= kHeevNoAllocMaxM && nev >= 1 && nev <= m && lda >= m && ldy >= m, DZHIP_ERR_ARG,
               "allocation-free Hermitian eigensolver: need 1 <= nev <= m <= 256 (its scratch is 6 MiB) and "
               "leading dimensions >= m");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = device >= 0 && device < 64 ? g_scratch_ptr[device] : nullptr;
    }
    DZ_REQUIRE(base, DZHIP_ERR_STATE, "eigensolver scratch not set up on this device (dzhip_ctx_create)");
    int G = 1, ncl = 1;
    trd_shape(m, &G, &ncl);
    size_t off[11];
    const size_t need = scratch_layout(m, nev, G, off);
    DZ_REQUIRE(need <= kHeevNoAllocBytes, DZHIP_ERR_ARG, "allocation-free Hermitian eigensolver: scratch too small");
    unsigned char* p = static_cast<unsigned char*>(base);
    const HeevBufs b{reinterpret_cast<double2*>(p + off[0]), reinterpret_cast<double2*>(p + off[1]),
                     reinterpret_cast<double2*>(p + off[2]), reinterpret_cast<double2*>(p + off[3]),
                     reinterpret_cast<double*>(p + off[4]),  reinterpret_cast<double*>(p + off[5]),
                     reinterpret_cast<double*>(p + off[6]),  reinterpret_cast<double*>(p + off[7]),
                     reinterpret_cast<double*>(p + off[8]),  reinterpret_cast<double*>(p + off[9]),
                     reinterpret_cast<unsigned*>(p + off[10]), G * 16};
    run_tridiag(b, G, ncl, A, lda, m, s);
    run_eig(b, m, nev, w_dev, Y, ldy, s);
}

void HeevWork::release() {
    DevBuf* bufs[] = {&vh, &tau, &d, &e, &pbuf, &nbuf, &sync, &z, &zt, &gram, &fac};
    for (DevBuf* b : bufs) b->release();
}

void heev_tridiag_only(HeevWork& w, const double2* A, int lda, int m, hipStream_t s) {
    DZ_REQUIRE(m >= 1 && m <= kHeevMaxM && lda >= m, DZHIP_ERR_ARG, "Hermitian eigensolver: bad size");
    int G = 1, ncl = 1;
    trd_shape(m, &G, &ncl);
    run_tridiag(work_bufs(w, m, 0, G), G, ncl, A, lda, m, s);
}

void heev_lowest(HeevWork& w, const double2* A, int lda, int m, int nev, double* w_dev, double2* Y, int ldy,
                 hipStream_t s) {
    DZ_REQUIRE(m >= 1 && m <= kHeevMaxM && nev >= 1 && nev <= m && lda >= m && ldy >= m, DZHIP_ERR_ARG,
               "Hermitian eigensolver: need 1 <= nev <= m <= 1024 and leading dimensions >= m");
    int G = 1, ncl = 1;
    trd_shape(m, &G, &ncl);
    const HeevBufs b = work_bufs(w, m, nev, G);
    run_tridiag(b, G, ncl, A, lda, m, s);
    run_eig(b, m, nev, w_dev, Y, ldy, s);
}

int heev_error_flag(HeevWork& w, hipStream_t s) {
    int flag = 0;
    if (!w.sync.ptr) return 0;
    DZ_HIP(hipMemcpyAsync(&flag, w.sync.as<int>() + w.err_offset, sizeof(int), hipMemcpyDeviceToHost, s));
    DZ_HIP(hipStreamSynchronize(s));
    return flag;
}

}  // namespace dzhip
