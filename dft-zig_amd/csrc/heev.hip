// Lowest eigenpairs of a dense Hermitian matrix on the device (the Rayleigh-Ritz step of the
// block Davidson and CG solvers; src/features/linalg/lobpcg/common.zig:440-537 calls LAPACK
// zheev there).  Own design for one GPU; see heev.h for the four stages.
#include "heev.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <mutex>

namespace dzhip {

namespace {

constexpr int kTrdThreads = 256;
constexpr int kTrdColsPerGroup = 6;   // trailing-matrix columns held in each workgroup's LDS
constexpr int kTrdMaxCols = 16;
constexpr int kTrdRowsPerThread = (1024 + 255) / 256;  // rows per thread at m <= kHeevMaxM
constexpr unsigned kSpinLimit = 1u << 22;  // barrier polls (each after s_sleep) before giving up
// Barrier polls of a tridiagonalisation admitted by the gate: its grid may find the device short
// of CUs for a long time while other streams' kernels keep taking them, and it is then repeated as
// a cooperative launch (heev_lowest_shared).  Si-64 concurrent-k SCF, mixed arm to 1e-6 Ha: 9.16 s
// at 2^16 polls (frequent repeats), 8.60 s at 2^19, 10.21 s with cooperative launches only.
constexpr unsigned kSpinLimitShared = 1u << 19;

// One DPP row rotation of a double (both halves moved by the same row_ror:N control).
template <int CTRL>
__device__ __forceinline__ double dpp_rot(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over the 64 lanes of a wavefront (all lanes active), the same value in every lane: rotations
// by 8, 4, 2, 1 within each 16-lane row leave every lane of the row holding the row's sum (the same
// pairs added in every lane, so bitwise equal), then the four row sums are read from lanes 0, 16,
// 32 and 48.  No LDS traffic (a __shfl_xor butterfly is six ds_bpermute round trips).
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_rot<0x128>(v);  // row_ror:8
    v += dpp_rot<0x124>(v);  // row_ror:4
    v += dpp_rot<0x122>(v);  // row_ror:2
    v += dpp_rot<0x121>(v);  // row_ror:1
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
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

// Data that one workgroup writes and others read inside the tridiagonalisation (the published
// p_j and column j+1) goes through relaxed agent-scope atomics: such stores write through the
// XCD's L2 and such loads do not hit a stale L2 copy, so the barrier below needs no L2 writeback
// or invalidate (an agent-scope fence costs several microseconds per workgroup), only that every
// store has completed before the flag goes out.
__device__ __forceinline__ void xstore(double2* p, double2 v) {
    double* q = reinterpret_cast<double*>(p);
    __hip_atomic_store(q, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double2 xload(const double2* p) {
    const double* q = reinterpret_cast<const double*>(p);
    return make_double2(__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                        __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Barrier over the whole (co-resident) grid.  Each workgroup owns a flag on its own 64-byte line
// and stores the barrier's epoch there (no read-modify-write on a shared counter, whose atomics
// serialise: about 70 ns per arriving workgroup); wavefront 0 of every workgroup then polls all
// flags (64 per load instruction) until each has reached the epoch.  Every thread first waits for
// its own memory operations to complete (s_waitcnt 0), so the exchanged data written with xstore
// is in place before the flag is.  With part != nullptr, wavefront 0 then also sums part[0:G]
// (one complex value per workgroup, written with xstore before the barrier) into sum[0:2].
// A wait longer than spin_limit polls (or an error raised elsewhere) sets *err; every thread
// returns false then and the kernel winds down.
__device__ bool grid_sync(unsigned* flags, unsigned epoch, int* err, int* s_ok, unsigned spin_limit,
                          const double2* part = nullptr, double* sum = nullptr) {
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = (int)threadIdx.x;
        const int G = (int)gridDim.x;
        if (lane == 0)
            __hip_atomic_store(flags + (size_t)blockIdx.x * 16, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        int ok = 1;
        while (true) {
            int arrived = 1;
            for (int q = lane; q < G; q += 64)
                arrived &= __hip_atomic_load(flags + (size_t)q * 16, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= epoch;
            if (__all(arrived)) break;
            const int bad = lane == 0 ? __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            if (__any(bad) || ++spins > spin_limit) {
                if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (part && ok) {
            double sx = 0.0, sy = 0.0;
            for (int q = lane; q < G; q += 64) {
                const double2 v = xload(part + q);
                sx += v.x;
                sy += v.y;
            }
            sx = wave_sum(sx);
            sy = wave_sum(sy);
            if (lane == 0) {
                sum[0] = sx;
                sum[1] = sy;
            }
        }
        if (lane == 0) *s_ok = ok;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    __syncthreads();
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
    double2* pbuf; // 2 x m: tau_j A v_j, then 2 x G: per-workgroup parts of p_j^H v_j; by step parity
    double2* nbuf; // 2 x m: the next column after the previous rank-2 update
    unsigned* flags;  // one per workgroup, 16 apart
    int* err;
    unsigned spin_limit;  // barrier polls before giving up
};

// Householder tridiagonalisation A = Q T Q^H, Q = H_0 ... H_{m-2}, H_j = I - tau_j v_j v_j^H, with
// zhetd2's reflectors (zlarfg: H_j^H x = beta e_1, beta real).  Workgroup g holds the trailing
// columns i = g + c G (c < NCL) in LDS for the whole reduction, and every workgroup keeps its own
// copy of the current column.  Step j (thread t works on rows r = j+1+t (mod 256) throughout, so
// that its phases need no workgroup barrier between them):
//   A. every workgroup forms v_j, tau_j from the current column j (norm of rows j+2.. summed at
//      the end of the previous step);
//   B. fused pass over its own columns i > j: the rank-2 update of step j-1
//      a_i -= v_{j-1} conj(w_{j-1}[i]) + w_{j-1} conj(v_{j-1}[i]), then p_j[i] = tau_j a_i^H v_j
//      (A is Hermitian, so (A v)_i = a_i^H v); the owner of column j+1 publishes it (updated), and
//      every workgroup publishes its part of p_j^H v_j;
//   C. grid barrier, which also sums those parts;
//   D. w_j = p_j - (tau_j / 2) (p_j^H v_j) v_j, from the published p_j;
//   E. column j+1 = published column - v_j conj(w_j[j+1]) - w_j conj(v_j[j+1]), and the norm of
//      its rows j+3.. for the next step.
// The owner never updates column j+1 in place at step j, so the published copy stays valid while
// the others read it; the published buffers alternate by step parity.
template <int NCL>
__global__ __launch_bounds__(kTrdThreads) void k_hetrd(TrdArgs a) {
    extern __shared__ double2 lds[];
    __shared__ double red[(kTrdThreads / 64) * 2 * kTrdMaxCols];
    __shared__ double s_dot[2];
    __shared__ int s_ok;
    constexpr int ncl = NCL;
    constexpr int NT = kTrdThreads;
    const int m = a.m, G = (int)gridDim.x, g = (int)blockIdx.x, t = (int)threadIdx.x;
    const int wv = t >> 6, lane = t & 63;
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
    double xs = 0.0;  // this thread's part of |column j, rows j+2..|^2
    for (int r = 1 + t; r < m; r += NT)
        if (r >= 2) xs += vc_b[r].x * vc_b[r].x + vc_b[r].y * vc_b[r].y;
    for (int j = 0; j + 1 < m; ++j) {
        double2* vc = vc_b;
        // A. reflector of column j (rows j+1..m-1).  Every thread reads alpha between the two
        // barriers; thread 0 overwrites its row only after the second.
        xs = wave_sum(xs);
        __syncthreads();
        const double2 alpha = vc[j + 1];
        const double dj = vc[j].x;
        if (lane == 0) red[wv] = xs;
        __syncthreads();
        double xn2 = 0.0;
#pragma unroll
        for (int q = 0; q < NT / 64; ++q) xn2 += red[q];
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
            a.d[j] = dj;
            a.e[j] = beta;
            a.tau[j] = tau;
        }
        for (int r = j + 1 + t; r < m; r += NT) {
            const double2 v = r == j + 1 ? make_double2(1.0, 0.0) : cm(vc[r], scale);
            vc[r] = v;
            if (g == 0) a.vh[(size_t)j * m + r] = v;
        }
        // B. fused update (step j-1) and product (step j) over the own columns i >= j+1.
        const double2* vp = vp_b;
        const double2* wp = wp_b;
        double part[2 * ncl];
        double2 wi[ncl], vi[ncl];
#pragma unroll
        for (int c = 0; c < ncl; ++c) {
            const int i = g + c * G;
            const bool live = j > 0 && i > j && i < m;
            wi[c] = live ? wp[i] : make_double2(0.0, 0.0);
            vi[c] = live ? vp[i] : make_double2(0.0, 0.0);
            part[2 * c] = 0.0;
            part[2 * c + 1] = 0.0;
        }
        double2* nb_out = a.nbuf + (size_t)((j + 1) & 1) * m;
        for (int r = j + 1 + t; r < m; r += NT) {
            const double2 vr = vc[r];
            const double2 vpr = j > 0 ? vp[r] : make_double2(0.0, 0.0);
            const double2 wpr = j > 0 ? wp[r] : make_double2(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < ncl; ++c) {
                const int i = g + c * G;
                if (i <= j || i >= m) continue;
                double2* col = cols + (size_t)c * m;
                double2 x = col[r];
                if (j > 0) {
                    // x -= vp[r] conj(wi) + wp[r] conj(vi)
                    const double2 u1 = cjm(wi[c], vpr);
                    const double2 u2 = cjm(vi[c], wpr);
                    x.x -= u1.x + u2.x;
                    x.y -= u1.y + u2.y;
                    col[r] = x;
                }
                const double2 pr = cjm(x, vr);
                part[2 * c] += pr.x;
                part[2 * c + 1] += pr.y;
                if (i == j + 1) xstore(nb_out + r, x);
            }
        }
        double sums[2 * ncl];
        block_sums<NT>(part, red, sums);
        double2* pb = a.pbuf + (size_t)(j & 1) * m;
        double2* dpart = a.pbuf + (size_t)2 * m + (size_t)(j & 1) * G;
        double2 pdot = make_double2(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < ncl; ++c) {
            const int i = g + c * G;
            if (i > j && i < m) {
                const double2 p = cm(tau, make_double2(sums[2 * c], sums[2 * c + 1]));
                if (t == c) xstore(pb + i, p);
                const double2 u = cjm(p, vc[i]);
                pdot.x += u.x;
                pdot.y += u.y;
            }
        }
        if (t == 0) xstore(dpart + g, pdot);
        // C.
        if (!grid_sync(a.flags, (unsigned)(j + 1), a.err, &s_ok, a.spin_limit)) return;
        // Every published value this thread needs, loaded together: p_j and column j+1 on its own
        // rows, p_j[j+1], and (wavefront 0) the parts of p_j^H v_j.
        const double2* nb_in = a.nbuf + (size_t)((j + 1) & 1) * m;
        double2 pr_in[kTrdRowsPerThread], nr_in[kTrdRowsPerThread];
#pragma unroll
        for (int q = 0; q < kTrdRowsPerThread; ++q) {
            const int r = j + 1 + t + q * NT;
            if (r < m) {
                pr_in[q] = xload(pb + r);
                nr_in[q] = xload(nb_in + r);
            }
        }
        const double2 pj1 = xload(pb + j + 1);
        if (wv == 0) {
            double sx = 0.0, sy = 0.0;
            for (int q = lane; q < G; q += 64) {
                const double2 v = xload(dpart + q);
                sx += v.x;
                sy += v.y;
            }
            sx = wave_sum(sx);
            sy = wave_sum(sy);
            if (lane == 0) {
                s_dot[0] = sx;
                s_dot[1] = sy;
            }
        }
        __syncthreads();
        // D. w_j = p_j - (tau_j / 2) (p_j^H v_j) v_j on the own rows; row j+1 by every thread.
        double2* wn = wn_b;
        const double2 al = cm(make_double2(-0.5 * tau.x, -0.5 * tau.y), make_double2(s_dot[0], s_dot[1]));
#pragma unroll
        for (int q = 0; q < kTrdRowsPerThread; ++q) {
            const int r = j + 1 + t + q * NT;
            if (r < m) {
                const double2 u = cm(al, vc[r]);
                wn[r] = make_double2(pr_in[q].x + u.x, pr_in[q].y + u.y);
            }
        }
        const double2 wj1 = make_double2(pj1.x + al.x, pj1.y + al.y);  // v_j[j+1] = 1
        // E. column j+1 after the step-j update, into the buffer of v_{j-1} (no longer needed).
        double2* vnext = vp_b;
        xs = 0.0;
#pragma unroll
        for (int q = 0; q < kTrdRowsPerThread; ++q) {
            const int r = j + 1 + t + q * NT;
            if (r >= m) continue;
            const double2 x = nr_in[q];
            const double2 u1 = cjm(wj1, vc[r]);
            const double2 w = wn[r];  // conj(v_j[j+1]) w_j[r] with v_j[j+1] = 1
            const double2 y = make_double2(x.x - u1.x - w.x, x.y - u1.y - w.y);
            vnext[r] = y;
            if (r >= j + 3) xs += y.x * y.x + y.y * y.y;
        }
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

// Sturm counts: numbers of eigenvalues of T(d, e) below x[0..NP) (e2 = e^2), with tiny pivots
// replaced by -pivmin (LAPACK dlaebz's rule).  The NP recurrences are independent and interleaved.
template <int NP>
__device__ __forceinline__ void sturm_counts(const double* d, const double* e2, int m, const double (&x)[NP],
                                             double pivmin, int (&cnt)[NP]) {
    double q[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        q[p] = d[0] - x[p];
        if (fabs(q[p]) < pivmin) q[p] = -pivmin;
        cnt[p] = q[p] < 0.0 ? 1 : 0;
    }
    for (int k = 1; k < m; ++k) {
        const double dk = d[k], ek = e2[k - 1];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            q[p] = dk - x[p] - ek / q[p];
            if (fabs(q[p]) < pivmin) q[p] = -pivmin;
            cnt[p] += q[p] < 0.0 ? 1 : 0;
        }
    }
}

// Eigenvalue idx (ascending) of T by multisection: each round the 64 lanes of a wavefront count
// at NP points each (64 NP interior points of [lo, hi]) and keep the sub-interval that holds
// eigenvalue idx.  Also writes the Gershgorin scale of T to scal[0] (block 0).
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
    constexpr int NP = 1;  // more points per lane lose: the FP64 divisions bound the issue rate
    constexpr double kParts = 64.0 * NP + 1.0;
    double lo = gl, hi = gu;
    for (int round = 0; round < 14; ++round) {
        const double width = hi - lo;
        if (width <= 2.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi)) + 2.0 * pivmin) break;
        double x[NP];
        int cnt[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) x[p] = lo + width * (double)(lane * NP + p + 1) / kParts;
        sturm_counts<NP>(d, e2, m, x, pivmin, cnt);
        // Highest point (lane * NP + p) that still has eigenvalue idx at or above it; -1 if none.
        int L = -1;
#pragma unroll
        for (int p = 0; p < NP; ++p)
            if (cnt[p] <= idx) L = lane * NP + p;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) L = max(L, __shfl_xor(L, o, 64));
        double nlo = lo, nhi = hi;
        if (L < 0) {
            nhi = lo + width * 1.0 / kParts;
        } else {
            nlo = lo + width * (double)(L + 1) / kParts;
            if (L < 64 * NP - 1) nhi = lo + width * (double)(L + 2) / kParts;
        }
        lo = nlo;
        hi = nhi;
    }
    if (lane == 0) w[idx] = 0.5 * (lo + hi);
}

// Eigenvalues closer than this (consecutive gaps, relative to the Gershgorin scale) form a cluster
// whose inverse-iteration vectors are orthonormalised together.
__device__ __forceinline__ double cluster_tol(double bnorm) { return 1e-12 * fmax(bnorm, 1.0); }

// Eigenvector of T for eigenvalue k by inverse iteration, one wavefront per eigenvalue (all
// independent).  T - lambda I = P L U with partial pivoting (dlagtf's factorisation, U with two
// superdiagonals; pivots below eps |T| are raised to that size), then three solves from a seeded
// start, each followed by scaling and normalisation.  T, the factors and the iterate live in LDS;
// the recurrences run on lane 0, the rest on all lanes.  LDS: 7 m doubles + m bytes.
__global__ __launch_bounds__(64) void k_stein(const double* __restrict__ d_in, const double* __restrict__ e_in, int m,
                                              int nev, const double* __restrict__ w, const double* __restrict__ scal,
                                              double* __restrict__ Z, int ldz) {
    extern __shared__ double sh[];
    const int k = (int)blockIdx.x, lane = (int)threadIdx.x;
    if (k >= nev) return;
    double* d = sh;
    double* e = d + m;
    double* u0i = e + m;  // 1 / u0
    double* u1 = u0i + m;
    double* u2 = u1 + m;
    double* lm = u2 + m;
    double* x = lm + m;
    unsigned char* pv = reinterpret_cast<unsigned char*>(x + m);
    for (int r = lane; r < m; r += 64) {
        d[r] = d_in[r];
        e[r] = r + 1 < m ? e_in[r] : 0.0;
        // Seeded start (splitmix64 of eigenvalue and row).
        unsigned long long h = 0x9E3779B97F4A7C15ull * (unsigned long long)(k + 1) + (unsigned long long)r;
        h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
        h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
        h ^= h >> 31;
        x[r] = (double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    __syncthreads();
    const double bnorm = fmax(scal[0], 1e-300);
    const double pert = 2.220446049250313e-16 * bnorm + 1e-300;
    const double lam = w[k];
    if (lane == 0) {
        double ca = d[0] - lam, cb = m > 1 ? e[0] : 0.0;
        for (int r = 0; r + 1 < m; ++r) {
            const double c = e[r];
            const double na = d[r + 1] - lam;
            const double nb = e[r + 1];
            double u0;
            if (fabs(ca) >= fabs(c)) {
                const double l = ca != 0.0 ? c / ca : 0.0;
                u0 = ca;
                u1[r] = cb;
                u2[r] = 0.0;
                lm[r] = l;
                pv[r] = 0;
                ca = na - l * cb;
                cb = nb;
            } else {
                const double l = ca / c;
                u0 = c;
                u1[r] = na;
                u2[r] = nb;
                lm[r] = l;
                pv[r] = 1;
                ca = cb - l * na;
                cb = -l * nb;
            }
            if (fabs(u0) < pert) u0 = u0 < 0.0 ? -pert : pert;
            u0i[r] = 1.0 / u0;
        }
        if (fabs(ca) < pert) ca = ca < 0.0 ? -pert : pert;
        u0i[m - 1] = 1.0 / ca;
    }
    __syncthreads();
    for (int it = 0; it < 3; ++it) {
        if (lane == 0) {
            // P L y = x, then U z = y, in place.
            double xr = x[0];
            for (int r = 0; r + 1 < m; ++r) {
                double xn = x[r + 1];
                if (pv[r]) {
                    const double tmp = xr;
                    xr = xn;
                    xn = tmp;
                }
                x[r] = xr;
                xr = xn - lm[r] * xr;
            }
            x[m - 1] = xr;
            double z1 = xr * u0i[m - 1];
            x[m - 1] = z1;
            if (m > 1) {
                double z0 = (x[m - 2] - u1[m - 2] * z1) * u0i[m - 2];
                x[m - 2] = z0;
                for (int r = m - 3; r >= 0; --r) {
                    const double zr = (x[r] - u1[r] * z0 - u2[r] * z1) * u0i[r];
                    x[r] = zr;
                    z1 = z0;
                    z0 = zr;
                }
            }
        }
        __syncthreads();
        // Scale by the largest entry (the solve amplifies by up to 1 / pert), then normalise.
        double mx = 0.0;
        for (int r = lane; r < m; r += 64) mx = fmax(mx, fabs(x[r]));
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
        const double inv_mx = mx > 0.0 ? 1.0 / mx : 1.0;
        double s2 = 0.0;
        for (int r = lane; r < m; r += 64) {
            const double v = x[r] * inv_mx;
            s2 += v * v;
        }
        s2 = wave_sum(s2);
        const double sc = inv_mx / sqrt(fmax(s2, 1e-300));
        for (int r = lane; r < m; r += 64) x[r] *= sc;
        __syncthreads();
    }
    for (int r = lane; r < m; r += 64) Z[(size_t)k * ldz + r] = x[r];
}

// Members of each cluster (consecutive eigenvalues closer than cluster_tol), in order, made
// orthonormal against the cluster's earlier members by two passes of classical Gram-Schmidt (one
// workgroup per cluster; the first member leads).  Inverse iteration gives each member a vector in
// the cluster's invariant subspace, so the orthonormal result spans the same subspace.
__global__ __launch_bounds__(256) void k_stein_cluster(const double* __restrict__ w, const double* __restrict__ scal,
                                                       int nev, int m, double* __restrict__ Z, int ldz) {
    __shared__ double dots[kHeevMaxM];
    __shared__ double red[4];
    const int i = (int)blockIdx.x;
    if (i >= nev) return;
    const double tau_c = cluster_tol(scal[0]);
    if (i > 0 && w[i] - w[i - 1] <= tau_c) return;
    int last = i;
    while (last + 1 < nev && w[last + 1] - w[last] <= tau_c) ++last;
    if (last == i) return;
    const int t = (int)threadIdx.x, wv = t >> 6, lane = t & 63;
    for (int k = i + 1; k <= last; ++k) {
        double* x = Z + (size_t)k * ldz;
        for (int pass = 0; pass < 2; ++pass) {
            for (int p = i + wv; p < k; p += 4) {
                const double* zp = Z + (size_t)p * ldz;
                double sdot = 0.0;
                for (int r = lane; r < m; r += 64) sdot += zp[r] * x[r];
                sdot = wave_sum(sdot);
                if (lane == 0) dots[p - i] = sdot;
            }
            __syncthreads();
            for (int r = t; r < m; r += 256) {
                double acc = x[r];
                for (int p = i; p < k; ++p) acc -= dots[p - i] * Z[(size_t)p * ldz + r];
                x[r] = acc;
            }
            __syncthreads();
        }
        double s2 = 0.0;
        for (int r = t; r < m; r += 256) s2 += x[r] * x[r];
        s2 = wave_sum(s2);
        if (lane == 0) red[wv] = s2;
        __syncthreads();
        const double nrm = sqrt(red[0] + red[1] + red[2] + red[3]);
        const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
        for (int r = t; r < m; r += 256) x[r] *= inv;
        __syncthreads();
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

// Y[:, cols] = Q Z[:, cols] = H_0 (H_1 (... H_{m-2} Z)).  One workgroup per kUnmtrWgCols columns,
// each wavefront holding kUnmtrWaveCols of them in registers (lane l: rows l + 64 q).  The
// reflectors are staged through LDS in blocks of unmtr_block(NQ) (rows <= k zero, row k+1 one), and the
// next block's global loads are issued into registers before the current block is applied.
// NQ: rows per lane (m <= 64 NQ).  One eigenvector per wavefront (40 workgroups for 160 of them):
// the reflector sequence is latency-bound, and halving each wavefront's dot products, sums and
// updates per reflector beat two per wavefront by 3 % at m = 164..1000.
constexpr int kUnmtrWaveCols = 1;
constexpr int kUnmtrWgCols = 4 * kUnmtrWaveCols;
// Reflectors per LDS stage (fewer for the larger register sets).
constexpr int unmtr_block(int nq) { return nq <= 8 ? 8 : 4; }

template <int NQ>
__global__ __launch_bounds__(256) void k_unmtr(const double2* __restrict__ vh, const double2* __restrict__ tau, int m,
                                               int nev, const double* __restrict__ Z, int ldz, double2* __restrict__ Y,
                                               int ldy) {
    extern __shared__ double2 vs[];  // kUnmtrBlk x m
    __shared__ double2 ts[8];
    static_assert(unmtr_block(NQ) <= 8, "ts holds one block's tau values");
    const int t = (int)threadIdx.x, wv = t >> 6, lane = t & 63;
    constexpr int kUnmtrRegs = NQ;
    constexpr int kUnmtrBlk = unmtr_block(NQ);
    constexpr int kUnmtrPre = kUnmtrBlk * 64 * NQ / 256;  // staged values per thread
    const int col0 = (int)blockIdx.x * kUnmtrWgCols + wv * kUnmtrWaveCols;
    double2 y[kUnmtrWaveCols][kUnmtrRegs];
#pragma unroll
    for (int c = 0; c < kUnmtrWaveCols; ++c)
#pragma unroll
        for (int q = 0; q < kUnmtrRegs; ++q) {
            const int r = lane + 64 * q;
            y[c][q] = make_double2(r < m && col0 + c < nev ? Z[(size_t)(col0 + c) * ldz + r] : 0.0, 0.0);
        }
    // Block b holds reflectors klo(b) .. khi(b), applied from khi down; blocks from the bottom up.
    double2 pre[kUnmtrPre];
    auto fetch = [&](int khi) {
        const int klo = khi - kUnmtrBlk + 1 < 0 ? 0 : khi - kUnmtrBlk + 1;
        const int cnt = (khi - klo + 1) * m;
#pragma unroll
        for (int q = 0; q < kUnmtrPre; ++q) {
            const int idx = t + 256 * q;
            double2 v = make_double2(0.0, 0.0);
            if (idx < cnt) {
                const int sft = idx / m, r = idx - sft * m, k = klo + sft;
                if (r == k + 1) v = make_double2(1.0, 0.0);
                else if (r > k + 1) v = vh[(size_t)k * m + r];
            }
            pre[q] = v;
        }
    };
    if (m >= 2) fetch(m - 2);
    for (int khi = m - 2; khi >= 0; khi -= kUnmtrBlk) {
        const int klo = khi - kUnmtrBlk + 1 < 0 ? 0 : khi - kUnmtrBlk + 1;
        const int nk = khi - klo + 1;
        __syncthreads();  // the previous block is consumed
#pragma unroll
        for (int q = 0; q < kUnmtrPre; ++q) {
            const int idx = t + 256 * q;
            if (idx < nk * m) vs[idx] = pre[q];
        }
        if (t < nk) ts[t] = tau[klo + t];
        __syncthreads();
        if (klo > 0) fetch(klo - 1);
        for (int sft = nk - 1; sft >= 0; --sft) {
            const int k = klo + sft;
            const double2 tk = ts[sft];
            if (tk.x == 0.0 && tk.y == 0.0) continue;
            const double2* v = vs + (size_t)sft * m;
            double acc[2 * kUnmtrWaveCols];
#pragma unroll
            for (int c = 0; c < 2 * kUnmtrWaveCols; ++c) acc[c] = 0.0;
#pragma unroll
            for (int q = 0; q < kUnmtrRegs; ++q) {
                const int r = lane + 64 * q;
                if (r < m && r > k) {
                    const double2 vr = v[r];
#pragma unroll
                    for (int c = 0; c < kUnmtrWaveCols; ++c) {
                        const double2 p = cjm(vr, y[c][q]);
                        acc[2 * c] += p.x;
                        acc[2 * c + 1] += p.y;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 2 * kUnmtrWaveCols; ++c) acc[c] = wave_sum(acc[c]);
            double2 tsc[kUnmtrWaveCols];
#pragma unroll
            for (int c = 0; c < kUnmtrWaveCols; ++c) tsc[c] = cm(tk, make_double2(acc[2 * c], acc[2 * c + 1]));
#pragma unroll
            for (int q = 0; q < kUnmtrRegs; ++q) {
                const int r = lane + 64 * q;
                if (r < m && r > k) {
                    const double2 vr = v[r];
#pragma unroll
                    for (int c = 0; c < kUnmtrWaveCols; ++c) {
                        const double2 u = cm(vr, tsc[c]);
                        y[c][q].x -= u.x;
                        y[c][q].y -= u.y;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < kUnmtrWaveCols; ++c)
#pragma unroll
        for (int q = 0; q < kUnmtrRegs; ++q) {
            const int r = lane + 64 * q;
            if (r < m && col0 + c < nev) Y[(size_t)(col0 + c) * ldy + r] = y[c][q];
        }
}

size_t unmtr_lds_bytes(int m) { return (size_t)unmtr_block((m + 63) / 64) * m * sizeof(double2); }

const void* unmtr_kernel(int m) {
    if (m <= 256) return (const void*)k_unmtr<4>;
    if (m <= 512) return (const void*)k_unmtr<8>;
    if (m <= 768) return (const void*)k_unmtr<12>;
    return (const void*)k_unmtr<16>;
}

size_t trd_lds_bytes(int m, int ncl) { return (size_t)(ncl + 4) * (size_t)m * sizeof(double2); }
size_t stein_lds_bytes(int m) { return (size_t)7 * m * sizeof(double) + (size_t)m; }

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
        set_max_dynamic_lds((const void*)k_stein, stein_lds_bytes(kHeevMaxM));
        for (int mm : {256, 512, 768, kHeevMaxM}) set_max_dynamic_lds(unmtr_kernel(mm), unmtr_lds_bytes(kHeevMaxM));
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
    double *d, *e, *z, *zt, *gram;
    unsigned* sync;  // (G + 1) * 16 words: barrier flags, then the error flag
    int err_offset;
};

// The tridiagonalisation gate of each device (see TrdAdmission): CUs per XCD taken by admitted
// grids, out of cap.
constexpr int kXcds = 8;
struct TrdGate {
    std::mutex mu;
    std::condition_variable cv;
    int used = 0, cap = 0;
};
TrdGate g_trd_gate[64];

int trd_units(int G) { return (G + kXcds - 1) / kXcds; }

void trd_gate_acquire(int dev, int units) {
    TrdGate& gt = g_trd_gate[dev];
    std::unique_lock<std::mutex> lk(gt.mu);
    if (gt.cap == 0) {
        int cus = 0;
        DZ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        gt.cap = std::max(1, cus / kXcds);
    }
    DZ_REQUIRE(units >= 1 && units <= gt.cap, DZHIP_ERR_HIP, "Hermitian eigensolver: grid does not fit on the device");
    gt.cv.wait(lk, [&] { return gt.used + units <= gt.cap; });
    gt.used += units;
}

void trd_gate_release(int dev, int units) {
    TrdGate& gt = g_trd_gate[dev];
    {
        std::lock_guard<std::mutex> lk(gt.mu);
        gt.used -= units;
    }
    gt.cv.notify_all();
}

// Ordinary rather than cooperative launches of the gated tridiagonalisation; DZHIP_HEEV_COOP=1
// forces the cooperative launch everywhere (for comparisons).
bool trd_gated_enabled() {
    static const bool on = [] {
        const char* env = std::getenv("DZHIP_HEEV_COOP");
        return !(env && std::atoi(env) != 0);
    }();
    return on;
}

// Spin limit of admitted launches; DZHIP_HEEV_SHARED_SPIN overrides it (a test sets 1 so that the
// barrier gives up at once and the cooperative repeat runs).
unsigned shared_spin_limit() {
    static const unsigned v = [] {
        const char* env = std::getenv("DZHIP_HEEV_SHARED_SPIN");
        return env ? (unsigned)std::max(1, std::atoi(env)) : kSpinLimitShared;
    }();
    return v;
}

void run_tridiag(const HeevBufs& b, int G, int ncl, const double2* A, int lda, int m, hipStream_t s,
                 bool cooperative = true) {
    // Barrier flags (one 64-byte line per workgroup) and the error flag after them, cleared by a
    // kernel (not a memset node) on the same stream.
    const int nsync = (G + 1) * 16;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((nsync + 255) / 256)), dim3(256), 0, s, b.sync, nsync);
    DZ_HIP(hipGetLastError());
    TrdArgs args{A, lda, m, ncl, b.vh, b.tau, b.d, b.e, b.pbuf, b.nbuf, b.sync,
                 reinterpret_cast<int*>(b.sync) + b.err_offset, cooperative ? kSpinLimit : shared_spin_limit()};
    void* kargs[] = {&args};
    if (cooperative) {
        DZ_HIP(hipLaunchCooperativeKernel(trd_kernel(ncl), dim3((unsigned)G), dim3(kTrdThreads), kargs,
                                          (unsigned)trd_lds_bytes(m, ncl), s));
    } else {
        DZ_HIP(hipLaunchKernel(trd_kernel(ncl), dim3((unsigned)G), dim3(kTrdThreads), kargs,
                               (unsigned)trd_lds_bytes(m, ncl), s));
    }
}

void run_eig(const HeevBufs& b, int m, int nev, double* w_dev, double2* Y, int ldy, hipStream_t s) {
    double* scal = b.gram + (size_t)nev * nev;
    hipLaunchKernelGGL(k_stebz, dim3((unsigned)((nev + 3) / 4)), dim3(256), 2 * m * sizeof(double), s, b.d, b.e, m,
                       nev, w_dev, scal, reinterpret_cast<const int*>(b.sync) + b.err_offset);
    DZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_stein, dim3((unsigned)nev), dim3(64), stein_lds_bytes(m), s, b.d, b.e, m, nev, w_dev, scal,
                       b.z, m);
    DZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_stein_cluster, dim3((unsigned)nev), dim3(256), 0, s, w_dev, scal, nev, m, b.z, m);
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
    {
        const double2* vh = b.vh;
        const double2* tau = b.tau;
        const double* zcc = zc;
        int mm = m, nn = nev, ldz = m, ldyy = ldy;
        void* kargs[] = {&vh, &tau, &mm, &nn, &zcc, &ldz, &Y, &ldyy};
        DZ_HIP(hipLaunchKernel(unmtr_kernel(m), dim3((unsigned)((nev + kUnmtrWgCols - 1) / kUnmtrWgCols)), dim3(256),
                               kargs, unmtr_lds_bytes(m), s));
    }
    DZ_HIP(hipGetLastError());
}

HeevBufs work_bufs(HeevWork& w, int m, int nev, int G) {
    w.vh.ensure((size_t)m * m * sizeof(double2));
    w.tau.ensure((size_t)m * sizeof(double2));
    w.d.ensure((size_t)m * sizeof(double));
    w.e.ensure((size_t)m * sizeof(double));
    w.pbuf.ensure((size_t)2 * (m + G) * sizeof(double2));
    w.nbuf.ensure((size_t)2 * m * sizeof(double2));
    w.sync.ensure((size_t)(G + 1) * 16 * sizeof(unsigned));
    if (nev > 0) {
        w.z.ensure((size_t)m * nev * sizeof(double));
        w.zt.ensure((size_t)m * nev * sizeof(double));
        w.gram.ensure((size_t)nev * nev * sizeof(double) + 64);
    }
    w.err_offset = G * 16;
    return HeevBufs{w.vh.as<double2>(), w.tau.as<double2>(), w.pbuf.as<double2>(), w.nbuf.as<double2>(),
                    w.d.as<double>(),   w.e.as<double>(),     w.z.as<double>(),     w.zt.as<double>(),
                    w.gram.as<double>(), w.sync.as<unsigned>(), G * 16};
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
size_t scratch_layout(int m, int nev, int G, size_t out[10]) {
    size_t off = 0;
    out[0] = carve(off, (size_t)m * m * sizeof(double2));       // vh
    out[1] = carve(off, (size_t)m * sizeof(double2));           // tau
    out[2] = carve(off, (size_t)2 * (m + G) * sizeof(double2)); // pbuf
    out[3] = carve(off, (size_t)2 * m * sizeof(double2));       // nbuf
    out[4] = carve(off, (size_t)m * sizeof(double));            // d
    out[5] = carve(off, (size_t)m * sizeof(double));            // e
    out[6] = carve(off, (size_t)m * nev * sizeof(double));      // z
    out[7] = carve(off, (size_t)m * nev * sizeof(double));      // zt
    out[8] = carve(off, (size_t)nev * nev * sizeof(double) + 64);  // gram + scale
    out[9] = carve(off, (size_t)(G + 1) * 16 * sizeof(unsigned));  // sync
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
    size_t off[10];
    const size_t need = scratch_layout(m, nev, G, off);
    DZ_REQUIRE(need <= kHeevNoAllocBytes, DZHIP_ERR_ARG, "allocation-free Hermitian eigensolver: scratch too small");
    unsigned char* p = static_cast<unsigned char*>(base);
    const HeevBufs b{reinterpret_cast<double2*>(p + off[0]), reinterpret_cast<double2*>(p + off[1]),
                     reinterpret_cast<double2*>(p + off[2]), reinterpret_cast<double2*>(p + off[3]),
                     reinterpret_cast<double*>(p + off[4]),  reinterpret_cast<double*>(p + off[5]),
                     reinterpret_cast<double*>(p + off[6]),  reinterpret_cast<double*>(p + off[7]),
                     reinterpret_cast<double*>(p + off[8]),  reinterpret_cast<unsigned*>(p + off[9]), G * 16};
    run_tridiag(b, G, ncl, A, lda, m, s);
    run_eig(b, m, nev, w_dev, Y, ldy, s);
}

void HeevWork::release() {
    DevBuf* bufs[] = {&vh, &tau, &d, &e, &pbuf, &nbuf, &sync, &z, &zt, &gram};
    for (DevBuf* b : bufs) b->release();
}

void heev_tridiag_only(HeevWork& w, const double2* A, int lda, int m, hipStream_t s) {
    DZ_REQUIRE(m >= 1 && m <= kHeevMaxM && lda >= m, DZHIP_ERR_ARG, "Hermitian eigensolver: bad size");
    int G = 1, ncl = 1;
    trd_shape(m, &G, &ncl);
    run_tridiag(work_bufs(w, m, 0, G), G, ncl, A, lda, m, s);
}

TrdAdmission::~TrdAdmission() {
    if (units <= 0) return;
    (void)hipStreamSynchronize(stream);
    trd_gate_release(device, units);
}

void heev_lowest(HeevWork& w, const double2* A, int lda, int m, int nev, double* w_dev, double2* Y, int ldy,
                 hipStream_t s, TrdAdmission* adm) {
    DZ_REQUIRE(m >= 1 && m <= kHeevMaxM && nev >= 1 && nev <= m && lda >= m && ldy >= m, DZHIP_ERR_ARG,
               "Hermitian eigensolver: need 1 <= nev <= m <= 1024 and leading dimensions >= m");
    int G = 1, ncl = 1;
    trd_shape(m, &G, &ncl);
    const HeevBufs b = work_bufs(w, m, nev, G);
    if (adm && trd_gated_enabled()) {
        DZ_REQUIRE(adm->units == 0, DZHIP_ERR_STATE, "Hermitian eigensolver: admission already held");
        int dev = 0;
        DZ_HIP(hipGetDevice(&dev));
        DZ_REQUIRE(dev >= 0 && dev < 64, DZHIP_ERR_ARG, "device index out of range");
        const int units = trd_units(G);
        trd_gate_acquire(dev, units);
        adm->device = dev;
        adm->units = units;
        adm->stream = s;
        run_tridiag(b, G, ncl, A, lda, m, s, false);
    } else {
        run_tridiag(b, G, ncl, A, lda, m, s);
    }
    run_eig(b, m, nev, w_dev, Y, ldy, s);
}

std::atomic<unsigned long long> g_trd_retries{0};

void heev_lowest_shared(HeevWork& w, const double2* A, int lda, int m, int nev, double* w_dev, double2* Y, int ldy,
                        hipStream_t s) {
    {
        TrdAdmission adm;
        heev_lowest(w, A, lda, m, nev, w_dev, Y, ldy, s, &adm);
        if (heev_error_flag(w, s) == 0) return;  // synchronises s
    }
    // The admitted grid gave up waiting for its workgroups (A is unchanged): cooperative launch.
    g_trd_retries.fetch_add(1, std::memory_order_relaxed);
    heev_lowest(w, A, lda, m, nev, w_dev, Y, ldy, s);
    DZ_REQUIRE(heev_error_flag(w, s) == 0, DZHIP_ERR_HIP,
               "Hermitian eigensolver: tridiagonalisation barrier did not complete");
}

unsigned long long heev_shared_retries() { return g_trd_retries.load(std::memory_order_relaxed); }

int heev_error_flag(HeevWork& w, hipStream_t s) {
    int flag = 0;
    if (!w.sync.ptr) return 0;
    DZ_HIP(hipMemcpyAsync(&flag, w.sync.as<int>() + w.err_offset, sizeof(int), hipMemcpyDeviceToHost, s));
    DZ_HIP(hipStreamSynchronize(s));
    return flag;
}

}  // namespace dzhip
