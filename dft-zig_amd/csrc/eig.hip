// Device-resident block eigensolver for H x = lambda x (S = I) on one k-point.
//
// Follows the reference's block-Davidson ("LOBPCG" there,
// src/features/linalg/lobpcg/serial.zig:55-307) in its outer semantics: subspace grows from
// block = nbands + 4 by appending preconditioned residuals of the unconverged bands, up to
// max_subspace = min(n, 4 nbands + 8); converged when max_b ||H x_b - lambda_b x_b||_2 < tol
// or when no vector can be added; kinetic-shift preconditioner r / (diag - lambda) with
// |denominator| clamped at 1e-8 (common.zig:214-226); random start from the reference LCG
// (common.zig:287-319, seed 0x6f3d3cbe).  The block operations are re-designed for the GPU:
//   * new vectors are orthogonalised against the basis with two passes of block classical
//     Gram-Schmidt (GEMMs) and orthonormalised with Cholesky QR (in place of one-vector MGS);
//   * the projected matrix T = V^H H V is extended by the new columns only;
//   * Rayleigh-Ritz takes the lowest nbands eigenpairs of T with the library's own device
//     Hermitian eigensolver (heev.hip: one-launch tridiagonalisation, multisection, inverse
//     iteration), rocSOLVER zheevd only for subspaces above 1024;
//   * a full subspace restarts from [Ritz vectors, new block], reusing H x = W c (no H applies
//     are discarded).
// All subspace algebra is FP64.  In mixed mode the H apply runs in FP32 until the largest
// residual falls below `mixed_switch_tol`, then the basis restarts from the Ritz vectors with an
// FP64 apply, so the final eigenpairs are those of the FP64 operator.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "ham.h"
#include "heev.h"
#include "herm3.h"

namespace dzhip {

#define DZ_BLAS(call)                                                                           \
    do {                                                                                        \
        rocblas_status st_ = (call);                                                            \
        if (st_ != rocblas_status_success)                                                      \
            throw Error{DZHIP_ERR_HIP, std::string(#call) + ": " + rocblas_status_to_string(st_)}; \
    } while (0)

using zc = rocblas_double_complex;

// ----------------------------------------------------------------- kernels
__global__ void k_col_norm2(const double2* __restrict__ v, size_t ld, int n, double* __restrict__ out) {
    __shared__ double red[256];
    const double2* c = v + (size_t)blockIdx.x * ld;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += c[i].x * c[i].x + c[i].y * c[i].y;
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// R[:, b] = HX[:, b] - lambda_b X[:, b]
__global__ void k_residual(const double2* __restrict__ hx, const double2* __restrict__ x,
                           const double* __restrict__ lam, double2* __restrict__ r, size_t ld, int n) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t o = (size_t)b * ld + i;
    const double l = lam[b];
    r[o] = make_double2(hx[o].x - l * x[o].x, hx[o].y - l * x[o].y);
}

// Q[:, j] = P(R[:, sel_j]): kinetic shift r / (diag - lambda) with the reference's 1e-8 clamp,
// or Teter-Payne-Allan K(x) r with x = diag / ekin_b (ekin = <x_b|T|x_b>).
__global__ void k_precondition(const double2* __restrict__ r, const int* __restrict__ sel,
                               const double* __restrict__ lam, const double* __restrict__ ekin,
                               const double* __restrict__ diag, double2* __restrict__ q, size_t ld,
                               int n, int mode) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = sel[j];
    const double2 v = r[(size_t)b * ld + i];
    const double f = mode == 0 ? precond_kinetic_shift(diag[i], lam[b]) : precond_tpa(diag[i], ekin[b]);
    q[(size_t)j * ld + i] = make_double2(v.x * f, v.y * f);
}

// ----------------------------------------------------------------- Gram products C = A^H B
// For the tall-skinny products of the solver (K = n_pw ~ 24k, M, N <= a few hundred), where the
// library GEMM runs at a few TF/s: split-K over workgroups of a 64 x 64 output tile, FP64 MFMA
// 16x16x4 with the complex product as four real MFMAs, operands staged through LDS with a
// register prefetch of the next K-stage, partials reduced in a fixed order by a second kernel.
typedef double v4d_t __attribute__((ext_vector_type(4)));
constexpr int kGramTile = 64, kGramK = 16, kGramPitch = kGramK + 1, kGramThreads = 256;
constexpr int kGramPer = kGramTile * kGramK / kGramThreads;  // complex elements per thread per operand

__global__ __launch_bounds__(kGramThreads) void k_gram_cn(const double2* __restrict__ A, int lda, int M,
                                                          const double2* __restrict__ B, int ldb, int N, int K,
                                                          int kchunk, int mtiles, double2* __restrict__ partial) {
    __shared__ double2 As[kGramTile * kGramPitch];
    __shared__ double2 Bs[kGramTile * kGramPitch];
    const int mt = blockIdx.x % mtiles, ks = blockIdx.x / mtiles;
    const int i_base = mt * kGramTile, j_base = blockIdx.y * kGramTile;
    const int k0 = ks * kchunk, k1 = min(K, k0 + kchunk);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave & 1, wj = wave >> 1;
    const int l16 = lane & 15, l4 = lane >> 4;
    double ar[kGramPer], ai[kGramPer], br[kGramPer], bi[kGramPer];
    auto fetch = [&](int kt) {
#pragma unroll
        for (int q = 0; q < kGramPer; ++q) {
            const int e = tid + q * kGramThreads;
            const int c = e / kGramK, kk = e % kGramK, k = kt + kk;
            const int i = i_base + c, j = j_base + c;
            const bool oka = i < M && k < k1, okb = j < N && k < k1;
            const double2 va = A[oka ? (size_t)i * lda + k : 0];
            const double2 vb = B[okb ? (size_t)j * ldb + k : 0];
            ar[q] = oka ? va.x : 0.0;
            ai[q] = oka ? va.y : 0.0;
            br[q] = okb ? vb.x : 0.0;
            bi[q] = okb ? vb.y : 0.0;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < kGramPer; ++q) {
            const int e = tid + q * kGramThreads;
            const int c = e / kGramK, kk = e % kGramK;
            As[c * kGramPitch + kk] = make_double2(ar[q], ai[q]);
            Bs[c * kGramPitch + kk] = make_double2(br[q], bi[q]);
        }
    };
    // conj(a) b = (ar br + ai bi) + i (ar bi - ai br) by three real MFMAs (Gauss):
    // p1 = sum ar br, p2 = sum ai bi, p3 = sum (ar - ai)(br + bi); Re = p1 + p2, Im = p3 - p1 + p2.
    v4d_t p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            p1[x][y] = v4d_t{0, 0, 0, 0};
            p2[x][y] = v4d_t{0, 0, 0, 0};
            p3[x][y] = v4d_t{0, 0, 0, 0};
        }
    if (k0 < k1) {
        fetch(k0);
        stash();
    }
    __syncthreads();
    for (int kt = k0; kt < k1; kt += kGramK) {
        const bool more = kt + kGramK < k1;
        if (more) fetch(kt + kGramK);
#pragma unroll
        for (int kk = 0; kk < kGramK; kk += 4) {
            double2 av[2], bv[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) av[x] = As[(wi * 32 + x * 16 + l16) * kGramPitch + kk + l4];
#pragma unroll
            for (int y = 0; y < 2; ++y) bv[y] = Bs[(wj * 32 + y * 16 + l16) * kGramPitch + kk + l4];
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    p1[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x].x, bv[y].x, p1[x][y], 0, 0, 0);
                    p2[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x].y, bv[y].y, p2[x][y], 0, 0, 0);
                    p3[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x].x - av[x].y, bv[y].x + bv[y].y, p3[x][y],
                                                                   0, 0, 0);
                }
        }
        __syncthreads();
        if (more) stash();
        __syncthreads();
    }
    // D slot r of a 16x16 block: row (l/16) + 4 r, column l % 16.
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i_base + wi * 32 + x * 16 + l4 + 4 * r;
                const int j = j_base + wj * 32 + y * 16 + l16;
                const double a1 = p1[x][y][r], a2 = p2[x][y][r];
                if (i < M && j < N) partial[((size_t)ks * N + j) * M + i] = make_double2(a1 + a2, p3[x][y][r] - a1 + a2);
            }
}

// C[:, 0:N] = alpha A[:, 0:K] B[0:K, 0:N] + beta C, alpha and beta real: A is n x K (leading
// dimension lda), B is K x N (ldb), C is n x N (ldc), all column-major.  For the solver's products
// of the tall basis with small coefficient blocks (n = n_pw ~ 24k rows; K, N up to a few hundred):
// one workgroup per 64 x 64 tile of C, K in stages of 16 through LDS with a register prefetch of
// the next stage, FP64 MFMA 16x16x4 with each complex product as three real MFMAs.  The MFMA
// computes C^T = B^T A^T so that a wavefront's accumulator slots hold 16 consecutive rows of a
// column of C (coalesced loads of the old C and stores).
__global__ __launch_bounds__(kGramThreads) void k_gemm_nn(const double2* __restrict__ A, int lda,
                                                          const double2* __restrict__ B, int ldb,
                                                          double2* __restrict__ C, int ldc, int n, int K, int N,
                                                          double alpha, double beta) {
    __shared__ double2 As[kGramTile * kGramPitch];  // [row][k]
    __shared__ double2 Bs[kGramTile * kGramPitch];  // [column][k]
    const int r_base = blockIdx.x * kGramTile, c_base = blockIdx.y * kGramTile;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wr = wave & 1, wc = wave >> 1;
    const int l16 = lane & 15, l4 = lane >> 4;
    double2 av[kGramPer], bv[kGramPer];
    auto fetch = [&](int kt) {
#pragma unroll
        for (int q = 0; q < kGramPer; ++q) {
            const int e = tid + q * kGramThreads;
            // A: consecutive threads read consecutive rows of one column (coalesced).
            const int ra = e % kGramTile, ka = e / kGramTile;
            const int row = r_base + ra, k = kt + ka;
            av[q] = (row < n && k < K) ? A[(size_t)k * lda + row] : make_double2(0.0, 0.0);
            // B: consecutive threads read consecutive k of one column.
            const int kb = e % kGramK, cb = e / kGramK;
            const int col = c_base + cb, k2 = kt + kb;
            bv[q] = (col < N && k2 < K) ? B[(size_t)col * ldb + k2] : make_double2(0.0, 0.0);
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < kGramPer; ++q) {
            const int e = tid + q * kGramThreads;
            As[(e % kGramTile) * kGramPitch + e / kGramTile] = av[q];
            Bs[(e / kGramK) * kGramPitch + e % kGramK] = bv[q];
        }
    };
    // b a = (br ar - bi ai) + i (br ai + bi ar): p1 = sum br ar, p2 = sum bi ai,
    // p3 = sum (br + bi)(ar + ai); Re = p1 - p2, Im = p3 - p1 - p2.
    v4d_t p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            p1[x][y] = v4d_t{0, 0, 0, 0};
            p2[x][y] = v4d_t{0, 0, 0, 0};
            p3[x][y] = v4d_t{0, 0, 0, 0};
        }
    fetch(0);
    stash();
    __syncthreads();
    for (int kt = 0; kt < K; kt += kGramK) {
        const bool more = kt + kGramK < K;
        if (more) fetch(kt + kGramK);
#pragma unroll
        for (int kk = 0; kk < kGramK; kk += 4) {
            double2 bo[2], ao[2];
            // Operand 0 (MFMA rows): columns of C; operand 1 (MFMA columns): rows of C.
#pragma unroll
            for (int x = 0; x < 2; ++x) bo[x] = Bs[(wc * 32 + x * 16 + l16) * kGramPitch + kk + l4];
#pragma unroll
            for (int y = 0; y < 2; ++y) ao[y] = As[(wr * 32 + y * 16 + l16) * kGramPitch + kk + l4];
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    p1[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(bo[x].x, ao[y].x, p1[x][y], 0, 0, 0);
                    p2[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(bo[x].y, ao[y].y, p2[x][y], 0, 0, 0);
                    p3[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(bo[x].x + bo[x].y, ao[y].x + ao[y].y, p3[x][y],
                                                                   0, 0, 0);
                }
        }
        __syncthreads();
        if (more) stash();
        __syncthreads();
    }
    // Slot r of a 16x16 block: MFMA row (l/16) + 4 r = column of C, MFMA column l % 16 = row of C.
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = c_base + wc * 32 + x * 16 + l4 + 4 * r;
                const int row = r_base + wr * 32 + y * 16 + l16;
                if (col < N && row < n) {
                    const double a1 = p1[x][y][r], a2 = p2[x][y][r];
                    double2 v = make_double2(alpha * (a1 - a2), alpha * (p3[x][y][r] - a1 - a2));
                    double2* cp = C + (size_t)col * ldc + row;
                    if (beta != 0.0) {
                        const double2 o = *cp;
                        v.x += beta * o.x;
                        v.y += beta * o.y;
                    }
                    *cp = v;
                }
            }
}

// C[j * ldc + i] = sum_ks partial[ks][j][i]
__global__ void k_gram_reduce(const double2* __restrict__ partial, int nks, int M, int N, double2* __restrict__ C,
                              int ldc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= M) return;
    double re = 0.0, im = 0.0;
    for (int ks = 0; ks < nks; ++ks) {
        const double2 v = partial[((size_t)ks * N + j) * M + i];
        re += v.x;
        im += v.y;
    }
    C[(size_t)j * ldc + i] = make_double2(re, im);
}

// p = r / (diag - lambda) with the 1e-8 clamp (single vector, CG).
__global__ void k_precondition_one(const double2* __restrict__ r, const double* __restrict__ diag, double lam,
                                   double2* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double f = precond_kinetic_shift(diag[i], lam);
    p[i] = make_double2(r[i].x * f, r[i].y * f);
}

// out = c0 a + c1 b (+ c2 c), complex coefficients (CG Ritz update of psi and H psi).
__global__ void k_combine3(const double2* __restrict__ a, const double2* __restrict__ b,
                           const double2* __restrict__ c, double2 c0, double2 c1, double2 c2, int use_c,
                           double2* __restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 x = a[i], y = b[i];
    double re = c0.x * x.x - c0.y * x.y + c1.x * y.x - c1.y * y.y;
    double im = c0.x * x.y + c0.y * x.x + c1.x * y.y + c1.y * y.x;
    if (use_c) {
        const double2 z = c[i];
        re += c2.x * z.x - c2.y * z.y;
        im += c2.x * z.y + c2.y * z.x;
    }
    out[i] = make_double2(re, im);
}

// <x_b|T|x_b> = sum_G |k+G|^2 |x_bG|^2 (one workgroup per band)
__global__ void k_kinetic_energy(const double2* __restrict__ x, const double* __restrict__ diag, size_t ld,
                                 int n, double* __restrict__ out) {
    __shared__ double red[256];
    const double2* c = x + (size_t)blockIdx.x * ld;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += diag[i] * (c[i].x * c[i].x + c[i].y * c[i].y);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// dst[:, j] = src[:, idx_j] * scale_j
__global__ void k_gather_cols(const double2* __restrict__ src, const int* __restrict__ idx,
                              const double* __restrict__ scale, double2* __restrict__ dst, size_t ld, int n) {
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = scale[j];
    const double2 v = src[(size_t)idx[j] * ld + i];
    dst[(size_t)j * ld + i] = make_double2(v.x * s, v.y * s);
}

// T[m0 + i, j] = conj(T[j, m0 + i]) for j < m0 (mirror of the new columns), and the new
// a x a diagonal block made Hermitian with a real diagonal (common.zig:247-258).
__global__ void k_mirror_new(double2* __restrict__ t, int ld, int m0, int a) {
    const int i = blockIdx.y;  // new column index 0..a-1
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m0 + a) return;
    const int c = m0 + i;
    if (j < m0) {
        const double2 v = t[(size_t)c * ld + j];
        t[(size_t)j * ld + c] = make_double2(v.x, -v.y);
    } else {
        const int jj = j - m0;
        if (jj > i) return;
        const double2 u = t[(size_t)c * ld + j];  // T[j][c]
        const double2 w = t[(size_t)j * ld + c];  // T[c][j]
        if (jj == i) {
            t[(size_t)c * ld + j] = make_double2(u.x, 0.0);
        } else {
            const double re = 0.5 * (u.x + w.x), im = 0.5 * (u.y - w.y);
            t[(size_t)c * ld + j] = make_double2(re, im);
            t[(size_t)j * ld + c] = make_double2(re, -im);
        }
    }
}

// ----------------------------------------------------------------- solver
struct EigWork {
    rocblas_handle blas = nullptr;
    hipStream_t stream = nullptr;
    size_t n = 0;
    int max_sub = 0, nb = 0;
    DevBuf V, W, X, HX, R, Q, HQ, T, Tw, S, G, evals, ework, info, norms, lam, ekin, diag, sel, scl, idx;
    DevBuf x32, y32;
    uint64_t x_cols = 0;  // columns of X holding the last solve's Ritz vectors (0: none)
    bool x_converged = false;
    // Learned preconditioner: residual norms of the current iteration; training snapshots
    // (Ritz vectors, residuals, Ritz values, residual norms, <T> per band) of the last solve.
    DevBuf rnorm, snapX, snapR, snapLam, snapRn, snapEk, bandsum, samples, gram_part;
    int snaps = 0, snap_nb = 0;
    // Section timing: events recorded on the stream at section starts (pooled).
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_tag;
    size_t ev_used = 0;
    HeevWork heev;  // dense Hermitian eigensolver workspace (subspace problems)
    ~EigWork() {
        heev.release();
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        if (blas) rocblas_destroy_handle(blas);
        DevBuf* bufs[] = {&V, &W, &X, &HX, &R, &Q, &HQ, &T, &Tw, &S, &G, &evals, &ework, &info,
                          &norms, &lam, &ekin, &diag, &sel, &scl, &idx, &x32, &y32, &rnorm,
                          &snapX, &snapR, &snapLam, &snapRn, &snapEk, &bandsum, &samples, &gram_part};
        for (DevBuf* b : bufs) b->release();
    }
};

// Sections of a solve timed with events: a mark starts a section and ends the previous one.
enum EigSection { SEC_RR = 0, SEC_RITZ, SEC_PRECOND, SEC_ORTHO, SEC_APPLY, SEC_EXTEND, SEC_COUNT, SEC_END = SEC_COUNT };

static void mark(EigWork& w, hipStream_t s, int tag) {
    if (w.ev_used == w.ev_pool.size()) {
        hipEvent_t e;
        DZ_HIP(hipEventCreate(&e));
        w.ev_pool.push_back(e);
        w.ev_tag.push_back(0);
    }
    DZ_HIP(hipEventRecord(w.ev_pool[w.ev_used], s));
    w.ev_tag[w.ev_used++] = tag;
}

static void section_times(EigWork& w, double* out) {
    for (int i = 0; i < SEC_COUNT; ++i) out[i] = 0.0;
    if (w.ev_used < 2) {
        w.ev_used = 0;
        return;
    }
    DZ_HIP(hipEventSynchronize(w.ev_pool[w.ev_used - 1]));
    for (size_t i = 0; i + 1 < w.ev_used; ++i) {
        float ms = 0.f;
        DZ_HIP(hipEventElapsedTime(&ms, w.ev_pool[i], w.ev_pool[i + 1]));
        if (w.ev_tag[i] < SEC_COUNT) out[w.ev_tag[i]] += ms;
    }
    w.ev_used = 0;
}

// Lowest nev eigenpairs of the Hermitian m x m matrix A (lower triangle, leading dimension lda)
// into Y (leading dimension ldy) and the host vector returned (ascending).  The library's device
// eigensolver up to m = kHeevMaxM; rocSOLVER zheevd beyond (it overwrites a copy of A in Y,
// which must then hold m columns).
// Upper Cholesky factorisation A = R^H R of a Hermitian k x k matrix (column-major, leading
// dimension lda; the upper triangle is read and overwritten with R, the lower one is left alone),
// k <= kPotrfMax, one workgroup of 256 threads, right-looking in blocks of kPotrfNb rows:
//   1. the diagonal block by wavefront 0, lane l holding column l of it in registers; the row
//      R[j, :] of each step reaches the other lanes by lane shuffles;
//   2. the block's rows of R right of it, R_d^{-H} A[block, rest] by forward substitution, one
//      thread per column (the rows also go to LDS);
//   3. the trailing upper triangle minus the Gram products of those rows, 4 x 4 tiles per thread.
// *info = 0, or j + 1 for the first pivot j that is not positive (the factorisation stops there),
// as LAPACK's zpotrf reports it.
constexpr int kPotrfMax = 256;
constexpr int kPotrfNb = 16;
__global__ __launch_bounds__(256) void k_potrf_upper(double2* __restrict__ A, int lda, int k, int* __restrict__ info) {
    __shared__ double2 dblk[kPotrfNb][kPotrfNb];              // R of the diagonal block [row][column]
    __shared__ double2 rows[kPotrfNb][kPotrfMax - kPotrfNb];  // its rows right of it [row][column - c0]
    __shared__ int s_fail;
    const int tid = (int)threadIdx.x;
    if (tid == 0) {
        s_fail = 0;
        *info = 0;
    }
    __syncthreads();
    for (int jb = 0; jb < k; jb += kPotrfNb) {
        const int nb = min(kPotrfNb, k - jb);
        if (tid < 64) {
            const int l = tid;
            double2 col[kPotrfNb];  // A[jb + i][jb + l], i <= l
#pragma unroll
            for (int i = 0; i < kPotrfNb; ++i)
                col[i] = (l < nb && i <= l) ? A[(size_t)(jb + l) * lda + jb + i] : double2{0.0, 0.0};
            int fail = 0;
#pragma unroll
            for (int j = 0; j < kPotrfNb; ++j) {
                if (j < nb && fail == 0) {  // uniform over the wavefront
                    const double d = __shfl(col[j].x, j);
                    if (!(d > 0.0)) {
                        fail = jb + j + 1;
                    } else {
                        const double r = sqrt(d), inv = 1.0 / r;
                        if (l == j) col[j] = double2{r, 0.0};
                        if (l > j) col[j] = double2{col[j].x * inv, col[j].y * inv};  // R[j][l]
                        // A[i][l] -= conj(R[j][i]) R[j][l] for j < i <= l
#pragma unroll
                        for (int i = j + 1; i < kPotrfNb; ++i) {
                            const double rx = __shfl(col[j].x, i), ry = __shfl(col[j].y, i);  // R[j][i]
                            if (l >= i && l < nb) {
                                col[i].x -= rx * col[j].x + ry * col[j].y;
                                col[i].y -= rx * col[j].y - ry * col[j].x;
                            }
                        }
                    }
                }
            }
            if (l < nb) {
#pragma unroll
                for (int i = 0; i < kPotrfNb; ++i) {
                    if (i <= l) {
                        A[(size_t)(jb + l) * lda + jb + i] = col[i];
                        dblk[i][l] = col[i];
                    }
                }
            }
            if (l == 0 && fail) {
                s_fail = fail;
                *info = fail;
            }
        }
        __syncthreads();
        if (s_fail) return;
        const int c0 = jb + nb, m = k - c0;
        if (m <= 0) break;
        // Rows of R right of the block: solve R_d^H y = A[jb:jb+nb, c] for each column c.
        if (tid < m) {
            const int c = c0 + tid;
            double2 y[kPotrfNb];
#pragma unroll
            for (int i = 0; i < kPotrfNb; ++i) {
                if (i < nb) {
                    double2 v = A[(size_t)c * lda + jb + i];
#pragma unroll
                    for (int t = 0; t < i; ++t) {  // v -= conj(R_d[t][i]) y[t]
                        const double2 r = dblk[t][i];
                        v.x -= r.x * y[t].x + r.y * y[t].y;
                        v.y -= r.x * y[t].y - r.y * y[t].x;
                    }
                    const double inv = 1.0 / dblk[i][i].x;
                    y[i] = double2{v.x * inv, v.y * inv};
                    A[(size_t)c * lda + jb + i] = y[i];
                    rows[i][tid] = y[i];
                }
            }
        }
        __syncthreads();
        // Trailing upper triangle: A[i][c] -= sum_t conj(R[t][i]) R[t][c], c0 <= i <= c < k, in
        // 4 x 4 tiles (ti <= tc) per thread: eight LDS reads per t feed 16 complex products.
        const int nt = (m + 3) >> 2, ntiles = nt * (nt + 1) / 2;
        for (int e = tid; e < ntiles; e += 256) {
            int tc = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);  // tile column: tc (tc + 1) / 2 <= e
            while (tc * (tc + 1) / 2 > e) --tc;
            while ((tc + 1) * (tc + 2) / 2 <= e) ++tc;
            const int ti = e - tc * (tc + 1) / 2;
            double2 acc[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = double2{0.0, 0.0};
            for (int t = 0; t < nb; ++t) {
                double2 ri[4], rc[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int i = 4 * ti + a, c = 4 * tc + a;
                    ri[a] = i < m ? rows[t][i] : double2{0.0, 0.0};
                    rc[a] = c < m ? rows[t][c] : double2{0.0, 0.0};
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        acc[a][b].x += ri[a].x * rc[b].x + ri[a].y * rc[b].y;
                        acc[a][b].y += ri[a].x * rc[b].y - ri[a].y * rc[b].x;
                    }
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int c = 4 * tc + b;
                if (c >= m) continue;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int i = 4 * ti + a;
                    if (i > c) continue;
                    double2* p = A + (size_t)(c0 + c) * lda + c0 + i;
                    const double2 v = *p;
                    *p = double2{v.x - acc[a][b].x, v.y - acc[a][b].y};
                }
            }
        }
        __syncthreads();
    }
}

// ok = 1 when the Cholesky factorisation of the Gram matrix succeeded (info == 0) and every pivot
// keeps at least 1e-7 of its column's norm (R_jj > 1e-7 sqrt(G_jj); G's diagonal from its copy
// gcopy, R in r, both k x k with leading dimension k).  One workgroup.
__global__ __launch_bounds__(256) void k_chol_check(const double2* __restrict__ gcopy, const double2* __restrict__ r,
                                                    int k, const rocblas_int* __restrict__ info, int* __restrict__ ok) {
    int bad = 0;
    for (int j = (int)threadIdx.x; j < k; j += 256) {
        const double g = gcopy[(size_t)j * (k + 1)].x, rj = r[(size_t)j * (k + 1)].x;
        if (!(rj > 1e-7 * sqrt(fmax(g, 0.0)))) bad = 1;
    }
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) *ok = (*info == 0 && !bad) ? 1 : 0;
}

static std::vector<double> subspace_eigh(EigWork& w, const double2* A, int lda, int m, int nev, double2* Y, int ldy,
                                         hipStream_t s) {
    std::vector<double> ev(nev);
    if (m <= kHeevMaxM) {
        heev_lowest_shared(w.heev, A, lda, m, nev, w.evals.as<double>(), Y, ldy, s);
        DZ_HIP(hipMemcpyAsync(ev.data(), w.evals.ptr, nev * sizeof(double), hipMemcpyDeviceToHost, s));
        DZ_HIP(hipStreamSynchronize(s));
        return ev;
    }
    DZ_HIP(hipMemcpy2DAsync(Y, (size_t)ldy * sizeof(double2), A, (size_t)lda * sizeof(double2),
                            (size_t)m * sizeof(double2), m, hipMemcpyDeviceToDevice, s));
    DZ_BLAS(rocsolver_zheevd(w.blas, rocblas_evect_original, rocblas_fill_lower, m, reinterpret_cast<zc*>(Y), ldy,
                             w.evals.as<double>(), w.ework.as<double>(), w.info.as<rocblas_int>()));
    rocblas_int info = 0;
    DZ_HIP(hipMemcpyAsync(ev.data(), w.evals.ptr, nev * sizeof(double), hipMemcpyDeviceToHost, s));
    DZ_HIP(hipMemcpyAsync(&info, w.info.ptr, sizeof(info), hipMemcpyDeviceToHost, s));
    DZ_HIP(hipStreamSynchronize(s));
    DZ_REQUIRE(info == 0, DZHIP_ERR_HIP, "zheevd did not converge");
    return ev;
}

static void eig_work_release(Ham& h) {
    delete static_cast<EigWork*>(h.eig_work);
    h.eig_work = nullptr;
}

struct Solver {
    Ham& h;
    EigWork& w;
    const dzhip_eig_opts& o;
    hipStream_t s;
    size_t n;
    int nb, max_sub;
    bool fp64_apply;
    dzhip_eig_stats st{};
    unsigned long long seed;

    double2* V() { return w.V.as<double2>(); }
    double2* W() { return w.W.as<double2>(); }
    double2* T() { return w.T.as<double2>(); }

    dim3 cols_grid(int cols) const { return dim3((unsigned)((n + 255) / 256), (unsigned)cols); }

    // W_out[:, 0:cols] = H V_in[:, 0:cols] in the current apply precision.
    void apply(const double2* vin, double2* wout, int cols) {
        if (cols <= 0) return;
        if (fp64_apply) {
            ham_apply_dev<double>(h, vin, wout, cols, true, true);
            st.fp64_applies += cols;
        } else {
            const size_t cnt = n * cols;
            w.x32.ensure(cnt * sizeof(float2));
            w.y32.ensure(cnt * sizeof(float2));
            convert_to_float(vin, w.x32.as<float2>(), cnt, s);
            ham_apply_dev<float>(h, w.x32.ptr, w.y32.ptr, cols, true, true);
            convert_to_double(w.y32.as<float2>(), wout, cnt, s);
            st.fp32_applies += cols;
        }
        st.apply_calls += cols;
    }

    // C[0:M, 0:N] (ld ldc) = A[:, 0:M]^H B[:, 0:N], columns of length n with leading dimension n.
    void gram(const double2* A, int M, const double2* B, int N, double2* C, int ldc) {
        if (M <= 0 || N <= 0) return;
        const int K = (int)n;
        const int mtiles = (M + kGramTile - 1) / kGramTile, ntiles = (N + kGramTile - 1) / kGramTile;
        // K-chunks (multiples of the stage depth) so that about three workgroups per CU run.
        const int want = std::max(1, 768 / (mtiles * ntiles));
        int kchunk = (K + want - 1) / want;
        kchunk = std::max(kGramK * 8, (kchunk + kGramK - 1) / kGramK * kGramK);
        const int nks = (K + kchunk - 1) / kchunk;
        w.gram_part.ensure((size_t)nks * M * N * sizeof(double2));
        hipLaunchKernelGGL(k_gram_cn, dim3((unsigned)(mtiles * nks), (unsigned)ntiles), dim3(kGramThreads), 0, s, A,
                           K, M, B, K, N, K, kchunk, mtiles, w.gram_part.as<double2>());
        DZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)((M + 255) / 256), (unsigned)N), dim3(256), 0, s,
                           w.gram_part.as<double2>(), nks, M, N, C, ldc);
        DZ_HIP(hipGetLastError());
    }

    void gemm(rocblas_operation ta, rocblas_operation tb, int m, int nn, int k, zc alpha, const double2* A,
              int lda, const double2* B, int ldb, zc beta, double2* C, int ldc) {
        if (m <= 0 || nn <= 0) return;
        if (ta == rocblas_operation_none && tb == rocblas_operation_none && alpha.imag() == 0.0 && beta.imag() == 0.0) {
            // Basis times coefficients: the library's own MFMA kernel.
            hipLaunchKernelGGL(k_gemm_nn, dim3((unsigned)((m + kGramTile - 1) / kGramTile),
                                               (unsigned)((nn + kGramTile - 1) / kGramTile)),
                               dim3(kGramThreads), 0, s, A, lda, B, ldb, C, ldc, m, k, nn, alpha.real(), beta.real());
            DZ_HIP(hipGetLastError());
            return;
        }
        DZ_BLAS(rocblas_zgemm(w.blas, ta, tb, m, nn, k, &alpha, reinterpret_cast<const zc*>(A), lda,
                              reinterpret_cast<const zc*>(B), ldb, &beta, reinterpret_cast<zc*>(C), ldc));
    }

    std::vector<double> col_norms(const double2* v, int cols) {
        std::vector<double> out(cols);
        if (cols == 0) return out;
        hipLaunchKernelGGL(k_col_norm2, dim3(cols), dim3(256), 0, s, v, n, (int)n, w.norms.as<double>());
        DZ_HIP(hipGetLastError());
        DZ_HIP(hipMemcpyAsync(out.data(), w.norms.ptr, cols * sizeof(double), hipMemcpyDeviceToHost, s));
        DZ_HIP(hipStreamSynchronize(s));
        for (double& v2 : out) v2 = std::sqrt(v2);
        return out;
    }

    // Orthogonalise Q[:, 0:a] against V[:, 0:m] (two block CGS passes) and orthonormalise it
    // with Cholesky QR.  A column is dropped when projection leaves less than 1e-8 of its norm
    // (the reference's MGS threshold, common.zig:263-284, applied relative to the direction's
    // own norm so that tolerances below 1e-8 Ry remain reachable).  Returns the number of
    // columns kept (compacted to the front).
    int orthonormalize_new(double2* Q, int a, int m) {
        const int ldn = (int)n;
        const std::vector<double> nrm0 = col_norms(Q, a);
        for (int pass = 0; pass < 2 && m > 0; ++pass) {
            // S = V^H Q ; Q -= V S
            gram(V(), m, Q, a, w.S.as<double2>(), max_sub);
            gemm(rocblas_operation_none, rocblas_operation_none, ldn, a, m, zc(-1.0, 0.0), V(), ldn,
                 w.S.as<double2>(), max_sub, zc(1.0, 0.0), Q, ldn);
        }
        std::vector<double> nrm = col_norms(Q, a);
        std::vector<int> keep;
        std::vector<double> inv;
        for (int j = 0; j < a; ++j)
            if (nrm0[j] > 0.0 && nrm[j] > 1e-8 * nrm0[j]) {
                keep.push_back(j);
                inv.push_back(1.0 / nrm[j]);
            }
        const int k = (int)keep.size();
        if (k == 0) return 0;
        // Compact and normalise into HQ (scratch), then copy back to Q.
        DZ_HIP(hipMemcpyAsync(w.idx.ptr, keep.data(), k * sizeof(int), hipMemcpyHostToDevice, s));
        DZ_HIP(hipMemcpyAsync(w.scl.ptr, inv.data(), k * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_gather_cols, cols_grid(k), dim3(256), 0, s, Q, w.idx.as<int>(), w.scl.as<double>(),
                           w.HQ.as<double2>(), n, (int)n);
        DZ_HIP(hipGetLastError());
        DZ_HIP(hipMemcpyAsync(Q, w.HQ.ptr, n * k * sizeof(double2), hipMemcpyDeviceToDevice, s));
        int kk = k;
        for (int pass = 0; pass < 2 && kk > 0; ++pass) kk = ortho_block(Q, kk);
        return kk;
    }

    // G = R^H R in place (upper triangle, leading dimension k), the verdict in w.info[0]: the
    // library's one-workgroup factorisation up to kPotrfMax, rocSOLVER's zpotrf beyond.
    void potrf_upper(double2* G, int k) {
        static_assert(sizeof(rocblas_int) == sizeof(int), "info is read as int");
        if (k <= kPotrfMax) {
            hipLaunchKernelGGL(k_potrf_upper, dim3(1), dim3(256), 0, s, G, k, k, w.info.as<int>());
            DZ_HIP(hipGetLastError());
        } else {
            DZ_BLAS(rocsolver_zpotrf(w.blas, rocblas_fill_upper, k, reinterpret_cast<zc*>(G), k,
                                     w.info.as<rocblas_int>()));
        }
    }

    // Orthonormalise Q[:, 0:k] in place: Cholesky QR, or, when the Gram matrix is not
    // numerically positive definite, its eigen-decomposition G = U L U^H with the directions of
    // L < 1e-12 max(L) dropped and Q <- Q U L^{-1/2} (compacted to the front).  Returns the
    // number of columns kept.  Cholesky counts as failed also when a pivot keeps less than 1e-7 of
    // its column's norm (R_jj < 1e-7 sqrt(G_jj)): such a column is numerically dependent on the
    // ones before it, and dividing by its pivot would turn rounding noise into a basis vector.
    int ortho_block(double2* Q, int k) {
        const int ldn = (int)n;
        gram(Q, k, Q, k, w.G.as<double2>(), k);
        DZ_HIP(hipMemcpyAsync(w.S.ptr, w.G.ptr, (size_t)k * k * sizeof(double2), hipMemcpyDeviceToDevice, s));
        potrf_upper(w.G.as<double2>(), k);
        // The factorisation's verdict on the device (info and every pivot against its column's norm,
        // the latter from the copy in S), one int back to the host.
        int* ok_dev = w.info.as<int>() + 1;
        hipLaunchKernelGGL(k_chol_check, dim3(1), dim3(256), 0, s, w.S.as<double2>(), w.G.as<double2>(), k,
                           w.info.as<rocblas_int>(), ok_dev);
        DZ_HIP(hipGetLastError());
        int ok = 0;
        DZ_HIP(hipMemcpyAsync(&ok, ok_dev, sizeof(int), hipMemcpyDeviceToHost, s));
        DZ_HIP(hipStreamSynchronize(s));
        const zc one(1.0, 0.0), zero(0.0, 0.0);
        if (ok) {
            DZ_BLAS(rocblas_ztrsm(w.blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                  rocblas_diagonal_non_unit, ldn, k, &one, w.G.as<zc>(), k, reinterpret_cast<zc*>(Q), ldn));
            return k;
        }
        st.ortho_fallbacks++;
        // Eigen-decomposition of the Gram matrix (its copy in S, lower triangle) into G.
        const std::vector<double> lam = subspace_eigh(w, w.S.as<double2>(), k, k, k, w.G.as<double2>(), k, s);
        const double thr = 1e-12 * std::max(lam[k - 1], 0.0);
        int first = 0;
        while (first < k && !(lam[first] > thr)) ++first;
        const int kk = k - first;
        if (kk == 0) return 0;
        gemm(rocblas_operation_none, rocblas_operation_none, ldn, kk, k, one, Q, ldn,
             w.G.as<double2>() + (size_t)first * k, k, zero, w.HQ.as<double2>(), ldn);
        std::vector<int> idx(kk);
        std::vector<double> scl(kk);
        for (int j = 0; j < kk; ++j) {
            idx[j] = j;
            scl[j] = 1.0 / std::sqrt(lam[first + j]);
        }
        DZ_HIP(hipMemcpyAsync(w.idx.ptr, idx.data(), kk * sizeof(int), hipMemcpyHostToDevice, s));
        DZ_HIP(hipMemcpyAsync(w.scl.ptr, scl.data(), kk * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_gather_cols, cols_grid(kk), dim3(256), 0, s, w.HQ.as<double2>(), w.idx.as<int>(),
                           w.scl.as<double>(), Q, n, (int)n);
        DZ_HIP(hipGetLastError());
        DZ_HIP(hipStreamSynchronize(s));  // idx / scl are host vectors
        return kk;
    }

    // Q <- Q R^{-1} with G = Q^H Q = R^H R.  Throws if G is not positive definite.
    void cholqr(double2* Q, int k) {
        const int ldn = (int)n;
        gram(Q, k, Q, k, w.G.as<double2>(), k);
        potrf_upper(w.G.as<double2>(), k);
        rocblas_int info = 0;
        DZ_HIP(hipMemcpyAsync(&info, w.info.ptr, sizeof(info), hipMemcpyDeviceToHost, s));
        DZ_HIP(hipStreamSynchronize(s));
        DZ_REQUIRE(info == 0, DZHIP_ERR_HIP, "Cholesky QR: Gram matrix not positive definite");
        const zc one(1.0, 0.0);
        DZ_BLAS(rocblas_ztrsm(w.blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                              rocblas_diagonal_non_unit, ldn, k, &one, w.G.as<zc>(), k, reinterpret_cast<zc*>(Q), ldn));
    }

    // T[0:m+a, m:m+a] = V[:, 0:m+a]^H W[:, m:m+a], mirrored into the rows.
    void extend_T(int m, int a) {
        gram(V(), m + a, W() + (size_t)m * n, a, T() + (size_t)m * max_sub, max_sub);
        hipLaunchKernelGGL(k_mirror_new, dim3((unsigned)((m + a + 255) / 256), (unsigned)a), dim3(256), 0, s, T(),
                           max_sub, m, a);
        DZ_HIP(hipGetLastError());
    }

    // Orthonormal start block, its H image and T.  Columns [0, cols) come from the host start
    // vectors or, with reuse, from the previous solve's Ritz vectors still in X; the rest are
    // unit vectors at the smallest diag entries or LCG values (common.zig:287-319), with the
    // generator advanced past the columns that were supplied.  A start block that is not of full
    // rank (duplicated or dependent columns) keeps its independent directions and is completed
    // with fresh LCG vectors orthogonalised against them, up to three rounds, as
    // orthonormalize_all replaces dependent columns (common.zig:298-312).
    void start(int m, const double* init, uint64_t init_cols, bool reuse, const double* diag_host) {
        size_t cols = 0;
        if (init && init_cols > 0) cols = std::min<uint64_t>(init_cols, (uint64_t)m);
        else if (reuse) cols = std::min<uint64_t>(w.x_cols, (uint64_t)m);
        const size_t rest = (size_t)m - cols;
        const unsigned long long A = 6364136223846793005ULL, C = 1ULL;
        // x_{k+j} = Aj x_k + Cj: affine map of j steps by repeated squaring.
        unsigned long long aj = 1, cj = 0, ap = A, cp = C;
        for (unsigned long long j = 2ULL * n * cols; j; j >>= 1) {
            if (j & 1) {
                aj = ap * aj;
                cj = ap * cj + cp;
            }
            cp = ap * cp + cp;
            ap = ap * ap;
        }
        unsigned long long x = aj * seed + cj;
        auto next = [&]() {
            x = x * A + C;
            return (double)(x >> 11) / 9007199254740992.0;
        };
        auto lcg_block = [&](std::vector<double2>& v, size_t ncol) {
            v.resize(n * ncol);
            for (size_t c = 0; c < ncol; ++c)
                for (size_t i = 0; i < n; ++i) {
                    const double re = next() - 0.5;
                    const double im = next() - 0.5;
                    v[c * n + i] = make_double2(re, im);
                }
        };
        std::vector<double2> v;
        if (o.init_diagonal) {
            std::vector<size_t> order(n);
            for (size_t i = 0; i < n; ++i) order[i] = i;
            std::stable_sort(order.begin(), order.end(),
                             [&](size_t a, size_t b) { return diag_host[a] < diag_host[b]; });
            v.assign(n * rest, make_double2(0.0, 0.0));
            for (size_t c = 0; c < rest; ++c) v[c * n + order[cols + c]] = make_double2(1.0, 0.0);
        } else {
            lcg_block(v, rest);
        }
        if (init && init_cols > 0)
            DZ_HIP(hipMemcpyAsync(V(), init, cols * n * sizeof(double2), hipMemcpyHostToDevice, s));
        else if (cols > 0)
            DZ_HIP(hipMemcpyAsync(V(), w.X.ptr, cols * n * sizeof(double2), hipMemcpyDeviceToDevice, s));
        if (rest > 0)
            DZ_HIP(hipMemcpyAsync(V() + cols * n, v.data(), v.size() * sizeof(double2), hipMemcpyHostToDevice, s));
        DZ_HIP(hipStreamSynchronize(s));
        int k = ortho_block(V(), m);
        if (k > 0) k = ortho_block(V(), k);
        for (int round = 0; round < 3 && k < m; ++round) {
            const int need = m - k;
            lcg_block(v, (size_t)need);
            DZ_HIP(hipMemcpyAsync(w.Q.ptr, v.data(), v.size() * sizeof(double2), hipMemcpyHostToDevice, s));
            const int got = orthonormalize_new(w.Q.as<double2>(), need, k);
            if (got > 0)
                DZ_HIP(hipMemcpyAsync(V() + (size_t)k * n, w.Q.ptr, (size_t)got * n * sizeof(double2),
                                      hipMemcpyDeviceToDevice, s));
            DZ_HIP(hipStreamSynchronize(s));  // v is refilled next round
            k += got;
        }
        DZ_REQUIRE(k == m, DZHIP_ERR_STATE, "start block: no orthonormal basis of the requested size");
        apply(V(), W(), m);
        extend_T(0, m);
    }

    // Lowest nb eigenpairs of T[0:m, 0:m] (m >= nb); eigenvectors land in Tw, eigenvalues (host)
    // returned.
    std::vector<double> rayleigh_ritz(int m) {
        return subspace_eigh(w, T(), max_sub, m, std::min(m, nb), w.Tw.as<double2>(), max_sub, s);
    }

    void ritz(int m) {
        const int ldn = (int)n;
        gemm(rocblas_operation_none, rocblas_operation_none, ldn, nb, m, zc(1.0, 0.0), V(), ldn, w.Tw.as<double2>(),
             max_sub, zc(0.0, 0.0), w.X.as<double2>(), ldn);
        gemm(rocblas_operation_none, rocblas_operation_none, ldn, nb, m, zc(1.0, 0.0), W(), ldn, w.Tw.as<double2>(),
             max_sub, zc(0.0, 0.0), w.HX.as<double2>(), ldn);
    }

    // Basis <- [X, Qnew] (already orthonormal), W <- [HX, HQnew], T rebuilt.
    void restart_with(int a) {
        DZ_HIP(hipMemcpyAsync(V(), w.X.ptr, n * nb * sizeof(double2), hipMemcpyDeviceToDevice, s));
        DZ_HIP(hipMemcpyAsync(W(), w.HX.ptr, n * nb * sizeof(double2), hipMemcpyDeviceToDevice, s));
        if (a > 0) {
            DZ_HIP(hipMemcpyAsync(V() + (size_t)nb * n, w.Q.ptr, n * a * sizeof(double2), hipMemcpyDeviceToDevice, s));
            DZ_HIP(hipMemcpyAsync(W() + (size_t)nb * n, w.HQ.ptr, n * a * sizeof(double2), hipMemcpyDeviceToDevice, s));
        }
        extend_T(0, nb + a);
    }
};

static void ensure_work(Ham& h, int max_sub, int nb) {
    auto* w = static_cast<EigWork*>(h.eig_work);
    if (!w) {
        w = new EigWork();
        h.eig_work = w;
        if (rocblas_create_handle(&w->blas) != rocblas_status_success) {
            w->blas = nullptr;
            throw Error{DZHIP_ERR_HIP, "rocblas_create_handle failed"};
        }
    }
    DZ_BLAS(rocblas_set_stream(w->blas, h.stream()));
    w->stream = h.stream();
    const size_t n = h.npw;
    const size_t c = sizeof(double2);
    if (w->X.bytes < n * nb * c) w->x_cols = 0;  // X is reallocated below
    w->V.ensure(n * max_sub * c);
    w->W.ensure(n * max_sub * c);
    w->X.ensure(n * nb * c);
    w->HX.ensure(n * nb * c);
    w->R.ensure(n * nb * c);
    w->Q.ensure(n * max_sub * c);
    w->HQ.ensure(n * max_sub * c);
    w->T.ensure((size_t)max_sub * max_sub * c);
    w->Tw.ensure((size_t)max_sub * max_sub * c);
    w->S.ensure((size_t)max_sub * max_sub * c);
    w->G.ensure((size_t)max_sub * max_sub * c);
    w->evals.ensure(max_sub * sizeof(double));
    w->ework.ensure(max_sub * sizeof(double));
    w->info.ensure(64);
    w->norms.ensure(max_sub * sizeof(double));
    w->lam.ensure(max_sub * sizeof(double));
    w->ekin.ensure(max_sub * sizeof(double));
    w->diag.ensure(n * sizeof(double));
    w->sel.ensure(max_sub * sizeof(int));
    w->scl.ensure(max_sub * sizeof(double));
    w->idx.ensure(max_sub * sizeof(int));
    w->rnorm.ensure(max_sub * sizeof(double));
}

void eig_solve(Ham& h, const double* diag, uint64_t nbands, const dzhip_eig_opts& o, const double* init,
               uint64_t init_cols, double* values, double* vectors, dzhip_eig_stats* stats_out) {
    const size_t n = h.npw;
    DZ_REQUIRE(nbands > 0 && nbands <= n, DZHIP_ERR_ARG, "nbands must be in 1..n");
    DZ_REQUIRE(o.precision == DZHIP_FP64 || o.precision == DZHIP_FP32, DZHIP_ERR_ARG, "bad precision");
    DZ_REQUIRE(o.preconditioner >= DZHIP_PRECOND_KINETIC && o.preconditioner <= DZHIP_PRECOND_MLP, DZHIP_ERR_ARG,
               "bad preconditioner");
    DZ_REQUIRE(o.preconditioner != DZHIP_PRECOND_MLP || h.has_mlp, DZHIP_ERR_STATE,
               "DZHIP_PRECOND_MLP without a model (dzhip_ham_set_precond_model)");
    const int nb = (int)nbands;
    const int max_sub = (int)(o.max_subspace == 0 ? std::min<uint64_t>(n, 4 * nbands + 8)
                                                   : std::min<uint64_t>(n, std::max<uint64_t>(o.max_subspace, nbands)));
    const int block = (int)(o.block_size == 0 ? std::min<uint64_t>(max_sub, nbands + 4)
                                              : std::min<uint64_t>(max_sub, std::max<uint64_t>(o.block_size, nbands)));
    ensure_work(h, max_sub, nb);
    EigWork& w = *static_cast<EigWork*>(h.eig_work);
    Solver S{h, w, o, h.stream(), n, nb, max_sub, o.precision == DZHIP_FP64, {}, o.seed ? o.seed : 0x6f3d3cbeULL};
    hipStream_t s = S.s;
    DZ_HIP(hipMemcpyAsync(w.diag.ptr, diag, n * sizeof(double), hipMemcpyHostToDevice, s));
    const double switch_tol = o.mixed_switch_tol > 0 ? o.mixed_switch_tol : 1e-5;

    int m = block;
    // Leave the FP32 apply: restart from the Ritz vectors with an FP64 image.
    auto switch_to_fp64 = [&](uint64_t it) {
        S.fp64_apply = true;
        S.st.precision_switch_iter = it + 1;
        DZ_HIP(hipMemcpyAsync(S.V(), w.X.ptr, n * nb * sizeof(double2), hipMemcpyDeviceToDevice, s));
        S.cholqr(S.V(), nb);
        S.apply(S.V(), S.W(), nb);
        S.extend_T(0, nb);
        m = nb;
    };
    const bool reuse = o.reuse_vectors != 0 && !(init && init_cols > 0) && w.x_cols > 0;
    S.start(m, init, init_cols, reuse, diag);
    w.x_cols = 0;
    w.x_converged = false;
    const int nsnap = (int)std::min<uint64_t>(o.collect_snapshots, 64);
    if (nsnap > 0) {
        const size_t blk = n * nb * sizeof(double2);
        w.snapX.ensure(nsnap * blk);
        w.snapR.ensure(nsnap * blk);
        w.snapLam.ensure(nsnap * nb * sizeof(double));
        w.snapRn.ensure(nsnap * nb * sizeof(double));
        w.snapEk.ensure(nsnap * nb * sizeof(double));
    }
    w.snaps = 0;
    w.snap_nb = nb;
    int precond = o.preconditioner;
    // A learned model that fell back in two solves in a row on this context sits out the next
    // kMlpRest solves (the base preconditioner runs; precond_fallback reports iteration 1).
    constexpr int kMlpRest = 8;
    constexpr size_t kCurveMax = 64;  // iterations of the base preconditioner's progress record
    if (precond == DZHIP_PRECOND_MLP && h.mlp_rest > 0) {
        h.mlp_rest--;
        precond = h.mlp_base;
        S.st.precond_fallback = 1;
    }
    const bool mlp_ran = precond == DZHIP_PRECOND_MLP;
    std::vector<double> res_hist;
    const int nconv = (o.nconv == 0 || o.nconv > (uint64_t)nb) ? nb : (int)o.nconv;
    std::vector<double> lam(nb), best(nb);
    bool have_best = false;
    uint64_t iter = 0;
    w.ev_used = 0;
    for (; iter < o.max_iter; ++iter) {
        mark(w, s, SEC_RR);
        std::vector<double> ev = S.rayleigh_ritz(m);
        for (int b = 0; b < nb; ++b) lam[b] = ev[b];
        best = lam;
        have_best = true;
        mark(w, s, SEC_RITZ);
        S.ritz(m);
        DZ_HIP(hipMemcpyAsync(w.lam.ptr, lam.data(), nb * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_residual, S.cols_grid(nb), dim3(256), 0, s, w.HX.as<double2>(), w.X.as<double2>(),
                           w.lam.as<double>(), w.R.as<double2>(), n, (int)n);
        DZ_HIP(hipGetLastError());
        std::vector<double> rn = S.col_norms(w.R.as<double2>(), nb);
        // Convergence is judged on the first nconv bands (all by default); the others are a buffer.
        const double max_res = *std::max_element(rn.begin(), rn.begin() + nconv);
        S.st.max_residual = max_res;
        if (iter == 0) S.st.initial_residual = max_res;
        {
            // Convergence progress: mean over the bands that must converge of log10 max(r_b, tol).
            double sl = 0.0;
            for (int b = 0; b < nconv; ++b) sl += std::log10(std::max(rn[b], std::max(o.tol, 1e-300)));
            res_hist.push_back(sl / nconv);
        }
        if (w.snaps < nsnap && S.fp64_apply) {
            const size_t blk = n * nb;
            double2* sx = w.snapX.as<double2>() + w.snaps * blk;
            double2* sr = w.snapR.as<double2>() + w.snaps * blk;
            DZ_HIP(hipMemcpyAsync(sx, w.X.ptr, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
            DZ_HIP(hipMemcpyAsync(sr, w.R.ptr, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
            DZ_HIP(hipMemcpyAsync(w.snapLam.as<double>() + w.snaps * nb, lam.data(), nb * sizeof(double),
                                  hipMemcpyHostToDevice, s));
            DZ_HIP(hipMemcpyAsync(w.snapRn.as<double>() + w.snaps * nb, rn.data(), nb * sizeof(double),
                                  hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_kinetic_energy, dim3(nb), dim3(256), 0, s, w.X.as<double2>(), w.diag.as<double>(), n,
                               (int)n, w.snapEk.as<double>() + w.snaps * nb);
            DZ_HIP(hipGetLastError());
            DZ_HIP(hipStreamSynchronize(s));  // lam / rn are host vectors rebuilt next iteration
            w.snaps++;
        }
        // Learned preconditioner: fall back to its base preconditioner for the rest of the solve
        // when, from the fourth iteration on, the progress so far (drop of the mean log residual
        // since the first iteration) trails the base's mean progress at the same iteration on
        // this context by more than 0.25 decades, or (no base record for this iteration yet)
        // when the last three iterations gained less than 0.05 decades per iteration.
        if (precond == DZHIP_PRECOND_MLP && res_hist.size() > 3) {
            const size_t i = res_hist.size() - 1;
            const auto& curve = h.precond_curve[h.mlp_base == DZHIP_PRECOND_TPA ? 1 : 0];
            const double drop = res_hist.front() - res_hist.back();
            const bool behind = i < curve.size() ? drop < curve[i] - 0.25
                                                 : (res_hist[i - 3] - res_hist[i]) / 3.0 < 0.05;
            if (behind) {
                precond = h.mlp_base;
                S.st.precond_fallback = iter + 1;
            }
        }
        // FP32 apply: a tolerance looser than the switch threshold is met in FP32; tighter
        // tolerances switch the apply to FP64 once the residual is below the threshold.
        if (max_res < o.tol && (S.fp64_apply || o.tol >= switch_tol)) {
            S.st.converged = 1;
            break;
        }
        if (!S.fp64_apply && max_res < switch_tol) {
            switch_to_fp64(iter);
            continue;
        }
        // New directions for the unconverged bands, in band order (serial.zig:249-261).
        const int room = std::min(nb, max_sub - nb);
        std::vector<int> sel;
        for (int b = 0; b < nb && (int)sel.size() < room; ++b)
            if (rn[b] >= o.tol) sel.push_back(b);
        int a = 0;
        if (!sel.empty()) {
            mark(w, s, SEC_PRECOND);
            DZ_HIP(hipMemcpyAsync(w.sel.ptr, sel.data(), sel.size() * sizeof(int), hipMemcpyHostToDevice, s));
            if (precond != DZHIP_PRECOND_KINETIC) {
                hipLaunchKernelGGL(k_kinetic_energy, dim3(nb), dim3(256), 0, s, w.X.as<double2>(), w.diag.as<double>(),
                                   n, (int)n, w.ekin.as<double>());
                DZ_HIP(hipGetLastError());
            }
            if (precond == DZHIP_PRECOND_MLP) {
                DZ_HIP(hipMemcpyAsync(w.rnorm.ptr, rn.data(), nb * sizeof(double), hipMemcpyHostToDevice, s));
                precondition_mlp(w.R.as<double2>(), w.sel.as<int>(), (int)sel.size(), w.lam.as<double>(),
                                 w.ekin.as<double>(), w.rnorm.as<double>(), w.diag.as<double>(), h.mlp_params.as<float>(),
                                 h.mlp_energy_scale, h.mlp_base, w.Q.as<double2>(), n, (int)n, s);
            } else {
                hipLaunchKernelGGL(k_precondition, S.cols_grid((int)sel.size()), dim3(256), 0, s, w.R.as<double2>(),
                                   w.sel.as<int>(), w.lam.as<double>(), w.ekin.as<double>(), w.diag.as<double>(),
                                   w.Q.as<double2>(), n, (int)n, precond == DZHIP_PRECOND_TPA ? 1 : 0);
                DZ_HIP(hipGetLastError());
            }
            mark(w, s, SEC_ORTHO);
            a = S.orthonormalize_new(w.Q.as<double2>(), (int)sel.size(), m);
        }
        if (a == 0) {
            // Nothing could be added: every residual is individually below tol
            // (serial.zig:262-273) or the new directions lie in the basis.
            if (!S.fp64_apply && o.tol < switch_tol) {
                switch_to_fp64(iter);
                continue;
            }
            S.st.converged = max_res < o.tol ? 1 : 0;
            break;
        }
        mark(w, s, SEC_APPLY);
        S.apply(w.Q.as<double2>(), w.HQ.as<double2>(), a);
        mark(w, s, SEC_EXTEND);
        if (m + a > max_sub) {
            // Thick restart from [Ritz vectors, new block]; their H images are W c and H Q.
            S.restart_with(a);
            m = nb + a;
            S.st.restarts++;
            continue;
        }
        DZ_HIP(hipMemcpyAsync(S.V() + (size_t)m * n, w.Q.ptr, n * a * sizeof(double2), hipMemcpyDeviceToDevice, s));
        DZ_HIP(hipMemcpyAsync(S.W() + (size_t)m * n, w.HQ.ptr, n * a * sizeof(double2), hipMemcpyDeviceToDevice, s));
        S.extend_T(m, a);
        m += a;
    }
    mark(w, s, SEC_END);
    S.st.iterations = iter;
    if (mlp_ran) {
        h.mlp_fail_streak = S.st.precond_fallback > 0 ? h.mlp_fail_streak + 1 : 0;
        if (h.mlp_fail_streak >= 2) {
            h.mlp_rest = kMlpRest;
            h.mlp_fail_streak = 0;
        }
    }
    // Mean progress per iteration index of kinetic-shift and TPA solves on this context (drop of
    // the mean log residual since the first iteration), the yardstick of the learned
    // preconditioner's fallback.  A converged solve's last value stands for the iterations it
    // did not need.
    if ((o.preconditioner == DZHIP_PRECOND_KINETIC || o.preconditioner == DZHIP_PRECOND_TPA) &&
        res_hist.size() >= 2) {
        const int c = o.preconditioner == DZHIP_PRECOND_TPA ? 1 : 0;
        auto& curve = h.precond_curve[c];
        auto& cnt = h.precond_curve_n[c];
        const size_t len = std::max(curve.size(), S.st.converged ? (size_t)kCurveMax : res_hist.size());
        const size_t keep = std::min(len, (size_t)kCurveMax);
        curve.resize(keep, 0.0);
        cnt.resize(keep, 0);
        for (size_t i = 0; i < keep; ++i) {
            if (i >= res_hist.size() && !S.st.converged) break;
            const double d = res_hist.front() - res_hist[std::min(i, res_hist.size() - 1)];
            cnt[i]++;
            curve[i] += (d - curve[i]) / cnt[i];
        }
    }
    if (have_best) {
        std::memcpy(values, best.data(), nb * sizeof(double));
        w.x_cols = nb;
        w.x_converged = S.st.converged != 0;
    }
    if (vectors) DZ_HIP(hipMemcpyAsync(vectors, w.X.ptr, n * nb * sizeof(double2), hipMemcpyDeviceToHost, s));
    DZ_HIP(hipStreamSynchronize(s));
    double sec[SEC_COUNT];
    section_times(w, sec);
    S.st.t_rr_ms = sec[SEC_RR];
    S.st.t_ritz_ms = sec[SEC_RITZ];
    S.st.t_precond_ms = sec[SEC_PRECOND];
    S.st.t_ortho_ms = sec[SEC_ORTHO];
    S.st.t_apply_ms = sec[SEC_APPLY];
    S.st.t_extend_ms = sec[SEC_EXTEND];
    if (stats_out) *stats_out = S.st;
}

// ----------------------------------------------------------------- band-by-band CG
namespace {

struct CgRun {
    Ham& h;
    EigWork& w;
    hipStream_t s;
    size_t n;
    int ldn;
    dzhip_eig_stats& st;

    dim3 grid1() const { return dim3((unsigned)((n + 255) / 256)); }
    cplx dot(const double2* a, const double2* b) {  // <a|b>
        zc r;
        DZ_BLAS(rocblas_zdotc(w.blas, ldn, reinterpret_cast<const zc*>(a), 1, reinterpret_cast<const zc*>(b), 1, &r));
        return cplx(r.real(), r.imag());
    }
    double nrm(const double2* a) {
        double r = 0.0;
        DZ_BLAS(rocblas_dznrm2(w.blas, ldn, reinterpret_cast<const zc*>(a), 1, &r));
        return r;
    }
    void axpy(cplx alpha, const double2* x, double2* y) {
        const zc al(alpha.real(), alpha.imag());
        DZ_BLAS(rocblas_zaxpy(w.blas, ldn, &al, reinterpret_cast<const zc*>(x), 1, reinterpret_cast<zc*>(y), 1));
    }
    void scal(double a, double2* x) {
        DZ_BLAS(rocblas_zdscal(w.blas, ldn, &a, reinterpret_cast<zc*>(x), 1));
    }
    void copy(double2* dst, const double2* src) {
        DZ_HIP(hipMemcpyAsync(dst, src, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    }
    void apply(const double2* x, double2* y) {
        ham_apply_dev<double>(h, x, y, 1, true, true);
        st.apply_calls++;
        st.fp64_applies++;
    }
    // v -= sum_{j<m} <V_j|v> V_j, one column at a time in order (modified Gram-Schmidt).
    void orth(double2* v, const double2* V, int m) {
        for (int j = 0; j < m; ++j) axpy(-dot(V + (size_t)j * n, v), V + (size_t)j * n, v);
    }
    void combine(const double2* a, const double2* b, const double2* c, cplx c0, cplx c1, cplx c2, bool use_c,
                 double2* out) {
        hipLaunchKernelGGL(k_combine3, grid1(), dim3(256), 0, s, a, b, c, make_double2(c0.real(), c0.imag()),
                           make_double2(c1.real(), c1.imag()), make_double2(c2.real(), c2.imag()), use_c ? 1 : 0, out,
                           (int)n);
        DZ_HIP(hipGetLastError());
    }
};

}  // namespace

void cg_solve(Ham& h, const double* diag, uint64_t nbands, const dzhip_eig_opts& o, const double* init,
              uint64_t init_cols, double* values, double* vectors, dzhip_eig_stats* stats_out) {
    const size_t n = h.npw;
    DZ_REQUIRE(nbands > 0 && nbands <= n, DZHIP_ERR_ARG, "nbands must be in 1..n");
    const int nb = (int)nbands;
    const uint64_t max_iter = o.max_iter ? o.max_iter : 20;
    const double tol = o.tol > 0.0 ? o.tol : 1e-6;
    ensure_work(h, nb, nb);
    EigWork& w = *static_cast<EigWork*>(h.eig_work);
    // Work vectors: psi block in V, H psi in W[:,0], r/p/d/Hd/psi_old/Hpsi_old/psi_prev/Hpsi_prev
    // in Q columns 0..7 (Q holds n x max_sub >= n x 8 when nb >= 8; else grow it).
    w.Q.ensure(n * std::max(nb, 8) * sizeof(double2));
    w.W.ensure(n * std::max(nb, 8) * sizeof(double2));
    dzhip_eig_stats st{};
    hipStream_t s = h.stream();
    CgRun R{h, w, s, n, (int)n, st};
    DZ_BLAS(rocblas_set_pointer_mode(w.blas, rocblas_pointer_mode_host));
    DZ_HIP(hipMemcpyAsync(w.diag.ptr, diag, n * sizeof(double), hipMemcpyHostToDevice, s));

    // Start: copied vectors, reused device vectors, or the LCG; orthonormalised (MGS).
    double2* V = w.V.as<double2>();
    {
        const bool reuse = o.reuse_vectors != 0 && !(init && init_cols > 0) && w.x_cols > 0;
        size_t cols = 0;
        if (init && init_cols > 0) cols = std::min<uint64_t>(init_cols, nbands);
        else if (reuse) cols = std::min<uint64_t>(w.x_cols, nbands);
        std::vector<double2> v(n * (nb - cols));
        unsigned long long x = o.seed ? o.seed : 0x6f3d3cbeULL;
        for (size_t i = 0; i < v.size(); ++i) {
            auto nxt = [&]() {
                x = x * 6364136223846793005ULL + 1ULL;
                return (double)(x >> 11) / 9007199254740992.0;
            };
            const double re = nxt() - 0.5;
            const double im = nxt() - 0.5;
            v[i] = make_double2(re, im);
        }
        if (init && init_cols > 0) DZ_HIP(hipMemcpyAsync(V, init, cols * n * sizeof(double2), hipMemcpyHostToDevice, s));
        else if (cols > 0) DZ_HIP(hipMemcpyAsync(V, w.X.ptr, cols * n * sizeof(double2), hipMemcpyDeviceToDevice, s));
        if (!v.empty())
            DZ_HIP(hipMemcpyAsync(V + cols * n, v.data(), v.size() * sizeof(double2), hipMemcpyHostToDevice, s));
        DZ_HIP(hipStreamSynchronize(s));
        for (int b = 0; b < nb; ++b) {
            double2* vb = V + (size_t)b * n;
            R.orth(vb, V, b);
            const double nv = R.nrm(vb);
            DZ_REQUIRE(nv > 1e-8, DZHIP_ERR_ARG, "start vectors are linearly dependent");
            R.scal(1.0 / nv, vb);
        }
    }
    w.x_cols = 0;
    w.x_converged = false;
    double2* Q = w.Q.as<double2>();
    double2 *hpsi = w.W.as<double2>(), *r = Q, *p = Q + n, *d = Q + 2 * n, *hd = Q + 3 * n, *pold = Q + 4 * n,
            *hpold = Q + 5 * n, *pp = Q + 6 * n, *hpp = Q + 7 * n;
    std::vector<double> vals(nb);
    bool all_conv = true;
    for (int b = 0; b < nb; ++b) {
        double2* psi = V + (size_t)b * n;
        if (b > 0) {
            R.orth(psi, V, b);
            R.scal(1.0 / R.nrm(psi), psi);
        }
        R.apply(psi, hpsi);
        double lam = R.dot(psi, hpsi).real();
        double prev_zr = 0.0;
        bool has_prev = false, conv = false;
        for (uint64_t it = 0; it < max_iter; ++it) {
            st.iterations++;
            R.copy(r, hpsi);
            R.axpy(cplx(-lam, 0.0), psi, r);
            R.orth(r, V, b);
            if (R.nrm(r) < tol) {
                conv = true;
                break;
            }
            hipLaunchKernelGGL(k_precondition_one, R.grid1(), dim3(256), 0, s, r, w.diag.as<double>(), lam, p, (int)n);
            DZ_HIP(hipGetLastError());
            R.orth(p, V, b + 1);
            const double zr = R.dot(p, r).real();
            if (has_prev && prev_zr > 1e-30) {
                R.scal(std::max(0.0, zr / prev_zr), d);
                R.axpy(cplx(1.0, 0.0), p, d);
            } else {
                R.copy(d, p);
            }
            prev_zr = zr;
            R.orth(d, V, b + 1);
            const double dn = R.nrm(d);
            if (dn < 1e-14) break;
            R.scal(1.0 / dn, d);
            R.apply(d, hd);
            // Three-term Rayleigh-Ritz on {psi, d, psi_prev} when the previous psi adds a direction.
            bool three = false;
            double lam_new = lam;
            cplx c0, c1, c2;
            if (has_prev) {
                R.copy(pp, pold);
                R.copy(hpp, hpold);
                const cplx ov = R.dot(psi, pp);
                R.axpy(-ov, psi, pp);
                R.axpy(-ov, hpsi, hpp);
                const cplx ovd = R.dot(d, pp);
                R.axpy(-ovd, d, pp);
                R.axpy(-ovd, hd, hpp);
                const double pn = R.nrm(pp);
                if (pn > 1e-6) {
                    R.scal(1.0 / pn, pp);
                    R.scal(1.0 / pn, hpp);
                    const double a22 = R.dot(d, hd).real(), a33 = R.dot(pp, hpp).real();
                    const cplx a12 = R.dot(psi, hd), a13 = R.dot(psi, hpp), a23 = R.dot(d, hpp);
                    const cplx m3[9] = {lam, a12, a13, std::conj(a12), a22, a23, std::conj(a13), std::conj(a23), a33};
                    cplx vec[3];
                    herm3_lowest(m3, &lam_new, vec);
                    // Phase convention: the psi coefficient real and non-negative.
                    if (std::abs(vec[0]) > 0.0) {
                        const cplx ph = std::abs(vec[0]) / vec[0];
                        for (auto& c : vec) c *= ph;
                    }
                    c0 = vec[0];
                    c1 = vec[1];
                    c2 = vec[2];
                    three = true;
                }
            }
            if (!three) {
                const double a11 = lam, a22 = R.dot(d, hd).real();
                const cplx a12 = R.dot(psi, hd);
                const double avg = 0.5 * (a11 + a22), diff = 0.5 * (a11 - a22), off = std::norm(a12);
                lam_new = avg - std::sqrt(diff * diff + off);
                if (off > 1e-30) {
                    c0 = -a12;
                    c1 = cplx(a11 - lam_new, 0.0);
                } else if (a11 <= a22) {
                    c0 = 1.0;
                    c1 = 0.0;
                } else {
                    c0 = 0.0;
                    c1 = 1.0;
                }
                const double cn = std::sqrt(std::norm(c0) + std::norm(c1));
                if (cn > 1e-30) {
                    c0 /= cn;
                    c1 /= cn;
                }
                c2 = 0.0;
            }
            R.copy(pold, psi);
            R.copy(hpold, hpsi);
            has_prev = true;
            R.combine(pold, d, pp, c0, c1, c2, three, psi);
            R.combine(hpold, hd, hpp, c0, c1, c2, three, hpsi);
            const double pn2 = R.nrm(psi);
            if (pn2 > 1e-14) {
                R.scal(1.0 / pn2, psi);
                R.scal(1.0 / pn2, hpsi);
            }
            const double change = std::fabs(lam_new - lam);
            lam = lam_new;
            if (change < tol) {
                conv = true;
                break;
            }
        }
        vals[b] = lam;
        all_conv = all_conv && conv;
        if (b > 0) {
            R.orth(psi, V, b);
            R.scal(1.0 / R.nrm(psi), psi);
        }
    }
    // Final subspace rotation: T = V^H H V, eigen-decomposition, X = V C (cg.zig:450-482).
    double2* HV = w.HX.as<double2>();
    ham_apply_dev<double>(h, V, HV, nb, true, true);
    st.apply_calls += nb;
    st.fp64_applies += nb;
    const zc one(1.0, 0.0), zero(0.0, 0.0);
    DZ_BLAS(rocblas_zgemm(w.blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, nb, nb, (int)n, &one,
                          reinterpret_cast<const zc*>(V), (int)n, reinterpret_cast<const zc*>(HV), (int)n, &zero,
                          w.Tw.as<zc>(), nb));
    // Eigenvectors of T (in Tw) into S (nb x nb).
    const std::vector<double> ev = subspace_eigh(w, w.Tw.as<double2>(), nb, nb, nb, w.S.as<double2>(), nb, s);
    DZ_BLAS(rocblas_zgemm(w.blas, rocblas_operation_none, rocblas_operation_none, (int)n, nb, nb, &one,
                          reinterpret_cast<const zc*>(V), (int)n, w.S.as<zc>(), nb, &zero, w.X.as<zc>(), (int)n));
    std::copy(ev.begin(), ev.end(), values);
    if (vectors) DZ_HIP(hipMemcpyAsync(vectors, w.X.ptr, n * nb * sizeof(double2), hipMemcpyDeviceToHost, s));
    DZ_HIP(hipStreamSynchronize(s));
    w.x_cols = nb;
    w.x_converged = all_conv;
    st.converged = all_conv ? 1 : 0;
    if (stats_out) *stats_out = st;
}

void eig_release(Ham& h) { eig_work_release(h); }

size_t eig_training_samples(Ham& h, double energy_scale, int base, int stride, float* out_host, size_t cap) {
    auto* w = static_cast<EigWork*>(h.eig_work);
    DZ_REQUIRE(w && w->snaps > 0, DZHIP_ERR_STATE, "no training snapshots (solve with collect_snapshots > 0)");
    DZ_REQUIRE(w->x_converged && w->x_cols == (uint64_t)w->snap_nb, DZHIP_ERR_STATE,
               "the recording solve did not converge");
    DZ_REQUIRE(stride >= 1 && energy_scale > 0.0, DZHIP_ERR_ARG, "stride >= 1 and energy_scale > 0");
    DZ_REQUIRE(base == DZHIP_PRECOND_KINETIC || base == DZHIP_PRECOND_TPA, DZHIP_ERR_ARG,
               "base must be DZHIP_PRECOND_KINETIC or DZHIP_PRECOND_TPA");
    hipStream_t s = h.stream();
    DZ_BLAS(rocblas_set_stream(w->blas, s));
    const int nb = w->snap_nb;
    const size_t n = h.npw, blk = n * nb;
    const int per_col = (int)((n + stride - 1) / stride);
    const size_t total = std::min<size_t>(cap, (size_t)w->snaps * nb * per_col);
    if (total == 0) return 0;
    w->samples.ensure(6 * total * sizeof(float));
    w->bandsum.ensure(3 * nb * sizeof(double));
    w->Q.ensure(blk * sizeof(double2));
    w->G.ensure((size_t)nb * nb * sizeof(double2));
    const zc one(1.0, 0.0), zero(0.0, 0.0), mone(-1.0, 0.0);
    size_t done = 0;
    for (int i = 0; i < w->snaps && done < total; ++i) {
        const double2* xi = w->snapX.as<double2>() + i * blk;
        // C = X^H x_i, E = X C - x_i (the part of x_i outside the converged subspace, negated).
        DZ_BLAS(rocblas_zgemm(w->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, nb, nb, (int)n,
                              &one, reinterpret_cast<const zc*>(w->X.ptr), (int)n, reinterpret_cast<const zc*>(xi),
                              (int)n, &zero, w->G.as<zc>(), nb));
        DZ_HIP(hipMemcpyAsync(w->Q.ptr, xi, blk * sizeof(double2), hipMemcpyDeviceToDevice, s));
        DZ_BLAS(rocblas_zgemm(w->blas, rocblas_operation_none, rocblas_operation_none, (int)n, nb, nb, &one,
                              reinterpret_cast<const zc*>(w->X.ptr), (int)n, w->G.as<zc>(), nb, &mone,
                              w->Q.as<zc>(), (int)n));
        done += mlp_write_samples(w->Q.as<double2>(), w->snapR.as<double2>() + i * blk, w->snapLam.as<double>() + i * nb,
                                  w->snapEk.as<double>() + i * nb, w->snapRn.as<double>() + i * nb,
                                  w->diag.as<double>(), n, (int)n, nb, stride, energy_scale, base, w->bandsum.as<double>(),
                                  w->samples.as<float>(), total, done, s);
    }
    for (int f = 0; f < 6; ++f)
        DZ_HIP(hipMemcpyAsync(out_host + (size_t)f * cap, w->samples.as<float>() + (size_t)f * total,
                              done * sizeof(float), hipMemcpyDeviceToHost, s));
    DZ_HIP(hipStreamSynchronize(s));
    return done;
}

const double2* eig_solution(const Ham& h, uint64_t* cols) {
    const auto* w = static_cast<const EigWork*>(h.eig_work);
    *cols = w ? w->x_cols : 0;
    return (w && w->x_cols > 0) ? w->X.as<double2>() : nullptr;
}

}  // namespace dzhip
