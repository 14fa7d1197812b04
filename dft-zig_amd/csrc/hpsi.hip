// H·psi on the device: kinetic + local (pruned 3D FFT with V_loc fused) + KB nonlocal.
//
// Reference: apply_hamiltonian_batched (src/features/scf/apply.zig:1584-1618), the fused
// local path apply_local_potential_safe (apply.zig:820-852) and the batched nonlocal
// (apply.zig:1118-1482).
//
// Local part, per band, as three launches over all bands at once (own design):
//   plane_inv : one workgroup per (x-plane of the G-sphere, band): scatter the sphere
//               points of the plane into an LDS [z][y] tile, inverse FFT along z and y,
//               write the plane to T[band][plane][z][y].  Only the ~35 of 72 x-planes that
//               hold plane waves exist, so T is half a grid.
//   xline     : one workgroup per (z, band): gather the planes' rows into an LDS [x][y]
//               tile, inverse FFT along x, multiply by V_loc(r)/N, forward FFT along x,
//               write back only the rows of sphere x-planes.  FP32 with nx = 2^a 3^b 5^c
//               up to 100 can instead run k_xline_reg, one x-line per thread in registers
//               (fft_reg.h), when the autotuner finds it faster.
//   plane_fwd : one workgroup per (plane, band): forward FFT along y and z, gather the
//               sphere points, add |k+G|^2 x (and the nonlocal term already in y).
// HBM traffic per band: psi + H psi on the sphere and two round trips of half a grid.
//
// Nonlocal part: complex projector matrix A[p][G] = phi_p(G) e^{+iG.tau_a} held on the
// device; proj = A x (split-K), q = D proj / Omega per atom block, y = A^H q (before the
// local part, which then accumulates into y).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include <type_traits>
#include <vector>

#include "fft_lds.h"
#include "fft_reg.h"
#include "ham.h"

namespace dzhip {

// Columns (bands) per launch sequence of the apply; larger blocks are processed in passes.
constexpr size_t kMaxColsPerPass = 256;

// LDS layout of the tile kernels: tile rows of pitch ny+1 (odd in elements), then the
// twiddle tables of the axes used, then (plane kernels) the list of active y columns.
template <typename C>
__host__ __device__ inline size_t plane_lds_bytes(int ny, int nz) {
    return ((size_t)nz * (ny + 1) + ny + nz) * sizeof(C) + round_up((size_t)ny * sizeof(short), 16);
}
template <typename C>
__host__ __device__ inline size_t xline_lds_bytes(int nx, int ny) {
    return ((size_t)nx * (ny + 1) + nx) * sizeof(C);
}

// Sphere entries per thread whose global loads the plane kernels issue ahead of their tile work
// (FP64 runs twice the threads per tile at a tighter register budget).
template <typename T> struct SpherePrefetch { static constexpr int value = 2; };
template <> struct SpherePrefetch<double> { static constexpr int value = 1; };

template <typename T>
__global__ __launch_bounds__(TileThreads<T>::value) void k_plane_inv(
    const typename C2T<T>::type* __restrict__ x, size_t ldx, typename C2T<T>::type* __restrict__ tbuf,
    const int* __restrict__ pstart, const int* __restrict__ ppw, const int* __restrict__ ppos,
    const int* __restrict__ ystart, const short* __restrict__ ycols, int ny, int nz, int nplanes,
    FftAxis ay, FftAxis az, const typename C2T<T>::type* __restrict__ tw) {
    using C = typename C2T<T>::type;
    extern __shared__ unsigned char smem_raw[];
    const int pitch = ny + 1;
    C* buf = reinterpret_cast<C*>(smem_raw);
    C* twy = buf + nz * pitch;
    C* twz = twy + ny;
    short* cols = reinterpret_cast<short*>(twz + nz);
    const int p = blockIdx.x;
    const size_t col = blockIdx.y;
    const int tid = threadIdx.x, nthr = blockDim.x;
    // The plane's first sphere entries are loaded before the tile is cleared, so the latency of
    // the two dependent loads (index, then coefficient) overlaps the clearing.
    const C* xc = x + col * ldx;
    const int j0 = pstart[p], j1 = pstart[p + 1];
    C pv[SpherePrefetch<T>::value];
    int pp[SpherePrefetch<T>::value];
#pragma unroll
    for (int u = 0; u < SpherePrefetch<T>::value; ++u) {
        const int j = j0 + tid + u * nthr;
        pp[u] = -1;
        if (j < j1) {
            pp[u] = ppos[j];
            pv[u] = xc[ppw[j]];
        }
    }
    for (int i = tid; i < nz * pitch; i += nthr) buf[i] = C{T(0), T(0)};
    load_twiddles(twy, tw + ay.tw_off, ny, tid, nthr);
    load_twiddles(twz, tw + az.tw_off, nz, tid, nthr);
    const int c0 = ystart[p], nact = ystart[p + 1] - c0;
    for (int i = tid; i < nact; i += nthr) cols[i] = ycols[c0 + i];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < SpherePrefetch<T>::value; ++u)
        if (pp[u] >= 0) buf[pp[u]] = pv[u];
    for (int j = j0 + tid + SpherePrefetch<T>::value * nthr; j < j1; j += nthr) buf[ppos[j]] = xc[ppw[j]];
    __syncthreads();
    // Inverse (e^{+i}) along z, only on the y-columns that hold plane waves.
    lds_fft_lines<T, C, true, 1>(buf, az, twz, LineSet{nact, pitch, 0, cols}, tid, nthr);
    // Inverse along y on every z-row.
    lds_fft_lines<T, C, true, 1>(buf, ay, twy, LineSet{nz, 1, pitch, nullptr}, tid, nthr);
    C* out = tbuf + (col * nplanes + p) * (size_t)ny * nz;
    const unsigned my = div_magic(ny);
    for (int i = tid; i < ny * nz; i += nthr) {
        const int z = fdiv(i, ny, my), y = i - z * ny;
        out[i] = buf[y + pitch * z];
    }
}

template <typename T>
__global__ __launch_bounds__(TileThreads<T>::value) void k_xline(
    typename C2T<T>::type* __restrict__ tbuf, const int* __restrict__ pxpos, int nplanes,
    const int* __restrict__ xabsent, int nabsent, const T* __restrict__ vloc, int nx, int ny, int nz,
    int ylen, FftAxis ax, const typename C2T<T>::type* __restrict__ tw) {
    // One workgroup: the x-lines of ylen consecutive y values (y0 = blockIdx.z * ylen) at one z of
    // one band; the x-transforms of different y are independent, so the slab is the tuning knob
    // between LDS per workgroup (occupancy) and the length of the contiguous row segments.
    using C = typename C2T<T>::type;
    extern __shared__ unsigned char smem_raw[];
    const int pitch = ylen + 1;
    C* buf = reinterpret_cast<C*>(smem_raw);  // [x][y]: element (x, y) at y + pitch*x
    C* twx = buf + nx * pitch;
    const int z = blockIdx.x;
    const size_t col = blockIdx.y;
    const int y0 = blockIdx.z * ylen;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const unsigned my = div_magic(ylen);
    // Rows of x values that hold no plane wave start (and stay) zero before the inverse.
    for (int i = tid; i < nabsent * ylen; i += nthr) {
        const int r = fdiv(i, ylen, my), y = i - r * ylen;
        buf[y + pitch * xabsent[r]] = C{T(0), T(0)};
    }
    load_twiddles(twx, tw + ax.tw_off, nx, tid, nthr);
    C* tcol = tbuf + col * (size_t)nplanes * ny * nz + y0;
    // Loads are issued kBatch at a time before their LDS stores, so each thread keeps several
    // global loads in flight (the kernel is latency-bound at one outstanding load per thread).
    constexpr int kBatch = LoadBatch<T>::value;
    const int ntot = nplanes * ylen;
    for (int base = 0; base < ntot; base += kBatch * nthr) {
        C v[kBatch];
        int dst[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int i = base + u * nthr + tid;
            const int ic = i < ntot ? i : 0;
            const int p = fdiv(ic, ylen, my), y = ic - p * ylen;
            v[u] = tcol[((size_t)p * nz + z) * ny + y];
            dst[u] = y + pitch * pxpos[p];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            if (base + u * nthr + tid < ntot) buf[dst[u]] = v[u];
    }
    __syncthreads();
    const LineSet xl{ylen, pitch, 1, nullptr};
    lds_fft_lines<T, C, true, 1>(buf, ax, twx, xl, tid, nthr);
    const T* vz = vloc + (size_t)nx * (ny * (size_t)z + y0);
    const unsigned mx = div_magic(nx);
    const int nv = nx * ylen;
    for (int base = 0; base < nv; base += kBatch * nthr) {
        T vv[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int i = base + u * nthr + tid;
            vv[u] = vz[i < nv ? i : 0];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int i = base + u * nthr + tid;
            if (i < nv) {
                const int y = fdiv(i, nx, mx), xx = i - y * nx;
                C& e = buf[y + pitch * xx];
                e = C{e.x * vv[u], e.y * vv[u]};
            }
        }
    }
    __syncthreads();
    lds_fft_lines<T, C, false, -1>(buf, ax, twx, xl, tid, nthr);
    for (int i = tid; i < nplanes * ylen; i += nthr) {
        const int p = fdiv(i, ylen, my), y = i - p * ylen;
        tcol[((size_t)p * nz + z) * ny + y] = buf[y + pitch * pxpos[p]];
    }
}

// Raw buffer loads and stores (gfx9 buffer resource: base, byte size, no stride): the row offset
// is a scalar operand and the per-thread offset one VGPR for every row, so no 64-bit address is
// formed per access; accesses at or past the size read zero and write nothing.
using BufU2 = unsigned int __attribute__((ext_vector_type(2)));
using BufU4 = unsigned int __attribute__((ext_vector_type(4)));
constexpr int kBufRsrcWord3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, kBufRsrcWord3);
}

template <typename X>
__device__ __forceinline__ X buf_load(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff) {
    if constexpr (sizeof(X) == 4) return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    else if constexpr (sizeof(X) == 8) return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    else return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

template <typename X>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, unsigned voff, int soff, X v) {
    if constexpr (sizeof(X) == 8) __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(BufU2, v), r, voff, soff, 0);
    else __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(BufU4, v), r, voff, soff, 0);
}

// Planes of the x-line pass of one workgroup: (first plane, plane count) of k-point blockIdx.z in a
// batched apply over several k-points (kpl != nullptr, the planes of all k-points concatenated per
// band), or (0, nplanes) for a single context.
__device__ __forceinline__ int2 xl_planes(const int2* __restrict__ kpl, int nplanes) {
    return kpl ? kpl[blockIdx.z] : make_int2(0, nplanes);
}

// x-line pass with one x-line per thread held in registers (nx fixed at compile time): thread
// l = z ny + y of band blockIdx.x loads its values of the np sphere planes (consecutive threads
// read consecutive addresses of each plane), runs the inverse x-transform, multiplies by V_loc/N
// from the x-major copy vt[x][z][y], runs the forward transform and stores the planes back.  No
// LDS and no barrier.  The planes are the cyclic run x = xs + p (mod NX), p < np, and the line is
// held rotated, a[u] = value at x = xs + u: the x-part of H is a convolution over the frequency
// index (a product in real space), so the shift passes through it unchanged and the result lands
// at the same rotated positions.  Row u is then plane u for u < np and zero above (never loaded),
// and the twiddles are compile-time constants.
template <typename T, int NX>
__global__ __launch_bounds__(256) void k_xline_reg(typename C2T<T>::type* __restrict__ tbuf, int nplanes, int lines,
                                                   const T* __restrict__ vt, const int2* __restrict__ kpl) {
    using C = typename C2T<T>::type;
    // Bands run fastest over the grid (blockIdx.x), so the workgroups that follow each other
    // read the same V_loc lines while they are still in cache.
    const unsigned l = blockIdx.y * blockDim.x + threadIdx.x;
    if (l >= (unsigned)lines) return;
    const int row = lines * (int)sizeof(C), vrow = lines * (int)sizeof(T);
    const int2 kp = xl_planes(kpl, nplanes);
    const __amdgpu_buffer_rsrc_t rt = buf_rsrc(tbuf + ((size_t)blockIdx.x * nplanes + kp.x) * lines, kp.y * row);
    const __amdgpu_buffer_rsrc_t rv = buf_rsrc(vt, NX * vrow);
    const unsigned vo = l * (unsigned)sizeof(C), vvo = l * (unsigned)sizeof(T);
    C a[NX];
#pragma unroll
    for (int u = 0; u < NX; ++u) {
        a[u] = C{T(0), T(0)};
        if (u < kp.y) a[u] = buf_load<C>(rt, vo, u * row);
    }
    freg::fft<NX, 1>(a);
#pragma unroll
    for (int x = 0; x < NX; ++x) {
        const T v = buf_load<T>(rv, vvo, x * vrow);
        a[x] = C{a[x].x * v, a[x].y * v};
    }
    freg::fft<NX, -1>(a);
    // Fresh copies of the scalars keep the compiler from holding the load side's row offsets and
    // comparisons live across both transforms.
    int np = kp.y, rs = row;
    asm volatile("" : "+s"(np), "+s"(rs));
#pragma unroll
    for (int u = 0; u < NX; ++u)
        if (u < np) buf_store<C>(rt, vo, u * rs, a[u]);
}

// Lane-pair variant for spheres that span at most half the line (np <= NX / 2, NX even): the
// two lanes 2l, 2l + 1 share x-line l and each runs transforms of length M = NX / 2 only.  With
// the rotated rows b[u] (u < M, zero above), the real-space values at r = 2m + t are
//   psi(2m + t) = sum_u (b[u] e^{+2 pi i u t / NX}) e^{+2 pi i u m / M},
// an inverse length-M transform of b twisted by e^{+2 pi i u t / NX}; lane t multiplies its values by
// V_loc(2m + t)/N, and the forward result at rows k < np is
//   X[k] = F_0[k] + e^{-2 pi i k / NX} F_1[k],  F_t = forward length-M transform of lane t,
// so lane 1 twists its transform, one DPP exchange adds it to lane 0's, and lane 0 stores.
// Both lanes load the same rows (one memory request per cache line).  Half the registers per lane
// of k_xline_reg: several waves per SIMD, and FP64 lines up to 100 fit.
template <typename X>
__device__ __forceinline__ X pair_partner(X v) {
    // quad_perm [1, 0, 3, 2]: every lane reads its pair partner's value.
    constexpr int kSwap = 1 | (0 << 2) | (3 << 4) | (2 << 6);
    if constexpr (sizeof(X) == 4) {
        return __builtin_bit_cast(X, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), kSwap, 0xf, 0xf, false));
    } else {
        const BufU2 w = __builtin_bit_cast(BufU2, v);
        BufU2 r;
        r.x = (unsigned)__builtin_amdgcn_mov_dpp((int)w.x, kSwap, 0xf, 0xf, false);
        r.y = (unsigned)__builtin_amdgcn_mov_dpp((int)w.y, kSwap, 0xf, 0xf, false);
        return __builtin_bit_cast(X, r);
    }
}

template <typename T, int NX>
__global__ __launch_bounds__(256) void k_xline_pair(typename C2T<T>::type* __restrict__ tbuf, int nplanes, int lines,
                                                    const T* __restrict__ vt, const int2* __restrict__ kpl) {
    using C = typename C2T<T>::type;
    constexpr int M = NX / 2;
    static_assert(NX % 2 == 0, "pair kernel needs an even line");
    // Pairs never straddle a wavefront and exit together (blockDim is even).
    const unsigned g = blockIdx.y * blockDim.x + threadIdx.x;
    const unsigned l = g >> 1;
    const bool odd = (g & 1u) != 0;
    if (l >= (unsigned)lines) return;
    const int row = lines * (int)sizeof(C), vrow = lines * (int)sizeof(T);
    const int2 kp = xl_planes(kpl, nplanes);
    const __amdgpu_buffer_rsrc_t rt = buf_rsrc(tbuf + ((size_t)blockIdx.x * nplanes + kp.x) * lines, kp.y * row);
    const __amdgpu_buffer_rsrc_t rv = buf_rsrc(vt, NX * vrow);
    const unsigned vo = l * (unsigned)sizeof(C);
    const unsigned vvo = l * (unsigned)sizeof(T) + (odd ? (unsigned)vrow : 0u);
    C a[M];
#pragma unroll
    for (int u = 0; u < M; ++u) {
        a[u] = C{T(0), T(0)};
        if (u < kp.y) a[u] = buf_load<C>(rt, vo, u * row);
    }
    // Lane 1: b[u] e^{+2 pi i u / NX} (the product with literal twiddles, then a per-lane select).
#pragma unroll
    for (int u = 1; u < M; ++u) {
        const T c = T(freg::kRoots<NX>.c[u]), sn = T(freg::kRoots<NX>.s[u]);
        const T px = a[u].x * c - a[u].y * sn, py = a[u].x * sn + a[u].y * c;
        a[u].x = odd ? px : a[u].x;
        a[u].y = odd ? py : a[u].y;
    }
    freg::fft<M, 1>(a);
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const T v = buf_load<T>(rv, vvo, 2 * m * vrow);
        a[m] = C{a[m].x * v, a[m].y * v};
    }
    freg::fft<M, -1>(a);
    // Lane 1: e^{-2 pi i k / NX} F_1[k]; lane 0 adds it to F_0[k].
#pragma unroll
    for (int k = 1; k < M; ++k) {
        const T c = T(freg::kRoots<NX>.c[k]), sn = T(freg::kRoots<NX>.s[k]);
        const T px = a[k].x * c + a[k].y * sn, py = a[k].y * c - a[k].x * sn;
        a[k].x = odd ? px : a[k].x;
        a[k].y = odd ? py : a[k].y;
    }
    int np = kp.y, rs = row;
    asm volatile("" : "+s"(np), "+s"(rs));
#pragma unroll
    for (int k = 0; k < M; ++k) {
        if (k < np) {
            const C o = C{pair_partner(a[k].x), pair_partner(a[k].y)};
            if (!odd) buf_store<C>(rt, vo, k * rs, C{a[k].x + o.x, a[k].y + o.y});
        }
    }
}

// Lane-triple variant for lines divisible by 3 (M = NX / 3), any sphere width: lanes 3l, 3l + 1,
// 3l + 2 of a wavefront share x-line l (21 lines per wavefront, its last lane idle).  With the
// rotated rows b[u] (u < np, zero above) and w = e^{2 pi i / NX}, the real-space values at
// r = 3m + t are an inverse length-M transform of the folded, twisted rows
//   c_t[u0] = sum_j b[u0 + j M] w^{(u0 + j M) t},   u0 < M,
// because w^{(u0 + jM) 3m} = e^{2 pi i u0 m / M}.  Lane t multiplies by V_loc(3m + t)/N and runs the
// forward length-M transform F_t; with H_t[k0] = w^{-k0 t} F_t[k0] the result is a 3-point
// transform across the lanes,
//   X[k0 + j M] = sum_t e^{-2 pi i j t / 3} H_t[k0],
// so lane j gathers its partners' H_t[k0] (ds_bpermute) and stores the rows k0 + j M < np.  A
// third of the registers per lane of k_xline_reg: odd lines (75 = 3 x 25) fit in FP64.
template <typename X>
__device__ __forceinline__ X lane_read(X v, int src_lane) {
    static_assert(sizeof(X) % 4 == 0, "32-bit words");
    constexpr int kWords = sizeof(X) / 4;
    struct Words {
        int w[kWords];
    };
    Words in = __builtin_bit_cast(Words, v), out;
#pragma unroll
    for (int i = 0; i < kWords; ++i) out.w[i] = __builtin_amdgcn_ds_bpermute(src_lane << 2, in.w[i]);
    return __builtin_bit_cast(X, out);
}

template <typename T, int NX>
__global__ __launch_bounds__(256) void k_xline_tri(typename C2T<T>::type* __restrict__ tbuf, int nplanes, int lines,
                                                   const T* __restrict__ vt, const int2* __restrict__ kpl) {
    using C = typename C2T<T>::type;
    static_assert(NX % 3 == 0, "lane-triple kernel needs a line divisible by 3");
    constexpr int M = NX / 3;
    constexpr int kLinesPerWave = 21;
    const int lane = (int)(threadIdx.x & 63);
    const int wave = (int)((blockIdx.y * blockDim.x + threadIdx.x) >> 6);
    // Every lane stays active to the end (the exchange reads partner registers); lanes past the last
    // line and the idle lane 63 work on a valid line and store nothing.
    const int t = lane < 63 ? lane % 3 : 0;
    const int base_lane = lane < 63 ? lane - t : 63;
    const int lraw = wave * kLinesPerWave + (lane < 63 ? lane / 3 : 0);
    const bool store_ok = lane < 63 && lraw < lines;
    const int l = lraw < lines ? lraw : lines - 1;
    const int row = lines * (int)sizeof(C), vrow = lines * (int)sizeof(T);
    const int2 kp = xl_planes(kpl, nplanes);
    const __amdgpu_buffer_rsrc_t rt = buf_rsrc(tbuf + ((size_t)blockIdx.x * nplanes + kp.x) * lines, kp.y * row);
    const __amdgpu_buffer_rsrc_t rv = buf_rsrc(vt, NX * vrow);
    const unsigned vo = (unsigned)l * (unsigned)sizeof(C);
    const unsigned vvo = (unsigned)l * (unsigned)sizeof(T) + (unsigned)(t * vrow);
    const bool t1 = t == 1, t2 = t == 2;
    C c[M];
#pragma unroll
    for (int u = 0; u < M; ++u) c[u] = C{T(0), T(0)};
    // Fold the present rows with the lane's twist w^{u t} (literal twiddles selected per lane).
#pragma unroll
    for (int u = 0; u < NX; ++u) {
        if (u < kp.y) {
            const C a = buf_load<C>(rt, vo, u * row);
            const int e1 = u % NX, e2 = (2 * u) % NX;
            const T wc = t1 ? T(freg::kRoots<NX>.c[e1]) : (t2 ? T(freg::kRoots<NX>.c[e2]) : T(1));
            const T ws = t1 ? T(freg::kRoots<NX>.s[e1]) : (t2 ? T(freg::kRoots<NX>.s[e2]) : T(0));
            C& d = c[u % M];
            d.x += a.x * wc - a.y * ws;
            d.y += a.x * ws + a.y * wc;
        }
    }
    freg::fft<M, 1>(c);
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const T v = buf_load<T>(rv, vvo, 3 * m * vrow);
        c[m] = C{c[m].x * v, c[m].y * v};
    }
    freg::fft<M, -1>(c);
    // H_t[k0] = w^{-k0 t} F_t[k0].
#pragma unroll
    for (int k = 1; k < M; ++k) {
        const int e1 = k % NX, e2 = (2 * k) % NX;
        const T wc = t1 ? T(freg::kRoots<NX>.c[e1]) : (t2 ? T(freg::kRoots<NX>.c[e2]) : T(1));
        const T ws = t1 ? T(freg::kRoots<NX>.s[e1]) : (t2 ? T(freg::kRoots<NX>.s[e2]) : T(0));
        const T px = c[k].x * wc + c[k].y * ws, py = c[k].y * wc - c[k].x * ws;
        c[k].x = px;
        c[k].y = py;
    }
    // Lane j holds H_j; its partners hold H_{j+1}, H_{j+2} (indices mod 3).  X[k0 + j M] =
    // H_0 + e^{-2 pi i j / 3} H_1 + e^{-4 pi i j / 3} H_2, written as own + f1 next + f2 after-next
    // with per-lane factors: j = 0: 1, 1, 1; j = 1: w3^-1, w3^+1, 1; j = 2: w3^-1, 1, w3^+1
    // (w3 = e^{2 pi i / 3}).
    const int p1 = base_lane + (t + 1) % 3, p2 = base_lane + (t + 2) % 3;
    constexpr T kC3 = T(-0.5), kS3 = T(0.86602540378443864676);
    const T own_c = t == 0 ? T(1) : kC3, own_s = t == 0 ? T(0) : -kS3;
    const T n1_c = t1 ? kC3 : T(1), n1_s = t1 ? kS3 : T(0);
    const T n2_c = t2 ? kC3 : T(1), n2_s = t2 ? kS3 : T(0);
    int np = kp.y, rs = row;
    asm volatile("" : "+s"(np), "+s"(rs));
#pragma unroll
    for (int k = 0; k < M; ++k) {
        if (k < np) {
            const C h1 = lane_read(c[k], p1), h2 = lane_read(c[k], p2);
            const C o = c[k];
            const T xr = o.x * own_c - o.y * own_s + h1.x * n1_c - h1.y * n1_s + h2.x * n2_c - h2.y * n2_s;
            const T xi = o.x * own_s + o.y * own_c + h1.x * n1_s + h1.y * n1_c + h2.x * n2_s + h2.y * n2_c;
            const int kk = k + t * M;
            if (store_ok && kk < np) buf_store<C>(rt, vo, kk * rs, C{xr, xi});
        }
    }
}

// Residue-sequential variant for lines divisible by 3 (M = NX / 3) whose sphere spans at most two
// thirds of the line (np <= 2M): one thread per x-line, like k_xline_reg, but the line is never
// held whole.  With the rotated rows b[u], w = e^{2 pi i / NX} and x = 3m + t, the three residues t
// are done one after another, so every twiddle is a compile-time constant:
//   c_t[u0] = sum_j b[u0 + jM] w^{(u0 + jM) t}        (u0 < M; fold of the present rows)
//   psi(3m + t) = inverse length-M transform of c_t;  times V_loc(3m + t)/N
//   H_t[k0] = w^{-k0 t} F_t[k0],  F_t = forward length-M transform
//   X[k0 + jM] = sum_t e^{-2 pi i j t / 3} H_t[k0]     (j = 0, 1: rows below np <= 2M)
// X[k0] accumulates in registers, X[k0 + M] (k0 < np - M) in this thread's LDS column
// acc[k0][threadIdx.x].  The rows are re-read for each residue (cache hits after the first).
// Registers: two length-M arrays, so FP64 lines of 75 fit at two waves per SIMD.
template <typename T, int NX>
__global__ __launch_bounds__(256) void k_xline_seq3(typename C2T<T>::type* __restrict__ tbuf, int nplanes, int lines,
                                                    const T* __restrict__ vt, const int2* __restrict__ kpl) {
    using C = typename C2T<T>::type;
    static_assert(NX % 3 == 0, "residue-sequential kernel needs a line divisible by 3");
    constexpr int M = NX / 3;
    extern __shared__ unsigned char smem_raw[];
    C* acc = reinterpret_cast<C*>(smem_raw);
    const int nt = (int)blockDim.x, tid = (int)threadIdx.x;
    const unsigned l = blockIdx.y * blockDim.x + threadIdx.x;
    if (l >= (unsigned)lines) return;  // no barrier below: each thread uses only its own LDS column
    const int row = lines * (int)sizeof(C), vrow = lines * (int)sizeof(T);
    const int2 kp = xl_planes(kpl, nplanes);
    const __amdgpu_buffer_rsrc_t rt = buf_rsrc(tbuf + ((size_t)blockIdx.x * nplanes + kp.x) * lines, kp.y * row);
    const __amdgpu_buffer_rsrc_t rv = buf_rsrc(vt, NX * vrow);
    const unsigned vo = l * (unsigned)sizeof(C), vvo = l * (unsigned)sizeof(T);
    const int n1 = kp.y - M;  // rows k0 + M present for k0 < n1 (n1 <= M)
    constexpr T kC3 = T(-0.5), kS3 = T(0.86602540378443864676);
    C x0[M];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        C c[M];
#pragma unroll
        for (int u = 0; u < M; ++u) c[u] = C{T(0), T(0)};
        // A fresh copy of the offset per residue, tied to the previous residue's result: the rows
        // are loaded again after that residue is done instead of being kept live in registers.
        unsigned vot = vo;
        if (t == 0) asm volatile("" : "+v"(vot));
        else asm volatile("" : "+v"(vot) : "v"(x0[M - 1].x));
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            if (u < kp.y) {
                const C b = buf_load<C>(rt, vot, u * row);
                C& d = c[u % M];
                if (t == 0 || (u * t) % NX == 0) {
                    d.x += b.x;
                    d.y += b.y;
                } else {
                    const T wc = T(freg::kRoots<NX>.c[(u * t) % NX]), ws = T(freg::kRoots<NX>.s[(u * t) % NX]);
                    d.x += b.x * wc - b.y * ws;
                    d.y += b.x * ws + b.y * wc;
                }
            }
        }
        freg::fft<M, 1>(c);
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const T v = buf_load<T>(rv, vvo, (3 * m + t) * vrow);
            c[m] = C{c[m].x * v, c[m].y * v};
        }
        freg::fft<M, -1>(c);
        if (t > 0) {
#pragma unroll
            for (int k = 1; k < M; ++k) {
                const T wc = T(freg::kRoots<NX>.c[(k * t) % NX]), ws = T(freg::kRoots<NX>.s[(k * t) % NX]);
                const T px = c[k].x * wc + c[k].y * ws, py = c[k].y * wc - c[k].x * ws;
                c[k].x = px;
                c[k].y = py;
            }
        }
        // j = 1 coefficient e^{-2 pi i t / 3}: 1, (-1/2, -sqrt3/2), (-1/2, +sqrt3/2).
        const T fc = t == 0 ? T(1) : kC3, fs = t == 0 ? T(0) : (t == 1 ? -kS3 : kS3);
#pragma unroll
        for (int k = 0; k < M; ++k) {
            if (t == 0) x0[k] = c[k];
            else {
                x0[k].x += c[k].x;
                x0[k].y += c[k].y;
            }
            if (k < n1) {
                C& a1 = acc[(size_t)k * nt + tid];
                const C h = C{c[k].x * fc - c[k].y * fs, c[k].x * fs + c[k].y * fc};
                if (t == 0) a1 = h;
                else a1 = C{a1.x + h.x, a1.y + h.y};
            }
        }
    }
    int np = kp.y, rs = row;
    asm volatile("" : "+s"(np), "+s"(rs));
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (k < np) buf_store<C>(rt, vo, k * rs, x0[k]);
#pragma unroll
    for (int k = 0; k < M; ++k)
        if (k < n1) buf_store<C>(rt, vo, (k + M) * rs, acc[(size_t)k * nt + tid]);
}

// Line lengths the register x-line kernel is built for: FP32 from 8 to 100, FP64 (twice the
// registers per line) up to 50, beyond which the line no longer fits the register file.
#define DZ_XREG_SIZES(X) X(8) X(12) X(15) X(16) X(18) X(20) X(24) X(25) X(27) X(30) X(32) X(36) X(40) \
    X(45) X(48) X(50) X(54) X(60) X(64) X(72) X(75) X(80) X(81) X(90) X(96) X(100)
#define DZ_XREG64_SIZES(X) X(8) X(12) X(15) X(16) X(18) X(20) X(24) X(25) X(27) X(30) X(32) X(36) X(40) \
    X(45) X(48) X(50)

// Even lines whose half length the register FFT takes, for the lane-pair kernel (both precisions).
#define DZ_XPAIR_SIZES(X) X(8) X(12) X(16) X(18) X(20) X(24) X(30) X(32) X(36) X(40) X(48) X(50) X(54) X(60) \
    X(64) X(72) X(80) X(90) X(96) X(100)

// Lines divisible by 3 whose third the register FFT takes, for the lane-triple kernel (both precisions).
#define DZ_XTRI_SIZES(X) X(12) X(15) X(18) X(24) X(27) X(30) X(36) X(45) X(48) X(54) X(60) X(72) X(75) X(81) \
    X(90) X(96)

bool xline_tri_supported(int nx) {
    switch (nx) {
#define DZ_XREG_CASE(N) case N:
        DZ_XTRI_SIZES(DZ_XREG_CASE)
#undef DZ_XREG_CASE
        return true;
        default:
            return false;
    }
}

bool xline_pair_supported(int nx) {
    switch (nx) {
#define DZ_XREG_CASE(N) case N:
        DZ_XPAIR_SIZES(DZ_XREG_CASE)
#undef DZ_XREG_CASE
        return true;
        default:
            return false;
    }
}

bool xline_reg_supported(int nx, bool fp64) {
    switch (nx) {
#define DZ_XREG_CASE(N) case N:
        DZ_XREG_SIZES(DZ_XREG_CASE)
#undef DZ_XREG_CASE
        break;
        default:
            return false;
    }
    if (!fp64) return true;
    switch (nx) {
#define DZ_XREG_CASE(N) case N:
        DZ_XREG64_SIZES(DZ_XREG_CASE)
        return true;
#undef DZ_XREG_CASE
        default:
            return false;
    }
}

bool xline_reg_usable(const Ham& h, bool fp64) {
    // The kernel's buffer resources hold a band's planes and V_loc in 32-bit byte ranges.
    return xline_reg_supported(h.nx, fp64) && h.plane_xs >= 0 && h.ngrid * 2 * sizeof(double) < ((size_t)1 << 31);
}

bool xline_pair_usable(const Ham& h) {
    return xline_pair_supported(h.nx) && h.plane_xs >= 0 && 2 * h.nplanes <= h.nx &&
           h.ngrid * 2 * sizeof(double) < ((size_t)1 << 31);
}

bool xline_tri_usable(const Ham& h) {
    return xline_tri_supported(h.nx) && h.plane_xs >= 0 && h.ngrid * 2 * sizeof(double) < ((size_t)1 << 31);
}

bool xline_seq3_usable(const Ham& h) {
    return xline_tri_supported(h.nx) && h.plane_xs >= 0 && h.nplanes <= 2 * (h.nx / 3) &&
           h.ngrid * 2 * sizeof(double) < ((size_t)1 << 31);
}

bool xline_code_usable(const Ham& h, int code, bool fp64) {
    if (code > 0) return h.ny % code == 0;
    if (code == kXlineConv) return xline_conv_usable(h);
    if (code < kXlineRegLast || code > kXlineReg) return false;
    switch (xline_reg_kind(code)) {
        case kXlineOneLane: return xline_reg_usable(h, fp64);
        case kXlinePair: return xline_pair_usable(h);
        case kXlineTriple: return xline_tri_usable(h);
        default: return xline_seq3_usable(h);
    }
}

template <typename T>
// Register x-line kernel `kind` on h's grid.  A batched apply over nk k-points passes kpl (per
// k-point: first plane and plane count within each band's planes), the planes per band and the
// largest per-k plane count; a single context passes kpl = nullptr, nk = 1 and its own counts.
void launch_xline_reg(Ham& h, typename C2T<T>::type* tb, size_t nc, int threads, int kind, hipStream_t s,
                      const int2* kpl = nullptr, int nk = 1, int planes_per_band = -1, int max_planes = -1) {
    const int lines = h.ny * h.nz;
    const int nplanes = planes_per_band >= 0 ? planes_per_band : h.nplanes;
    const int np_max = max_planes >= 0 ? max_planes : h.nplanes;
    const int nlanes = kind == kXlineSeq3 ? 1 : kind + 1;
    const bool pair = nlanes == 2;
    // The triple kernel puts 21 lines on each wavefront (63 of its 64 lanes).
    const int waves_per_group = threads / 64;
    const unsigned groups = nlanes == 3 ? (unsigned)((lines + 21 * waves_per_group - 1) / (21 * waves_per_group))
                                        : (unsigned)((nlanes * lines + threads - 1) / threads);
    const dim3 grid((unsigned)nc, groups, (unsigned)nk);
    const T* vt = std::is_same<T, float>::value ? (const T*)h.vloc32_t.ptr : (const T*)h.vloc64_t.ptr;
    bool done = false;
    if (kind == kXlineSeq3) {
        const size_t lds = (size_t)std::max(0, np_max - h.nx / 3) * threads * sizeof(typename C2T<T>::type);
        switch (h.nx) {
#define DZ_XREG_CASE(N)                                                                                      \
    case N:                                                                                                  \
        hipLaunchKernelGGL((k_xline_seq3<T, N>), grid, dim3(threads), lds, s, tb, nplanes, lines, vt, kpl);    \
        done = true;                                                                                         \
        break;
            DZ_XTRI_SIZES(DZ_XREG_CASE)
#undef DZ_XREG_CASE
            default:
                break;
        }
    } else if (nlanes == 3) {
        switch (h.nx) {
#define DZ_XREG_CASE(N)                                                                                      \
    case N:                                                                                                  \
        hipLaunchKernelGGL((k_xline_tri<T, N>), grid, dim3(threads), 0, s, tb, nplanes, lines, vt, kpl);       \
        done = true;                                                                                         \
        break;
            DZ_XTRI_SIZES(DZ_XREG_CASE)
#undef DZ_XREG_CASE
            default:
                break;
        }
    } else if (pair) {
        switch (h.nx) {
#define DZ_XREG_CASE(N)                                                                                      \
    case N:                                                                                                  \
        hipLaunchKernelGGL((k_xline_pair<T, N>), grid, dim3(threads), 0, s, tb, nplanes, lines, vt, kpl);      \
        done = true;                                                                                         \
        break;
            DZ_XPAIR_SIZES(DZ_XREG_CASE)
#undef DZ_XREG_CASE
            default:
                break;
        }
    } else if constexpr (std::is_same<T, float>::value) {
        switch (h.nx) {
#define DZ_XREG_CASE(N)                                                                                      \
    case N:                                                                                                  \
        hipLaunchKernelGGL((k_xline_reg<T, N>), grid, dim3(threads), 0, s, tb, nplanes, lines, vt, kpl);       \
        done = true;                                                                                         \
        break;
            DZ_XREG_SIZES(DZ_XREG_CASE)
#undef DZ_XREG_CASE
            default:
                break;
        }
    } else {
        switch (h.nx) {
#define DZ_XREG_CASE(N)                                                                                      \
    case N:                                                                                                  \
        hipLaunchKernelGGL((k_xline_reg<T, N>), grid, dim3(threads), 0, s, tb, nplanes, lines, vt, kpl);       \
        done = true;                                                                                         \
        break;
            DZ_XREG64_SIZES(DZ_XREG_CASE)
#undef DZ_XREG_CASE
            default:
                break;
        }
    }
    DZ_REQUIRE(done, DZHIP_ERR_STATE, "no register x-line kernel for this nx");
    DZ_HIP(hipGetLastError());
}

template <typename T>
__global__ __launch_bounds__(TileThreads<T>::value) void k_plane_fwd(
    const typename C2T<T>::type* __restrict__ tbuf, const typename C2T<T>::type* __restrict__ x,
    size_t ldx, typename C2T<T>::type* __restrict__ y, size_t ldy, const T* __restrict__ kinetic,
    const int* __restrict__ pstart, const int* __restrict__ ppw, const int* __restrict__ ppos,
    const int* __restrict__ ystart, const short* __restrict__ ycols, int ny, int nz, int nplanes,
    FftAxis ay, FftAxis az, const typename C2T<T>::type* __restrict__ tw, int accumulate) {
    using C = typename C2T<T>::type;
    extern __shared__ unsigned char smem_raw[];
    const int pitch = ny + 1;
    C* buf = reinterpret_cast<C*>(smem_raw);
    C* twy = buf + nz * pitch;
    C* twz = twy + ny;
    short* cols = reinterpret_cast<short*>(twz + nz);
    const int p = blockIdx.x;
    const size_t col = blockIdx.y;
    const int tid = threadIdx.x, nthr = blockDim.x;
    // The gather's operands (index, LDS position, kinetic factor, x and the nonlocal y) do not
    // depend on the transforms: the plane's first entries are loaded before the tile, so their
    // latency overlaps the tile load and the transforms.
    const C* xc = x + col * ldx;
    C* yc = y + col * ldy;
    const int j0 = pstart[p], j1 = pstart[p + 1];
    int gg[SpherePrefetch<T>::value], gp[SpherePrefetch<T>::value];
    T gk[SpherePrefetch<T>::value];
    C gx[SpherePrefetch<T>::value], gy[SpherePrefetch<T>::value];
#pragma unroll
    for (int u = 0; u < SpherePrefetch<T>::value; ++u) {
        const int j = j0 + tid + u * nthr;
        gg[u] = -1;
        if (j < j1) {
            const int g = ppw[j];
            gg[u] = g;
            gp[u] = ppos[j];
            gk[u] = kinetic[g];
            gx[u] = xc[g];
            gy[u] = accumulate ? yc[g] : C{T(0), T(0)};
        }
    }
    load_twiddles(twy, tw + ay.tw_off, ny, tid, nthr);
    load_twiddles(twz, tw + az.tw_off, nz, tid, nthr);
    const int c0 = ystart[p], nact = ystart[p + 1] - c0;
    for (int i = tid; i < nact; i += nthr) cols[i] = ycols[c0 + i];
    const C* in = tbuf + (col * nplanes + p) * (size_t)ny * nz;
    const unsigned my = div_magic(ny);
    // Tile load in batches of loads per thread (several in flight before the LDS stores).
    constexpr int kBatch = LoadBatch<T>::value;
    const int ntot = ny * nz;
    for (int base = 0; base < ntot; base += kBatch * nthr) {
        C v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int i = base + u * nthr + tid;
            v[u] = in[i < ntot ? i : 0];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int i = base + u * nthr + tid;
            if (i < ntot) {
                const int z = fdiv(i, ny, my), yy = i - z * ny;
                buf[yy + pitch * z] = v[u];
            }
        }
    }
    __syncthreads();
    // Forward along y on every z-row (frequency y lands at py[y]), then along z only on the
    // columns whose frequencies are gathered.
    lds_fft_lines<T, C, false, -1>(buf, ay, twy, LineSet{nz, 1, pitch, nullptr}, tid, nthr);
    lds_fft_lines<T, C, false, -1>(buf, az, twz, LineSet{nact, pitch, 0, cols}, tid, nthr);
#pragma unroll
    for (int u = 0; u < SpherePrefetch<T>::value; ++u)
        if (gg[u] >= 0) {
            const C v = buf[gp[u]];
            yc[gg[u]] = C{v.x + gk[u] * gx[u].x + gy[u].x, v.y + gk[u] * gx[u].y + gy[u].y};
        }
    for (int j = j0 + tid + SpherePrefetch<T>::value * nthr; j < j1; j += nthr) {
        const int g = ppw[j];
        const C v = buf[ppos[j]];
        const T kin = kinetic[g];
        const C xv = xc[g];
        C out = C{v.x + kin * xv.x, v.y + kin * xv.y};
        if (accumulate) {
            const C prev = yc[g];
            out = C{out.x + prev.x, out.y + prev.y};
        }
        yc[g] = out;
    }
}

// Plane kernels with the z-transform in registers.  The LDS holds only the rows of z-frequencies
// that hold plane waves in the plane (its "active rows": about half of nz for a sphere in a grid of
// twice its diameter), each a y-line of odd pitch ny | 1, so an FP64 75 x 75 plane needs 43 KB of LDS
// instead of the 91 KB full tile and three workgroups share a CU.  The z-transform over all ny
// y-lines runs in registers with three lanes per line (21 lines per wavefront; NZ = 3M with
// M = 2^a 3^b 5^c): lane t holds the residue z = 3u + t, u < M, read from or written to the [z][y]
// plane in global memory (consecutive lanes, consecutive y).  With w = e^{2 pi i / NZ}:
//   inverse:  c_t[k0] = w^{t k0} sum_j Y[k0 + jM] e^{+2 pi i j t / 3}  (the three lanes of a line
//             read the Y they need from the active rows), psi(3u + t) = inverse length-M
//             transform of c_t;
//   forward:  F_t = forward length-M transform of psi(3u + t), H_t[k0] = w^{-t k0} F_t[k0],
//             X[k0 + jM] = sum_t e^{-2 pi i j t / 3} H_t[k0] (lane j gathers H_{j+1}, H_{j+2} from
//             its partners and writes X to its active row).
// The y-transforms run on the active rows only, in LDS, with the plane waves at the digit-reversed
// y positions of the full-tile kernels.
// Three wavefronts per SIMD (at most 168 VGPRs): with the LDS of a 75^3 or 90^3 plane's active
// rows, two or three workgroups of four or five wavefronts then fit a CU.
constexpr int kZrLinesPerWave = 21;
constexpr int kZrMaxThreads = 384;  // ny <= 126
// Sphere entries per thread whose loads are issued together, ahead of the transforms (a 75^3
// plane holds about 670 entries for 256 threads, so one round of loads instead of one latency
// per entry).
template <typename T> struct ZrPrefetch { static constexpr int value = 3; };

__host__ __device__ inline int plane_zr_threads(int ny) {
    return 64 * ((ny + kZrLinesPerWave - 1) / kZrLinesPerWave);
}
// Row pitch of the active rows, odd in elements: the y-stages' threads take consecutive rows of
// one butterfly, and an odd pitch spreads their 16-byte (FP64) accesses over distinct banks.
__host__ __device__ inline int zr_pitch(int ny) { return ny | 1; }
template <typename C>
__host__ __device__ inline size_t plane_zr_lds_bytes(int ny, int nz, int nrow_max) {
    return ((size_t)nrow_max * zr_pitch(ny) + ny) * sizeof(C) + round_up((size_t)nz * sizeof(short), 16);
}

template <typename T, int NZ>
__global__ __launch_bounds__(kZrMaxThreads, 3) void k_plane_inv_zr(
    const typename C2T<T>::type* __restrict__ x, size_t ldx, typename C2T<T>::type* __restrict__ tbuf,
    const int* __restrict__ pstart, const int* __restrict__ ppw, const int* __restrict__ ppw0,
    const int* __restrict__ pposr, const int* __restrict__ zstart, const short* __restrict__ zrows, int ny,
    int nplanes, FftAxis ay, const typename C2T<T>::type* __restrict__ tw) {
    using C = typename C2T<T>::type;
    static_assert(NZ % 3 == 0, "z-register plane kernel needs nz divisible by 3");
    constexpr int M = NZ / 3;
    extern __shared__ unsigned char smem_raw[];
    const int pitch = zr_pitch(ny);
    const int p = blockIdx.x;
    const size_t col = blockIdx.y;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int r0 = zstart[p], nrow = zstart[p + 1] - r0;
    C* buf = reinterpret_cast<C*>(smem_raw);
    C* twy = buf + nrow * pitch;
    short* zb = reinterpret_cast<short*>(twy + ny);  // active row of z-frequency k, or -1
    const C* xc = x + col * ldx;
    const int j0 = pstart[p], j1 = pstart[p + 1];
    // A plane whose entries are one run of the basis (h-major order) is read without the index.
    const int pw0 = ppw0[p];
    // Every plane holds at least one entry, so j0 is a valid index to read in place of a missing one.
    C pv[ZrPrefetch<T>::value];
    int pp[ZrPrefetch<T>::value];
#pragma unroll
    for (int u = 0; u < ZrPrefetch<T>::value; ++u) {
        const int j = j0 + tid + u * nthr;
        const bool in = j < j1;
        const int jc = in ? j : j0;
        const int q = pposr[jc];
        const C v = xc[pw0 >= 0 ? pw0 + (jc - j0) : ppw[jc]];
        pp[u] = in ? q : -1;
        pv[u] = v;
    }
    for (int i = tid; i < nrow * pitch; i += nthr) buf[i] = C{T(0), T(0)};
    load_twiddles(twy, tw + ay.tw_off, ny, tid, nthr);
    for (int k = tid; k < NZ; k += nthr) zb[k] = -1;
    __syncthreads();
    for (int i = tid; i < nrow; i += nthr) zb[zrows[r0 + i]] = (short)i;
#pragma unroll
    for (int u = 0; u < ZrPrefetch<T>::value; ++u)
        if (pp[u] >= 0) buf[pp[u]] = pv[u];
    for (int j = j0 + tid + ZrPrefetch<T>::value * nthr; j < j1; j += nthr)
        buf[pposr[j]] = xc[pw0 >= 0 ? pw0 + (j - j0) : ppw[j]];
    __syncthreads();
    // Inverse along y on the active rows (ends with a barrier).
    lds_fft_lines<T, C, true, 1>(buf, ay, twy, LineSet{nrow, 1, pitch, nullptr}, tid, nthr);
    const int lane = tid & 63, wave = tid >> 6;
    const int yl = wave * kZrLinesPerWave + lane / 3;
    if (lane == 63 || yl >= ny) return;  // no barrier below
    const int t = lane % 3;
    const bool t1 = t == 1, t2 = t == 2;
    constexpr T kC3 = T(-0.5), kS3 = T(0.86602540378443864676);
    // e^{+2 pi i t / 3} = (fc, fs) multiplies Y[k0 + M]; its conjugate multiplies Y[k0 + 2M].
    const T fc = t == 0 ? T(1) : kC3, fs = t1 ? kS3 : (t2 ? -kS3 : T(0));
    C c[M];
#pragma unroll
    for (int k0 = 0; k0 < M; ++k0) {
        C sum = C{T(0), T(0)};
        const int b0 = zb[k0], b1 = zb[k0 + M], b2 = zb[k0 + 2 * M];
        if (b0 >= 0) sum = buf[b0 * pitch + yl];
        if (b1 >= 0) {
            const C v = buf[b1 * pitch + yl];
            sum.x += v.x * fc - v.y * fs;
            sum.y += v.x * fs + v.y * fc;
        }
        if (b2 >= 0) {
            const C v = buf[b2 * pitch + yl];
            sum.x += v.x * fc + v.y * fs;
            sum.y += v.y * fc - v.x * fs;
        }
        const int e1 = k0 % NZ, e2 = (2 * k0) % NZ;
        const T wc = t1 ? T(freg::kRoots<NZ>.c[e1]) : (t2 ? T(freg::kRoots<NZ>.c[e2]) : T(1));
        const T ws = t1 ? T(freg::kRoots<NZ>.s[e1]) : (t2 ? T(freg::kRoots<NZ>.s[e2]) : T(0));
        c[k0] = C{sum.x * wc - sum.y * ws, sum.x * ws + sum.y * wc};
    }
    freg::fft<M, 1>(c);
    C* out = tbuf + (col * nplanes + p) * (size_t)ny * NZ;
#pragma unroll
    for (int u = 0; u < M; ++u) out[(3 * u + t) * ny + yl] = c[u];
}

template <typename T, int NZ>
__global__ __launch_bounds__(kZrMaxThreads, 3) void k_plane_fwd_zr(
    const typename C2T<T>::type* __restrict__ tbuf, const typename C2T<T>::type* __restrict__ x, size_t ldx,
    typename C2T<T>::type* __restrict__ y, size_t ldy, const T* __restrict__ kinetic,
    const int* __restrict__ pstart, const int* __restrict__ ppw, const int* __restrict__ ppw0,
    const int* __restrict__ pposr, const int* __restrict__ zstart, const short* __restrict__ zrows, int ny,
    int nplanes, FftAxis ay, const typename C2T<T>::type* __restrict__ tw, int accumulate) {
    using C = typename C2T<T>::type;
    static_assert(NZ % 3 == 0, "z-register plane kernel needs nz divisible by 3");
    constexpr int M = NZ / 3;
    extern __shared__ unsigned char smem_raw[];
    const int pitch = zr_pitch(ny);
    const int p = blockIdx.x;
    const size_t col = blockIdx.y;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int r0 = zstart[p], nrow = zstart[p + 1] - r0;
    C* buf = reinterpret_cast<C*>(smem_raw);
    C* twy = buf + nrow * pitch;
    short* zb = reinterpret_cast<short*>(twy + ny);
    const C* xc = x + col * ldx;
    C* yc = y + col * ldy;
    const int j0 = pstart[p], j1 = pstart[p + 1];
    // Register lines: every lane takes part in the exchange; lanes past the last line and the idle
    // lane 63 work on a valid line and write nothing.
    const int lane = tid & 63, wave = tid >> 6;
    const int t = lane < 63 ? lane % 3 : 0;
    const int base_lane = lane < 63 ? lane - t : 63;
    const int yraw = wave * kZrLinesPerWave + (lane < 63 ? lane / 3 : 0);
    const bool own = lane < 63 && yraw < ny;
    const int yl = yraw < ny ? yraw : ny - 1;
    const C* in = tbuf + (col * nplanes + p) * (size_t)ny * NZ;
    C a[M];
#pragma unroll
    for (int u = 0; u < M; ++u) a[u] = in[(3 * u + t) * ny + yl];
    load_twiddles(twy, tw + ay.tw_off, ny, tid, nthr);
    for (int k = tid; k < NZ; k += nthr) zb[k] = -1;
    __syncthreads();
    for (int i = tid; i < nrow; i += nthr) zb[zrows[r0 + i]] = (short)i;
    __syncthreads();
    freg::fft<M, -1>(a);
    const bool t1 = t == 1, t2 = t == 2;
#pragma unroll
    for (int k = 1; k < M; ++k) {
        const int e1 = k % NZ, e2 = (2 * k) % NZ;
        const T wc = t1 ? T(freg::kRoots<NZ>.c[e1]) : (t2 ? T(freg::kRoots<NZ>.c[e2]) : T(1));
        const T ws = t1 ? T(freg::kRoots<NZ>.s[e1]) : (t2 ? T(freg::kRoots<NZ>.s[e2]) : T(0));
        const T px = a[k].x * wc + a[k].y * ws, py = a[k].y * wc - a[k].x * ws;
        a[k].x = px;
        a[k].y = py;
    }
    // The transform is finished here: otherwise the compiler sinks its last stage into the
    // conditional blocks below and keeps the earlier stages' values live across all of them.
#pragma unroll
    for (int k = 0; k < M; ++k) asm volatile("" : "+v"(a[k].x), "+v"(a[k].y));
    // Lane j holds H_j, its partners H_{j+1} and H_{j+2} (indices mod 3):
    // X[k0 + jM] = own f0 + next f1 + after-next f2 with j = 0: 1, 1, 1; j = 1: w3^-1, w3, 1;
    // j = 2: w3^-1, 1, w3 (w3 = e^{2 pi i / 3}).
    const int p1 = base_lane + (t + 1) % 3, p2 = base_lane + (t + 2) % 3;
    constexpr T kC3 = T(-0.5), kS3 = T(0.86602540378443864676);
    const T own_c = t == 0 ? T(1) : kC3, own_s = t == 0 ? T(0) : -kS3;
    const T n1_c = t1 ? kC3 : T(1), n1_s = t1 ? kS3 : T(0);
    const T n2_c = t2 ? kC3 : T(1), n2_s = t2 ? kS3 : T(0);
#pragma unroll
    for (int k0 = 0; k0 < M; ++k0) {
        const int b0 = zb[k0], b1 = zb[k0 + M], b2 = zb[k0 + 2 * M];
        if ((b0 & b1 & b2) < 0) continue;  // none of the three frequencies is active (uniform)
        const C h1 = lane_read(a[k0], p1), h2 = lane_read(a[k0], p2);
        const C o = a[k0];
        const T xr = o.x * own_c - o.y * own_s + h1.x * n1_c - h1.y * n1_s + h2.x * n2_c - h2.y * n2_s;
        const T xi = o.x * own_s + o.y * own_c + h1.x * n1_s + h1.y * n1_c + h2.x * n2_s + h2.y * n2_c;
        const int bm = t == 0 ? b0 : (t1 ? b1 : b2);
        if (own && bm >= 0) buf[bm * pitch + yl] = C{xr, xi};
    }
    const int pw0 = ppw0[p];
    int gg[ZrPrefetch<T>::value], gp[ZrPrefetch<T>::value];
    T gk[ZrPrefetch<T>::value];
    C gx[ZrPrefetch<T>::value], gy[ZrPrefetch<T>::value];
#pragma unroll
    for (int u = 0; u < ZrPrefetch<T>::value; ++u) {
        const int j = j0 + tid + u * nthr;
        gg[u] = -1;
        if (j < j1) {
            const int g = pw0 >= 0 ? pw0 + (j - j0) : ppw[j];
            gg[u] = g;
            gp[u] = pposr[j];
            gk[u] = kinetic[g];
            gx[u] = xc[g];
            gy[u] = accumulate ? yc[g] : C{T(0), T(0)};
        }
    }
    __syncthreads();
    // Forward along y on the active rows; frequency y lands at its digit-reversed position.
    lds_fft_lines<T, C, false, -1>(buf, ay, twy, LineSet{nrow, 1, pitch, nullptr}, tid, nthr);
#pragma unroll
    for (int u = 0; u < ZrPrefetch<T>::value; ++u)
        if (gg[u] >= 0) {
            const C v = buf[gp[u]];
            yc[gg[u]] = C{v.x + gk[u] * gx[u].x + gy[u].x, v.y + gk[u] * gx[u].y + gy[u].y};
        }
    for (int j = j0 + tid + ZrPrefetch<T>::value * nthr; j < j1; j += nthr) {
        const int g = pw0 >= 0 ? pw0 + (j - j0) : ppw[j];
        const C v = buf[pposr[j]];
        const T kin = kinetic[g];
        const C xv = xc[g];
        C o = C{v.x + kin * xv.x, v.y + kin * xv.y};
        if (accumulate) {
            const C prev = yc[g];
            o = C{o.x + prev.x, o.y + prev.y};
        }
        yc[g] = o;
    }
}

// Density: one workgroup per (z, band group).  For each band of the group, the x-lines of the
// z-plane are gathered from T (sphere planes after the z/y inverse transforms), inverse
// transformed along x, and w_b |psi_b(r)|^2 is accumulated in FP64 registers; each thread owns
// fixed grid points of the plane.  Groups are summed by k_density_sum (no atomics).
constexpr int kDensThreads = 512;
constexpr int kDensMaxPts = 32;  // nx*ny <= kDensThreads * kDensMaxPts

__global__ __launch_bounds__(kDensThreads) void k_xline_density(
    const double2* __restrict__ tbuf, const int* __restrict__ pxpos, int nplanes,
    const int* __restrict__ xabsent, int nabsent,
    const double* __restrict__ wts, int nbands, int per_group, int nx, int ny, int nz, FftAxis ax,
    const double2* __restrict__ tw, double inv_vol, double* __restrict__ partial) {
    extern __shared__ unsigned char smem_raw[];
    const int pitch = ny + 1;
    double2* buf = reinterpret_cast<double2*>(smem_raw);
    double2* twx = buf + nx * pitch;
    const int z = blockIdx.x, grp = blockIdx.y;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int b0 = grp * per_group, b1 = min(nbands, b0 + per_group);
    const int npts = nx * ny;
    load_twiddles(twx, tw + ax.tw_off, nx, tid, nthr);
    double acc[kDensMaxPts];
#pragma unroll
    for (int k = 0; k < kDensMaxPts; ++k) acc[k] = 0.0;
    const unsigned my = div_magic(ny), mx = div_magic(nx);
    const LineSet xl{ny, pitch, 1, nullptr};
    for (int b = b0; b < b1; ++b) {
        const double wb = wts[b];
        if (wb == 0.0) continue;  // same value for the whole workgroup
        for (int i = tid; i < nabsent * ny; i += nthr) {
            const int r = fdiv(i, ny, my), y = i - r * ny;
            buf[y + pitch * xabsent[r]] = make_double2(0.0, 0.0);
        }
        const double2* tcol = tbuf + (size_t)b * nplanes * ny * nz;
        for (int i = tid; i < nplanes * ny; i += nthr) {
            const int p = fdiv(i, ny, my), y = i - p * ny;
            buf[y + pitch * pxpos[p]] = tcol[((size_t)p * nz + z) * ny + y];
        }
        __syncthreads();
        lds_fft_lines<double, double2, true, 1>(buf, ax, twx, xl, tid, nthr);
#pragma unroll
        for (int k = 0; k < kDensMaxPts; ++k) {
            const int i = tid + k * nthr;
            if (i < npts) {
                const int y = fdiv(i, nx, mx), x = i - y * nx;
                const double2 e = buf[y + pitch * x];
                acc[k] += wb * (e.x * e.x + e.y * e.y);
            }
        }
        __syncthreads();
    }
    double* out = partial + (size_t)grp * nx * ny * nz + (size_t)nx * ny * z;
#pragma unroll
    for (int k = 0; k < kDensMaxPts; ++k) {
        const int i = tid + k * nthr;
        if (i < npts) out[i] = acc[k] * inv_vol;
    }
}

__global__ void k_density_sum(const double* __restrict__ partial, int ngroups, size_t n,
                              double* __restrict__ rho, int accumulate) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = accumulate ? rho[i] : 0.0;
    for (int g = 0; g < ngroups; ++g) s += partial[(size_t)g * n + i];
    rho[i] = s;
}

// ----------------------------------------------------------------- nonlocal
__global__ void k_build_a(double2* __restrict__ a, const double* __restrict__ phi_all,
                          const int* __restrict__ row_of, const int* __restrict__ atom_of,
                          const double* __restrict__ gcart, const double* __restrict__ apos,
                          int npw, int nproj) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)npw * nproj) return;
    const int p = (int)(i / npw);
    const int g = (int)(i - (size_t)p * npw);
    const int at = atom_of[p];
    // Same operation order as Vec3.dot: x*x' + y*y' + z*z'.
    const double arg = gcart[3 * g] * apos[3 * at] + gcart[3 * g + 1] * apos[3 * at + 1] +
                       gcart[3 * g + 2] * apos[3 * at + 2];
    double s, c;
    sincos(arg, &s, &c);
    const double ph = phi_all[(size_t)row_of[p] * npw + g];
    a[i] = make_double2(ph * c, ph * s);
}

// MFMA 16x16x4 (f32: v_mfma_f32_16x16x4f32, f64: v_mfma_f64_16x16x4f64).  Per wavefront:
// A operand lane l holds A[i = l%16][k = l/16], B operand B[k = l/16][j = l%16].
// Accumulator element r of lane l is D[row(l/16, r)][j = l%16] with
//   f32: row = 4*(l/16) + r,   f64: row = (l/16) + 4*r
// (both measured on gfx950 with a one-wavefront probe of the builtins).
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
template <typename T> struct Mfma;
template <> struct Mfma<float> {
    using acc = v4f;
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int k4, int r) { return 4 * k4 + r; }
};
template <> struct Mfma<double> {
    using acc = v4d;
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int k4, int r) { return k4 + 4 * r; }
};

// x-line pass as a convolution over the sphere's x-frequencies, on the matrix cores.  The x-part of
// H is diagonal in x, so on the present frequencies kx = x_s + p (p < np, the sphere's cyclic
// x-run) of a line it is
//   Y[q] = sum_{p < np} c[(q - p) mod nx] X[p],   c[d] = sum_x (V_loc(x)/N) e^{-2 pi i d x / nx},
// a Toeplitz matrix that does not depend on the run's start x_s.  Over all bands of a line it is
// an np x np by np x bands complex product, which replaces the two x-transforms.  One workgroup
// per ConvLines<T> consecutive lines (one 128-byte segment of every row) and 16 bands; each
// wavefront owns (line, 16-row block) jobs and keeps their accumulators in registers.  The rows
// stream through LDS in chunks of kConvChunk (every load of a chunk issued before its LDS
// stores); per chunk the jobs take their k-steps of 4 with four real MFMAs per complex product
// into two accumulators (A from the workgroup's c values in LDS, B from the chunk).  In FP64
// launches of at most two 16-row blocks the last n_p mod 16 <= 8 rows are not padded to a block:
// (line, 4-row group) jobs take them with the 4x4x4 MFMA (four blocks = four groups of four
// bands, a third of the 16x16x4's time), whose B operand sits in the same lanes as the
// 16x16x4's (band = lane % 16, k = lane / 16); A[i][k] is lane i + 4 q + 16 k of block q and
// D[i][j] lane 16 i + 4 q + j.  The results go back in place through the same chunk buffer,
// kConvChunk rows at a time, as coalesced stores.
template <typename T> struct ConvLines { static constexpr int value = 16; };
template <> struct ConvLines<double> { static constexpr int value = 8; };
constexpr int kConvBands = 16;
constexpr int kConvMaxPlanes = 48;  // three 16-row blocks: a wavefront's accumulators stay in VGPRs
constexpr int kConvChunk = 12;      // rows per LDS chunk (three k-steps)
constexpr int kConvThreads = 512;   // eight wavefronts share a tile; each holds a few jobs

// LDS index of (row p, band b, line li) in a chunk [p][b][li]; the line index is XORed with the
// band so that the 16 bands of one (p, li), read by 16 lanes, fall in different banks.
template <int LT>
__device__ __forceinline__ int conv_idx(int p, int b, int li) {
    return (p * kConvBands + b) * LT + (li ^ (b & (LT - 1)));
}

// 16-row blocks of n_p rows: FP64 sends a remainder of at most 8 rows to 4-row groups.
__host__ __device__ constexpr int conv_blocks(int np, bool fp64) {
    return fp64 && (np & 15) > 0 && (np & 15) <= 8 ? np >> 4 : (np + 15) >> 4;
}

// NMT: the most 16-row blocks of any line's n_p in the launch (1..3); it sizes the accumulators.
// With at most two blocks the FP64 kernel has the registers for the 4-row tail groups and for
// three real products per complex one in its 16-row blocks (t1 = ar br, t2 = ai bi,
// t3 = (ar + ai)(br + bi); Re = t1 - t2, Im = t3 - t1 - t2); the tail groups keep four.
// Four wavefronts per SIMD (128 VGPRs) except FP64 with three blocks, which fits in 123 on its own
// and runs faster scheduled without the bound.
template <typename T, int NMT>
constexpr int conv_waves_per_eu() { return NMT <= 2 || !std::is_same<T, double>::value ? 4 : 2; }

template <typename T, int NMT>
__global__ __launch_bounds__(kConvThreads, (conv_waves_per_eu<T, NMT>())) void k_xline_conv(typename C2T<T>::type* __restrict__ tbuf, int nplanes, int lines,
                                                    int nc, const typename C2T<T>::type* __restrict__ chat, int nx,
                                                    const int2* __restrict__ kpl) {
    using C = typename C2T<T>::type;
    using M = Mfma<T>;
    constexpr int LT = ConvLines<T>::value;
    constexpr int NB = kConvBands;
    constexpr int kWaves = kConvThreads / 64;
    constexpr int kMaxJobs = LT * NMT / kWaves;                      // per wavefront
    // 4-row tail groups and three products per complex one: FP64 with at most two 16-row blocks
    // (a launch whose largest n_p needs three blocks pads every line's last block, as in FP32).
    constexpr bool kTail = std::is_same<T, double>::value && NMT <= 2;
    constexpr int kMaxTail = kTail ? LT * 2 / kWaves : 1;  // tail jobs per wavefront
    constexpr bool k3 = kTail;
    constexpr int kChunkElems = kConvChunk * NB * LT;
    constexpr int kPer = kChunkElems / kConvThreads;           // loads per thread per chunk
    static_assert(kChunkElems % kConvThreads == 0 && (LT * NMT) % kWaves == 0 && NMT >= 1 && NMT * 16 <= kConvMaxPlanes,
                  "chunk and jobs must split evenly over the workgroup");
    extern __shared__ unsigned char smem_raw[];
    C* xs = reinterpret_cast<C*>(smem_raw);  // one chunk [kConvChunk][NB][LT]
    C* cs = xs + kChunkElems;                // c of the workgroup's lines [LT][nx]
    const int tid = (int)threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int2 kp = xl_planes(kpl, nplanes);
    const int np = kp.y;
    // Tiles in (band tile fastest, line tile) order, each XCD taking one contiguous run of them:
    // workgroups are dealt to the 8 XCDs round-robin, and neighbouring line tiles share the cache
    // lines at their boundary (a plane of ny nz elements is not a whole number of lines), which
    // then meet in one L2 instead of being fetched and written back by two.
    const unsigned ntile = gridDim.x * gridDim.y, lin = blockIdx.x + gridDim.x * blockIdx.y;
    const unsigned xcd = lin & 7u, idx = lin >> 3, per = ntile >> 3, rem = ntile & 7u;
    const unsigned tile = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
    const int b0 = (int)(tile % gridDim.x) * NB, l0 = (int)(tile / gridDim.x) * LT;
    C* base = tbuf + ((size_t)b0 * nplanes + kp.x) * lines + l0;
    const int nlt = min(LT, lines - l0);
    for (int i = tid; i < LT * nx; i += kConvThreads)
        cs[i] = i < nlt * nx ? chat[(size_t)l0 * nx + i] : C{T(0), T(0)};
    // 16-row blocks (the last one padded unless a tail of <= 8 rows goes to 4-row groups).
    const int nrem = np & 15;
    const bool tail = kTail && nrem > 0 && nrem <= 8;
    const int nmt = conv_blocks(np, kTail);  // <= NMT: n_p <= the launch's largest
    const int njobs = LT * nmt;
    const int ntjobs = tail ? LT * ((nrem + 3) >> 2) : 0;
    // Accumulators: (re, im) with four products, (t1, t2, t3) in (re, im, t3) with three.
    typename M::acc re[kMaxJobs], im[kMaxJobs], t3[k3 ? kMaxJobs : 1];
#pragma unroll
    for (int jj = 0; jj < kMaxJobs; ++jj) {
        re[jj] = typename M::acc{0, 0, 0, 0};
        im[jj] = typename M::acc{0, 0, 0, 0};
        if constexpr (k3) t3[jj] = typename M::acc{0, 0, 0, 0};
    }
    double tre[kMaxTail], tim[kMaxTail];
#pragma unroll
    for (int tt = 0; tt < kMaxTail; ++tt) tre[tt] = tim[tt] = 0.0;
    // Chunk rows k0 .. k0 + kConvChunk - 1 (rows >= np zero), loaded into registers one chunk ahead:
    // the next chunk's loads are in flight while the current one is multiplied.
    C v[kPer];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = tid + u * kConvThreads;
            const int li = e % LT, r = e / LT, b = r % NB, p = k0 + r / NB;
            v[u] = (p < np && b0 + b < nc && li < nlt) ? base[((size_t)b * nplanes + p) * lines + li] : C{T(0), T(0)};
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < np; k0 += kConvChunk) {
        __syncthreads();  // the previous chunk is consumed
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = tid + u * kConvThreads;
            const int li = e % LT, r = e / LT, b = r % NB;
            xs[conv_idx<LT>(r / NB, b, li)] = v[u];
        }
        __syncthreads();
        if (k0 + kConvChunk < np) fetch(k0 + kConvChunk);
#pragma unroll
        for (int jj = 0; jj < kMaxJobs; ++jj) {
            const int job = wave + kWaves * jj;
            if (job >= njobs) break;
            const int li = job % LT, mt = job / LT;
            const int q = mt * 16 + l16;  // output row of this lane's A operand
            const C* cl = cs + li * nx;
#pragma unroll
            for (int ks = 0; ks < kConvChunk; ks += 4) {
                const int p = k0 + ks + l4;
                C a = C{T(0), T(0)};
                if (q < np && p < np) {
                    int d = q - p;
                    d += d < 0 ? nx : 0;
                    a = cl[d];
                }
                const C bv = xs[conv_idx<LT>(ks + l4, l16, li)];
                if constexpr (k3) {
                    re[jj] = M::mma(a.x, bv.x, re[jj]);  // t1
                    im[jj] = M::mma(a.y, bv.y, im[jj]);  // t2
                    t3[jj] = M::mma(a.x + a.y, bv.x + bv.y, t3[jj]);
                } else {
                    // a b = (ar br - ai bi) + i (ar bi + ai br): four real products, two accumulators
                    re[jj] = M::mma(a.x, bv.x, re[jj]);
                    re[jj] = M::mma(-a.y, bv.y, re[jj]);
                    im[jj] = M::mma(a.x, bv.y, im[jj]);
                    im[jj] = M::mma(a.y, bv.x, im[jj]);
                }
            }
        }
        if constexpr (kTail) {
#pragma unroll
            for (int tt = 0; tt < kMaxTail; ++tt) {
                const int job = wave + kWaves * tt;
                if (job >= ntjobs) break;
                const int li = job % LT;
                const int q = 16 * nmt + 4 * (job / LT) + (lane & 3);  // A row of this lane
                const C* cl = cs + li * nx;
#pragma unroll
                for (int ks = 0; ks < kConvChunk; ks += 4) {
                    const int p = k0 + ks + l4;
                    C a = C{T(0), T(0)};
                    if (q < np && p < np) {
                        int d = q - p;
                        d += d < 0 ? nx : 0;
                        a = cl[d];
                    }
                    const C bv = xs[conv_idx<LT>(ks + l4, l16, li)];
                    tre[tt] = __builtin_amdgcn_mfma_f64_4x4x4f64(a.x, bv.x, tre[tt], 0, 0, 0);
                    tre[tt] = __builtin_amdgcn_mfma_f64_4x4x4f64(-a.y, bv.y, tre[tt], 0, 0, 0);
                    tim[tt] = __builtin_amdgcn_mfma_f64_4x4x4f64(a.x, bv.y, tim[tt], 0, 0, 0);
                    tim[tt] = __builtin_amdgcn_mfma_f64_4x4x4f64(a.y, bv.x, tim[tt], 0, 0, 0);
                }
            }
        }
    }
    // Results back in place through the chunk buffer, kConvChunk rows at a time.
    for (int r0 = 0; r0 < np; r0 += kConvChunk) {
        __syncthreads();  // the buffer is free
#pragma unroll
        for (int jj = 0; jj < kMaxJobs; ++jj) {
            const int job = wave + kWaves * jj;
            if (job >= njobs) break;
            const int li = job % LT, mt = job / LT;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = mt * 16 + M::row(l4, r);
                if (q >= r0 && q < r0 + kConvChunk && q < np) {
                    if constexpr (k3)
                        xs[conv_idx<LT>(q - r0, l16, li)] = C{re[jj][r] - im[jj][r], t3[jj][r] - re[jj][r] - im[jj][r]};
                    else
                        xs[conv_idx<LT>(q - r0, l16, li)] = C{re[jj][r], im[jj][r]};
                }
            }
        }
        if constexpr (kTail) {
#pragma unroll
            for (int tt = 0; tt < kMaxTail; ++tt) {
                const int job = wave + kWaves * tt;
                if (job >= ntjobs) break;
                const int li = job % LT;
                const int q = 16 * nmt + 4 * (job / LT) + l4;  // D row of this lane
                if (q >= r0 && q < r0 + kConvChunk && q < np) xs[conv_idx<LT>(q - r0, l16, li)] = C{T(tre[tt]), T(tim[tt])};
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int e = tid + u * kConvThreads;
            const int li = e % LT, r = e / LT, b = r % NB, q = r0 + r / NB;
            if (q < np && b0 + b < nc && li < nlt)
                base[((size_t)b * nplanes + q) * lines + li] = xs[conv_idx<LT>(r / NB, b, li)];
        }
    }
}

template <typename T>
size_t xline_conv_lds_bytes(int nx) {
    return ((size_t)kConvChunk * kConvBands * ConvLines<T>::value + (size_t)ConvLines<T>::value * nx) *
           sizeof(typename C2T<T>::type);
}

bool xline_conv_usable(const Ham& h) {
    return h.plane_xs >= 0 && h.nplanes >= 1 && h.nplanes <= kConvMaxPlanes &&
           xline_conv_lds_bytes<double>(h.nx) <= kMaxLdsBytes;
}

// Convolution x-line kernel on h's grid (kpl / nk / planes_per_band / max_planes as for
// launch_xline_reg); reads h's per-line transform of V_loc/N.
template <typename T>
void launch_xline_conv(Ham& h, typename C2T<T>::type* tb, size_t nc, hipStream_t s, const int2* kpl = nullptr,
                       int nk = 1, int planes_per_band = -1, int max_planes = -1) {
    using C = typename C2T<T>::type;
    const int lines = h.ny * h.nz;
    const int nplanes = planes_per_band >= 0 ? planes_per_band : h.nplanes;
    const int np_max = max_planes >= 0 ? max_planes : h.nplanes;
    DZ_REQUIRE(np_max >= 1 && np_max <= kConvMaxPlanes, DZHIP_ERR_STATE, "convolution x-line kernel: too many planes");
    const size_t lds = xline_conv_lds_bytes<T>(h.nx);
    // n_p of every line of the launch is <= np_max, and conv_blocks is non-decreasing in n_p.
    const int nmt = conv_blocks(np_max, std::is_same<T, double>::value);
    const C* chat = std::is_same<T, float>::value ? (const C*)h.chat32.ptr : (const C*)h.chat64.ptr;
    const dim3 grid((unsigned)((nc + kConvBands - 1) / kConvBands),
                    (unsigned)((lines + ConvLines<T>::value - 1) / ConvLines<T>::value), (unsigned)nk);
    auto launch = [&](auto kern) {
        set_max_dynamic_lds((const void*)kern, lds);
        hipLaunchKernelGGL(kern, grid, dim3(kConvThreads), lds, s, tb, nplanes, lines, (int)nc, chat, h.nx, kpl);
    };
    if (nmt <= 1)
        launch(k_xline_conv<T, 1>);
    else if (nmt == 2)
        launch(k_xline_conv<T, 2>);
    else
        launch(k_xline_conv<T, 3>);
    DZ_HIP(hipGetLastError());
}

// Workgroup tiles of the nonlocal GEMMs: 10 wavefronts = 5 band groups of 32 bands x 2 halves
// of the projector (projection) or G (back-projection) tile, so one workgroup covers a 160-band
// block, each projector row is read once per K-chunk, and two workgroups fit on a CU with the
// wavefronts spread evenly over its four SIMDs.
constexpr int kNlBandGroups = 5;
constexpr int kNlWaves = 2 * kNlBandGroups;
constexpr int kNlThreads = 64 * kNlWaves;
constexpr int kNlBands = 32 * kNlBandGroups;  // bands per workgroup tile
constexpr int kNlK = 32;                 // K (G or p) per LDS stage
constexpr int kNlPitch = kNlK + 1;       // odd pitches (in complex elements) against bank conflicts
// 16-wide MFMA blocks per workgroup tile along the projector (projection) and G
// (back-projection) axes, split over the two wavefront halves: 4 in FP32; 2 in FP64, whose
// accumulators are twice as wide.
template <typename T> struct NlTile { static constexpr int kBlocks = 4; };
template <> struct NlTile<double> { static constexpr int kBlocks = 2; };
template <typename T> constexpr int nl_p_tile() { return 16 * NlTile<T>::kBlocks; }  // projectors / tile
// Back-projection G tiles of two 16-wide blocks (one per wavefront half): at the bench size 746
// tiles of 32 G even out the work per CU better than 373 tiles of 64 G on 256 CUs (two
// workgroups fit per CU either way).
template <typename T> struct NlBackTile { static constexpr int kBlocks = 2; };
template <typename T> constexpr int nl_g_tile() { return 16 * NlBackTile<T>::kBlocks; }  // G / tile

template <typename T>
__host__ __device__ inline size_t nl_proj_lds_bytes() {
    return (size_t)(kNlBands + nl_p_tile<T>()) * kNlPitch * sizeof(typename C2T<T>::type);
}
template <typename T>
__host__ __device__ inline size_t nl_back_lds_bytes() {
    return ((size_t)kNlBands * kNlPitch + (size_t)kNlK * (nl_g_tile<T>() + 1)) * sizeof(typename C2T<T>::type);
}

// Block index -> logical tile so that consecutive logical tiles (which share operands)
// land on the same XCD: hardware dispatches block i to XCD i % 8.
__device__ __forceinline__ int xcd_tile(int bid, int nblocks) {
    constexpr int kXcd = 8;
    const int per = nblocks / kXcd;
    if (per == 0 || bid >= per * kXcd) return bid;
    return (bid % kXcd) * per + bid / kXcd;
}

// partial[ks][b][p] = sum_{G in chunk ks} X[b][G] * A[p][G]   (= <beta_p|psi_b>, no 1/Omega yet)
template <typename T>
__global__ __launch_bounds__(kNlThreads) void k_nl_proj(const typename C2T<T>::type* __restrict__ a,
                                                        const typename C2T<T>::type* __restrict__ x,
                                                        size_t ldx, typename C2T<T>::type* __restrict__ partial,
                                                        int P, int npw, int ncols, int kchunk, int ptiles,
                                                        const int2* __restrict__ chunks, int lda) {
    using C = typename C2T<T>::type;
    using M = Mfma<T>;
    extern __shared__ unsigned char smem_raw[];
    C* xs = reinterpret_cast<C*>(smem_raw);  // [kNlBands][kNlPitch]
    constexpr int kNlP = nl_p_tile<T>();
    constexpr int PB = NlTile<T>::kBlocks / 2;  // projector blocks per wavefront
    C* as = xs + kNlBands * kNlPitch;        // [kNlP][kNlPitch]
    const int lin = xcd_tile(blockIdx.x, gridDim.x);
    const int pt = lin % ptiles, ks = lin / ptiles;
    const int p0 = pt * kNlP;
    const int b0 = blockIdx.z * kNlBands;
    // G range of K-chunk ks: a fixed-size split of one context's sphere, or (batched apply over
    // several k-points, G of all k-points concatenated) the chunk table's range within one k-point.
    int k0 = ks * kchunk, k1 = min(npw, ks * kchunk + kchunk);
    if (chunks) {
        const int2 ck = chunks[ks];
        k0 = ck.x;
        k1 = ck.y;
    }
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int bg = wave % kNlBandGroups, half = wave / kNlBandGroups;
    const int i16 = lane & 15, k4 = lane >> 4;
    const C zero = C{T(0), T(0)};
    constexpr int kXPer = (kNlBands * kNlK + kNlThreads - 1) / kNlThreads;
    constexpr int kAPer = (kNlP * kNlK + kNlThreads - 1) / kNlThreads;
    T xre[kXPer], xim[kXPer], are[kAPer], aim[kAPer];
#define DZ_PROJ_FETCH(KT)                                                                     \
    _Pragma("unroll") for (int i = 0; i < kXPer; ++i) {                                      \
        const int e = tid + i * kNlThreads;                                                   \
        const int r = e >> 5, c = e & 31, g = (KT) + c, b = b0 + r;                           \
        const bool ok = r < kNlBands && b < ncols && g < k1;                                  \
        const C v = x[ok ? (size_t)b * ldx + g : 0];                                          \
        xre[i] = ok ? v.x : T(0);                                                             \
        xim[i] = ok ? v.y : T(0);                                                             \
    }                                                                                         \
    _Pragma("unroll") for (int i = 0; i < kAPer; ++i) {                                      \
        const int e = tid + i * kNlThreads;                                                   \
        const int r = e >> 5, c = e & 31, g = (KT) + c, p = p0 + r;                           \
        const bool ok = r < kNlP && p < P && g < k1;                                          \
        const C v = a[ok ? (size_t)p * lda + g : 0];                                          \
        are[i] = ok ? v.x : T(0);                                                             \
        aim[i] = ok ? v.y : T(0);                                                             \
    }
#define DZ_PROJ_STASH()                                                                       \
    _Pragma("unroll") for (int i = 0; i < kXPer; ++i) {                                      \
        const int e = tid + i * kNlThreads;                                                   \
        if (e < kNlBands * kNlK) xs[(e >> 5) * kNlPitch + (e & 31)] = C{xre[i], xim[i]};     \
    }                                                                                         \
    _Pragma("unroll") for (int i = 0; i < kAPer; ++i) {                                      \
        const int e = tid + i * kNlThreads;                                                   \
        if (e < kNlP * kNlK) as[(e >> 5) * kNlPitch + (e & 31)] = C{are[i], aim[i]};         \
    }
    // Complex products by three real MFMAs (Gauss): with x = xr + i xi, a = ar + i ai,
    // acc1 = sum xr ar, acc2 = sum xi ai, acc3 = sum (xr + xi)(ar + ai);
    // Re = acc1 - acc2, Im = acc3 - acc1 - acc2.
    typename M::acc acc1[2][PB], acc2[2][PB], acc3[2][PB];
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int pb = 0; pb < PB; ++pb) {
            acc1[bb][pb] = typename M::acc{0, 0, 0, 0};
            acc2[bb][pb] = typename M::acc{0, 0, 0, 0};
            acc3[bb][pb] = typename M::acc{0, 0, 0, 0};
        }
    if (k0 < k1) {
        DZ_PROJ_FETCH(k0)
        DZ_PROJ_STASH()
    }
    __syncthreads();
    for (int kt = k0; kt < k1; kt += kNlK) {
        const bool more = kt + kNlK < k1;
        if (more) {
            DZ_PROJ_FETCH(kt + kNlK)
        }
#pragma unroll
        for (int kk = 0; kk < kNlK; kk += 4) {
            C xa[2], ab[PB];
#pragma unroll
            for (int bb = 0; bb < 2; ++bb) xa[bb] = xs[(bg * 32 + bb * 16 + i16) * kNlPitch + kk + k4];
#pragma unroll
            for (int pb = 0; pb < PB; ++pb) ab[pb] = as[((half * PB + pb) * 16 + i16) * kNlPitch + kk + k4];
            T xsum[2], asum[PB];
#pragma unroll
            for (int bb = 0; bb < 2; ++bb) xsum[bb] = xa[bb].x + xa[bb].y;
#pragma unroll
            for (int pb = 0; pb < PB; ++pb) asum[pb] = ab[pb].x + ab[pb].y;
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int pb = 0; pb < PB; ++pb) {
                    acc1[bb][pb] = M::mma(xa[bb].x, ab[pb].x, acc1[bb][pb]);
                    acc2[bb][pb] = M::mma(xa[bb].y, ab[pb].y, acc2[bb][pb]);
                    acc3[bb][pb] = M::mma(xsum[bb], asum[pb], acc3[bb][pb]);
                }
        }
        __syncthreads();
        if (more) {
            DZ_PROJ_STASH()
        }
        __syncthreads();
    }
#undef DZ_PROJ_FETCH
#undef DZ_PROJ_STASH
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int pb = 0; pb < PB; ++pb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int b = b0 + bg * 32 + bb * 16 + M::row(k4, r);
                const int p = p0 + (half * PB + pb) * 16 + i16;
                const T c1 = acc1[bb][pb][r], c2 = acc2[bb][pb][r];
                if (b < ncols && p < P)
                    partial[((size_t)ks * ncols + b) * P + p] = C{c1 - c2, acc3[bb][pb][r] - c1 - c2};
            }
}

// One workgroup per band: sum the split-K partials of every projector into LDS (each partial
// read once, coalesced over p; `split` interleaved subsets of the K-chunks per projector so
// more loads are in flight, combined in a fixed order), then
// q[b][p] = inv_vol * sum_r D[m][r] * sum[base + r] within the projector's atom block.
constexpr int kDijSplit = 4;
constexpr int kDijThreads = 1024;
// K-chunk subsets the D kernel can stage for P projectors of element size esz, and its LDS bytes.
inline int dij_split(int P, size_t esz) {
    for (int sp = kDijSplit; sp > 1; --sp)
        if ((size_t)(sp + 1) * P * esz <= kMaxLdsBytes) return sp;
    return 1;
}
inline size_t dij_lds_bytes(int P, size_t esz) {
    const int sp = dij_split(P, esz);
    return (size_t)(sp == 1 ? 1 : sp + 1) * P * esz;
}

template <typename T>
__global__ __launch_bounds__(kDijThreads) void k_nl_dij(const typename C2T<T>::type* __restrict__ partial,
                                                        int nks, typename C2T<T>::type* __restrict__ q,
                                                        const ProjMeta* __restrict__ meta,
                                                        const T* __restrict__ dm, int P, int ncols, T inv_vol,
                                                        int split, const int2* __restrict__ kseg) {
    using C = typename C2T<T>::type;
    extern __shared__ unsigned char smem_raw[];
    C* part = reinterpret_cast<C*>(smem_raw);                     // [split][P]
    C* sum = split == 1 ? part : part + (size_t)split * P;       // [P]
    const int b = blockIdx.x;
    const int tid = threadIdx.x, nthr = blockDim.x;
    // K-chunks to sum: all of them, or (batched apply) those of k-point blockIdx.y, whose q block
    // follows the earlier k-points'.
    int ks_lo = 0, ks_hi = nks;
    if (kseg) {
        const int2 r = kseg[blockIdx.y];
        ks_lo = r.x;
        ks_hi = r.y;
        q += (size_t)blockIdx.y * ncols * P;
    }
    for (int idx = tid; idx < split * P; idx += nthr) {
        const int j = idx / P, p = idx - j * P;
        C s = C{T(0), T(0)};
        for (int ks = ks_lo + j; ks < ks_hi; ks += split) {
            const C v = partial[((size_t)ks * ncols + b) * P + p];
            s.x += v.x;
            s.y += v.y;
        }
        part[(size_t)j * P + p] = s;
    }
    __syncthreads();
    if (split > 1) {
        for (int p = tid; p < P; p += nthr) {
            C s = part[p];
            for (int j = 1; j < split; ++j) {
                s.x += part[(size_t)j * P + p].x;
                s.y += part[(size_t)j * P + p].y;
            }
            sum[p] = s;
        }
        __syncthreads();
    }
    for (int p = tid; p < P; p += nthr) {
        const ProjMeta mt = meta[p];
        const T* drow = dm + mt.dm_off + mt.m * mt.mt;
        C acc = C{T(0), T(0)};
        for (int r = 0; r < mt.mt; ++r) {
            const C v = sum[mt.base + r];
            acc.x += drow[r] * v.x;
            acc.y += drow[r] * v.y;
        }
        q[(size_t)b * P + p] = C{acc.x * inv_vol, acc.y * inv_vol};
    }
}

// y[b][G] = sum_p q[b][p] * conj(A[p][G])   (written, not accumulated: runs before k_plane_fwd)
template <typename T>
__global__ __launch_bounds__(kNlThreads) void k_nl_back(const typename C2T<T>::type* __restrict__ a,
                                                        const typename C2T<T>::type* __restrict__ q,
                                                        typename C2T<T>::type* __restrict__ y, size_t ldy,
                                                        int P, int npw, int ncols, const int4* __restrict__ tiles,
                                                        int lda) {
    using C = typename C2T<T>::type;
    using M = Mfma<T>;
    extern __shared__ unsigned char smem_raw[];
    C* qs = reinterpret_cast<C*>(smem_raw);  // [kNlBands][kNlPitch]  (band, p)
    constexpr int kNlG = nl_g_tile<T>();
    constexpr int kNlGPitch = kNlG + 1;
    constexpr int GB = NlBackTile<T>::kBlocks / 2;  // G blocks per wavefront
    C* as = qs + kNlBands * kNlPitch;        // [kNlK][kNlGPitch]     (p, G)
    // G tile: fixed-size tiles of one context's sphere, or (batched apply) the tile table's
    // (first G, end of its k-point's G range, k-point) with that k-point's q block.
    int g0 = blockIdx.x * kNlG, gend = npw;
    if (tiles) {
        const int4 tl = tiles[blockIdx.x];
        g0 = tl.x;
        gend = tl.y;
        q += (size_t)tl.z * ncols * P;
    }
    const int b0 = blockIdx.z * kNlBands;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int bg = wave % kNlBandGroups, half = wave / kNlBandGroups;
    const int i16 = lane & 15, k4 = lane >> 4;
    const C zero = C{T(0), T(0)};
    constexpr int kQPer = (kNlBands * kNlK + kNlThreads - 1) / kNlThreads;
    constexpr int kAPer = (kNlK * kNlG + kNlThreads - 1) / kNlThreads;
    T qre[kQPer], qim[kQPer], are[kAPer], aim[kAPer];
#define DZ_BACK_FETCH(PT)                                                                     \
    _Pragma("unroll") for (int i = 0; i < kQPer; ++i) {                                      \
        const int e = tid + i * kNlThreads;                                                   \
        const int r = e >> 5, c = e & 31, b = b0 + r, p = (PT) + c;                           \
        const bool ok = r < kNlBands && b < ncols && p < P;                                   \
        const C v = q[ok ? b * P + p : 0];                                                    \
        qre[i] = ok ? v.x : T(0);                                                             \
        qim[i] = ok ? v.y : T(0);                                                             \
    }                                                                                         \
    _Pragma("unroll") for (int i = 0; i < kAPer; ++i) {                                      \
        const int e = tid + i * kNlThreads;                                                   \
        const int r = e / kNlG, c = e % kNlG, p = (PT) + r, g = g0 + c;                       \
        const bool ok = r < kNlK && p < P && g < gend;                                        \
        const C v = a[ok ? (size_t)p * lda + g : 0];                                          \
        are[i] = ok ? v.x : T(0);                                                             \
        aim[i] = ok ? v.y : T(0);                                                             \
    }
#define DZ_BACK_STASH()                                                                       \
    _Pragma("unroll") for (int i = 0; i < kQPer; ++i) {                                      \
        const int e = tid + i * kNlThreads;                                                   \
        if (e < kNlBands * kNlK) qs[(e >> 5) * kNlPitch + (e & 31)] = C{qre[i], qim[i]};     \
    }                                                                                         \
    _Pragma("unroll") for (int i = 0; i < kAPer; ++i) {                                      \
        const int e = tid + i * kNlThreads;                                                   \
        if (e < kNlK * kNlG) as[(e / kNlG) * kNlGPitch + (e % kNlG)] = C{are[i], aim[i]};   \
    }
    // q conj(a) by three real MFMAs (Gauss) with a' = ar - i ai: acc1 = sum qr ar,
    // acc2 = sum qi (-ai), acc3 = sum (qr + qi)(ar - ai); Re = acc1 - acc2, Im = acc3 - acc1 - acc2.
    typename M::acc acc1[2][GB], acc2[2][GB], acc3[2][GB];
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int gb = 0; gb < GB; ++gb) {
            acc1[bb][gb] = typename M::acc{0, 0, 0, 0};
            acc2[bb][gb] = typename M::acc{0, 0, 0, 0};
            acc3[bb][gb] = typename M::acc{0, 0, 0, 0};
        }
    DZ_BACK_FETCH(0)
    DZ_BACK_STASH()
    __syncthreads();
    for (int pt = 0; pt < P; pt += kNlK) {
        const bool more = pt + kNlK < P;
        if (more) {
            DZ_BACK_FETCH(pt + kNlK)
        }
#pragma unroll
        for (int kk = 0; kk < kNlK; kk += 4) {
            C qa[2], ab[GB];
#pragma unroll
            for (int bb = 0; bb < 2; ++bb) qa[bb] = qs[(bg * 32 + bb * 16 + i16) * kNlPitch + kk + k4];
#pragma unroll
            for (int gb = 0; gb < GB; ++gb) ab[gb] = as[(kk + k4) * kNlGPitch + (half * GB + gb) * 16 + i16];
            T qsum[2], adif[GB];
#pragma unroll
            for (int bb = 0; bb < 2; ++bb) qsum[bb] = qa[bb].x + qa[bb].y;
#pragma unroll
            for (int gb = 0; gb < GB; ++gb) adif[gb] = ab[gb].x - ab[gb].y;
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int gb = 0; gb < GB; ++gb) {
                    acc1[bb][gb] = M::mma(qa[bb].x, ab[gb].x, acc1[bb][gb]);
                    acc2[bb][gb] = M::mma(qa[bb].y, -ab[gb].y, acc2[bb][gb]);
                    acc3[bb][gb] = M::mma(qsum[bb], adif[gb], acc3[bb][gb]);
                }
        }
        __syncthreads();
        if (more) {
            DZ_BACK_STASH()
        }
        __syncthreads();
    }
#undef DZ_BACK_FETCH
#undef DZ_BACK_STASH
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int gb = 0; gb < GB; ++gb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int b = b0 + bg * 32 + bb * 16 + M::row(k4, r);
                const int g = g0 + (half * GB + gb) * 16 + i16;
                const T c1 = acc1[bb][gb][r], c2 = acc2[bb][gb][r];
                if (b < ncols && g < gend) y[(size_t)b * ldy + g] = C{c1 - c2, acc3[bb][gb][r] - c1 - c2};
            }
}

template <typename T>
__global__ void k_fill_zero(typename C2T<T>::type* __restrict__ y, size_t n) {
    using C = typename C2T<T>::type;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = C{T(0), T(0)};
}

__global__ void k_d2f(const double* __restrict__ in, float* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// ----------------------------------------------------------------- host side
size_t Ham::device_bytes() const {
    const DevBuf* bufs[] = {&plane_start, &plane_pw, &plane_pos, &plane_xpos, &plane_ystart,
                            &plane_ycols, &plane_xabsent, &kinetic64,
                            &plane_zstart, &plane_zrows, &plane_posr, &plane_pw0,
                            &kinetic32, &vloc64, &vloc32, &a64, &a32, &meta, &block_rows, &dm64, &dm32,
                            &xstage, &ystage, &x32, &y32, &tbuf, &partial, &qbuf,
                            &vloc32_t, &vloc64_t, &chat64, &chat32};
    size_t s = tab.bytes;
    for (const DevBuf* b : bufs) s += b->bytes;
    return s;
}

template <typename V>
static void upload(DevBuf& buf, const std::vector<V>& v) {
    buf.ensure(std::max<size_t>(v.size() * sizeof(V), 16));
    if (!v.empty()) DZ_HIP(hipMemcpy(buf.ptr, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
}

void ham_build(Ham& h, const dzhip_ham_desc& d) {
    DZ_REQUIRE(d.nx > 0 && d.ny > 0 && d.nz > 0, DZHIP_ERR_ARG, "grid dimensions must be positive");
    DZ_REQUIRE(d.n_pw > 0 && d.fft_index && d.kinetic, DZHIP_ERR_ARG,
               "n_pw, fft_index and kinetic are required");
    h.nx = (int)d.nx;
    h.ny = (int)d.ny;
    h.nz = (int)d.nz;
    h.npw = d.n_pw;
    h.ngrid = (size_t)d.nx * d.ny * d.nz;
    h.inv_volume = d.inv_volume;
    DZ_REQUIRE(h.ngrid < (size_t)1 << 31, DZHIP_ERR_ARG, "grid too large");
    DZ_REQUIRE(plane_lds_bytes<double2>(h.ny, h.nz) <= kMaxLdsBytes &&
                   xline_lds_bytes<double2>(h.nx, h.ny) <= kMaxLdsBytes,
               DZHIP_ERR_ARG, "grid too large for one LDS tile");
    h.tab.build(h.nx, h.ny, h.nz);

    // Host copies of the permutations to place plane-wave entries.
    std::vector<int> rx, ry, rz;
    factor_fft_len(h.nx, rx);
    factor_fft_len(h.ny, ry);
    factor_fft_len(h.nz, rz);
    const std::vector<int> px = digit_reverse_perm(h.nx, rx);
    const std::vector<int> py = digit_reverse_perm(h.ny, ry);
    const std::vector<int> pz = digit_reverse_perm(h.nz, rz);

    // Split the sphere into x-planes; keep basis order inside a plane (contiguous runs
    // for the reference's h-major ordering).
    std::vector<int> xs(h.npw), ys(h.npw), zs(h.npw);
    for (size_t i = 0; i < h.npw; ++i) {
        const int64_t f = d.fft_index[i];
        DZ_REQUIRE(f >= 0 && (size_t)f < h.ngrid, DZHIP_ERR_ARG, "fft_index out of range");
        xs[i] = (int)(f % h.nx);
        ys[i] = (int)((f / h.nx) % h.ny);
        zs[i] = (int)(f / ((int64_t)h.nx * h.ny));
    }
    std::vector<int> count(h.nx, 0);
    for (size_t i = 0; i < h.npw; ++i) count[xs[i]]++;
    // The x values that hold plane waves normally form one cyclic run x0, x0 + 1, ... (mod nx)
    // (the G-sphere around k = 0 wraps past nx).  Planes are numbered along that run, so plane p
    // is x = x0 + p (mod nx) and the register x-line kernel reads its rows as a prefix; any other
    // pattern is numbered in increasing x and runs the LDS x-line kernel.
    int x0 = -1, run = 0;
    {
        int present = 0, starts = 0;
        for (int xx = 0; xx < h.nx; ++xx) {
            if (count[xx] == 0) continue;
            ++present;
            if (count[(xx + h.nx - 1) % h.nx] == 0) {
                ++starts;
                x0 = xx;
            }
        }
        if (present == h.nx) x0 = 0;
        else if (starts != 1) x0 = -1;
        run = x0 >= 0 ? present : 0;
    }
    std::vector<int> plane_of_x(h.nx, -1), xval;
    for (int k = 0; k < h.nx; ++k) {
        const int xx = x0 >= 0 ? (x0 + k) % h.nx : k;
        if (count[xx] > 0) {
            plane_of_x[xx] = (int)xval.size();
            xval.push_back(xx);
        }
    }
    h.nplanes = (int)xval.size();
    h.plane_xs = x0;
    DZ_REQUIRE(x0 < 0 || run == h.nplanes, DZHIP_ERR_STATE, "x-plane run mismatch");
    std::vector<int> start(h.nplanes + 1, 0);
    for (int p = 0; p < h.nplanes; ++p) start[p + 1] = start[p] + count[xval[p]];
    std::vector<int> fill(start.begin(), start.end() - 1);
    std::vector<int> pw(h.npw), pos(h.npw);
    const int pitch = h.ny + 1;  // LDS row pitch of the [z][y] plane tile
    std::vector<std::vector<char>> yused(h.nplanes, std::vector<char>(h.ny, 0));
    std::vector<std::vector<int>> zrow(h.nplanes, std::vector<int>(h.nz, -1));
    std::vector<int> pidx(h.npw);
    for (size_t i = 0; i < h.npw; ++i) {
        const int p = plane_of_x[xs[i]];
        const int j = fill[p]++;
        pw[j] = (int)i;
        pidx[j] = p;
        pos[j] = py[ys[i]] + pitch * pz[zs[i]];
        yused[p][ys[i]] = 1;
        zrow[p][zs[i]] = 0;
    }
    // Active z-rows per plane (ascending z-frequency) and the entries' positions in them.
    std::vector<int> zstart(h.nplanes + 1, 0);
    std::vector<short> zrows;
    h.zrow_max = 0;
    for (int p = 0; p < h.nplanes; ++p) {
        int b = 0;
        for (int zz = 0; zz < h.nz; ++zz)
            if (zrow[p][zz] >= 0) {
                zrow[p][zz] = b++;
                zrows.push_back((short)zz);
            }
        zstart[p + 1] = (int)zrows.size();
        h.zrow_max = std::max(h.zrow_max, b);
    }
    std::vector<int> posr(h.npw);
    for (size_t j = 0; j < h.npw; ++j) {
        const size_t i = (size_t)pw[j];
        posr[j] = py[ys[i]] + zr_pitch(h.ny) * zrow[pidx[j]][zs[i]];
    }
    // First plane-wave index of each plane whose entries are one run of the basis, else -1.
    std::vector<int> pw0(h.nplanes, -1);
    for (int p = 0; p < h.nplanes; ++p) {
        bool run = start[p + 1] > start[p];
        for (int j = start[p] + 1; run && j < start[p + 1]; ++j) run = pw[j] == pw[j - 1] + 1;
        if (run) pw0[p] = pw[start[p]];
    }
    upload(h.plane_pw0, pw0);
    upload(h.plane_zstart, zstart);
    upload(h.plane_zrows, zrows);
    upload(h.plane_posr, posr);
    // Per plane, the LDS columns (digit-reversed y) that hold plane waves: only these need
    // the z-transform.
    std::vector<int> ystart(h.nplanes + 1, 0);
    std::vector<short> ycols;
    for (int p = 0; p < h.nplanes; ++p) {
        for (int yy = 0; yy < h.ny; ++yy)
            if (yused[p][yy]) ycols.push_back((short)py[yy]);
        ystart[p + 1] = (int)ycols.size();
    }
    std::vector<int> xpos(h.nplanes);
    for (int p = 0; p < h.nplanes; ++p) xpos[p] = px[xval[p]];
    upload(h.plane_start, start);
    upload(h.plane_pw, pw);
    upload(h.plane_pos, pos);
    upload(h.plane_xpos, xpos);
    upload(h.plane_ystart, ystart);
    upload(h.plane_ycols, ycols);
    std::vector<int> xabsent;
    for (int xx = 0; xx < h.nx; ++xx)
        if (plane_of_x[xx] < 0) xabsent.push_back(px[xx]);
    h.nxabsent = (int)xabsent.size();
    upload(h.plane_xabsent, xabsent);
    std::vector<double> kin(d.kinetic, d.kinetic + h.npw);
    std::vector<float> kin32(kin.begin(), kin.end());
    upload(h.kinetic64, kin);
    upload(h.kinetic32, kin32);

    // Nonlocal tables.
    h.nproj = 0;
    if (d.n_species > 0 && d.n_atoms > 0) {
        DZ_REQUIRE(d.species && d.atom_pos && d.atom_species && d.gcart, DZHIP_ERR_ARG,
                   "nonlocal needs species, atom_pos, atom_species and gcart");
        std::vector<int> sp_row0(d.n_species), sp_dm0(d.n_species);
        std::vector<double> phi_all, dm_all;
        size_t rows = 0;
        for (uint64_t s = 0; s < d.n_species; ++s) {
            const dzhip_species_desc& sd = d.species[s];
            uint64_t mt = 0;
            for (uint64_t b = 0; b < sd.n_beta; ++b) {
                DZ_REQUIRE(sd.l_list[b] >= 0 && sd.l_list[b] <= 4, DZHIP_ERR_ARG, "l must be 0..4");
                mt += 2 * sd.l_list[b] + 1;
            }
            DZ_REQUIRE(mt == sd.m_total, DZHIP_ERR_ARG, "m_total != sum(2l+1)");

            DZ_REQUIRE(mt == 0 || sd.phi, DZHIP_ERR_ARG, "species phi table missing");
            DZ_REQUIRE(sd.n_beta == 0 || sd.dij || sd.dij_m, DZHIP_ERR_ARG, "species D missing");
            sp_row0[s] = (int)rows;
            sp_dm0[s] = (int)dm_all.size();
            phi_all.insert(phi_all.end(), sd.phi, sd.phi + mt * h.npw);
            rows += mt;
            // Dense m-resolved D: either given, or expanded from the per-l radial form
            // (only same-l pairs couple, same m).
            std::vector<double> dmv(mt * mt, 0.0);
            if (sd.dij_m) {
                std::memcpy(dmv.data(), sd.dij_m, mt * mt * sizeof(double));
            } else {
                std::vector<int> off(sd.n_beta);
                int o = 0;
                for (uint64_t b = 0; b < sd.n_beta; ++b) {
                    off[b] = o;
                    o += 2 * sd.l_list[b] + 1;
                }
                for (uint64_t b = 0; b < sd.n_beta; ++b)
                    for (uint64_t j = 0; j < sd.n_beta; ++j) {
                        if (sd.l_list[j] != sd.l_list[b]) continue;
                        const double dv = sd.dij[b * sd.n_beta + j];
                        for (int m = 0; m < 2 * sd.l_list[b] + 1; ++m)
                            dmv[(off[b] + m) * mt + off[j] + m] = dv;
                    }
            }
            dm_all.insert(dm_all.end(), dmv.begin(), dmv.end());
        }
        std::vector<int> row_of, atom_of;
        std::vector<ProjMeta> meta;
        for (uint64_t a = 0; a < d.n_atoms; ++a) {
            const int s = d.atom_species[a];
            DZ_REQUIRE(s >= 0 && (uint64_t)s < d.n_species, DZHIP_ERR_ARG, "atom species out of range");
            const int mt = (int)d.species[s].m_total;
            const int base = (int)row_of.size();
            for (int m = 0; m < mt; ++m) {
                row_of.push_back(sp_row0[s] + m);
                atom_of.push_back((int)a);
                meta.push_back(ProjMeta{base, m, mt, sp_dm0[s]});
            }
        }
        h.nproj = (int)row_of.size();
        DZ_REQUIRE((size_t)h.nproj * h.npw < ((size_t)1 << 31) && h.npw * kMaxColsPerPass < ((size_t)1 << 31),
                   DZHIP_ERR_ARG, "projector matrix or column block exceeds 2^31 elements");
        DZ_REQUIRE((size_t)h.nproj * sizeof(double2) <= kMaxLdsBytes, DZHIP_ERR_ARG,
                   "more than 10240 projectors per k-point is not supported");
        std::vector<int> block_rows;
        for (int p = 0; p < h.nproj; ++p)
            if (meta[p].m == 0) block_rows.push_back(p);
        h.nblocks = (int)block_rows.size();
        upload(h.block_rows, block_rows);
        if (h.nproj > 0) {
            DevBuf dphi, drow, datom, dg, dpos;
            upload(dphi, phi_all);
            upload(drow, row_of);
            upload(datom, atom_of);
            upload(dg, std::vector<double>(d.gcart, d.gcart + 3 * h.npw));
            upload(dpos, std::vector<double>(d.atom_pos, d.atom_pos + 3 * d.n_atoms));
            const size_t na = (size_t)h.nproj * h.npw;
            h.a64.ensure(na * sizeof(double2));
            h.a32.ensure(na * sizeof(float2));
            hipLaunchKernelGGL(k_build_a, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, h.ctx->stream,
                               h.a64.as<double2>(), dphi.as<double>(), drow.as<int>(),
                               datom.as<int>(), dg.as<double>(), dpos.as<double>(), (int)h.npw,
                               h.nproj);
            DZ_HIP(hipGetLastError());
            convert_to_float(h.a64.as<double2>(), h.a32.as<float2>(), na, h.ctx->stream);
            upload(h.meta, meta);
            upload(h.dm64, dm_all);
            std::vector<float> dm32(dm_all.begin(), dm_all.end());
            upload(h.dm32, dm32);
            DZ_HIP(hipStreamSynchronize(h.ctx->stream));
            dphi.release();
            drow.release();
            datom.release();
            dg.release();
            dpos.release();
        }
    }
    if (d.local_r) ham_set_local(h, d.local_r);
}

// vt[x][z][y] = v[z][y][x]: the x-major copies the register x-line kernel reads.
__global__ void k_local_xmajor(const float* __restrict__ v, float* __restrict__ vt, const double* __restrict__ v64,
                               double* __restrict__ vt64, int nx, int ny, int nz) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (i >= n) return;
    const size_t x = i % nx, line = i / nx;  // line = y + ny z
    const size_t o = x * ((size_t)ny * nz) + line;
    vt[o] = v[i];
    vt64[o] = v64[i];
}

// chat[l][d] = sum_x vt[x][l] w^(d x), w = e^{-2 pi i / nx} (the x axis' twiddle table): the forward
// x-transform of each line l = y + ny z of V_loc/N, the convolution x-line kernel's operand, in FP64
// and FP32.
__global__ void k_line_dft(const double* __restrict__ vt, const double2* __restrict__ tw, int nx, int lines,
                           double2* __restrict__ c64, float2* __restrict__ c32) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)lines * nx) return;
    const int d = (int)(i % nx);
    const size_t l = i / nx;
    double sr = 0.0, si = 0.0;
    int t = 0;
    for (int x = 0; x < nx; ++x) {
        const double v = vt[(size_t)x * lines + l];
        const double2 w = tw[t];
        sr += v * w.x;
        si += v * w.y;
        t += d;
        if (t >= nx) t -= nx;
    }
    c64[i] = make_double2(sr, si);
    c32[i] = make_float2((float)sr, (float)si);
}

static void local_xmajor(Ham& h) {
    h.vloc32_t.ensure(h.ngrid * sizeof(float));
    h.vloc64_t.ensure(h.ngrid * sizeof(double));
    hipLaunchKernelGGL(k_local_xmajor, dim3((unsigned)((h.ngrid + 255) / 256)), dim3(256), 0, h.stream(),
                       h.vloc32.as<float>(), h.vloc32_t.as<float>(), h.vloc64.as<double>(), h.vloc64_t.as<double>(),
                       h.nx, h.ny, h.nz);
    DZ_HIP(hipGetLastError());
    h.chat64.ensure(h.ngrid * sizeof(double2));
    h.chat32.ensure(h.ngrid * sizeof(float2));
    hipLaunchKernelGGL(k_line_dft, dim3((unsigned)((h.ngrid + 255) / 256)), dim3(256), 0, h.stream(),
                       h.vloc64_t.as<double>(), h.tab.tw64 + h.tab.ax[0].tw_off, h.nx, h.ny * h.nz,
                       h.chat64.as<double2>(), h.chat32.as<float2>());
    DZ_HIP(hipGetLastError());
}

void ham_set_local(Ham& h, const double* local_r) {
    std::vector<double> v(h.ngrid);
    const double inv_n = 1.0 / (double)h.ngrid;
    for (size_t i = 0; i < h.ngrid; ++i) v[i] = local_r[i] * inv_n;
    std::vector<float> v32(v.begin(), v.end());
    upload(h.vloc64, v);
    upload(h.vloc32, v32);
    local_xmajor(h);
    DZ_HIP(hipStreamSynchronize(h.stream()));
    h.has_local = true;
}

__global__ void k_scale_local(const double* __restrict__ in, double inv_n, double* __restrict__ v64,
                              float* __restrict__ v32, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = in[i] * inv_n;
    v64[i] = v;
    v32[i] = (float)v;
}

void ham_set_local_dev(Ham& h, const double* local_dev) {
    h.vloc64.ensure(h.ngrid * sizeof(double));
    h.vloc32.ensure(h.ngrid * sizeof(float));
    hipLaunchKernelGGL(k_scale_local, dim3((unsigned)((h.ngrid + 255) / 256)), dim3(256), 0, h.stream(), local_dev,
                       1.0 / (double)h.ngrid, h.vloc64.as<double>(), h.vloc32.as<float>(), h.ngrid);
    DZ_HIP(hipGetLastError());
    local_xmajor(h);
    h.has_local = true;
}

// The second stream and fork/join events of the overlapped apply (created on first use, default
// priority, on the context's device).
void ham_ensure_side_stream(Ham& h) {
    if (h.nl_stream) return;
    DZ_HIP(hipStreamCreateWithFlags(&h.nl_stream, hipStreamNonBlocking));
    DZ_HIP(hipEventCreateWithFlags(&h.ev_fork, hipEventDisableTiming));
    DZ_HIP(hipEventCreateWithFlags(&h.ev_join, hipEventDisableTiming));
}

void ham_release(Ham& h) {
    eig_release(h);
    if (h.nl_stream) {
        (void)hipStreamSynchronize(h.nl_stream);
        (void)hipStreamDestroy(h.nl_stream);
        (void)hipEventDestroy(h.ev_fork);
        (void)hipEventDestroy(h.ev_join);
        h.nl_stream = nullptr;
    }
    if (h.ev_switch) {
        (void)hipEventDestroy(h.ev_switch);
        h.ev_switch = nullptr;
    }
    DevBuf* bufs[] = {&h.plane_start, &h.plane_pw, &h.plane_pos, &h.plane_xpos, &h.plane_ystart,
                      &h.plane_ycols, &h.plane_xabsent, &h.kinetic64,
                      &h.kinetic32, &h.vloc64, &h.vloc32, &h.a64, &h.a32, &h.meta, &h.block_rows, &h.dm64,
                      &h.dm32, &h.xstage, &h.ystage, &h.x32, &h.y32, &h.tbuf, &h.partial, &h.qbuf,
                      &h.dens_rho, &h.dens_w, &h.vloc32_t, &h.vloc64_t, &h.chat64, &h.chat32,
                      &h.plane_zstart, &h.plane_zrows, &h.plane_posr, &h.plane_pw0};
    for (DevBuf* b : bufs) b->release();
    h.tab.release();
    for (auto& set : h.ev_sets)
        for (auto& e : set)
            if (e) (void)hipEventDestroy(e);
    h.ev_sets.clear();
    h.ev_used = 0;
}

template <typename T> struct Sel;
template <> struct Sel<double> {
    static const double* kin(const Ham& h) { return h.kinetic64.as<double>(); }
    static const double* vloc(const Ham& h) { return h.vloc64.as<double>(); }
    static const double2* a(const Ham& h) { return h.a64.as<double2>(); }
    static const double* dm(const Ham& h) { return h.dm64.as<double>(); }
    static const double2* tw(const Ham& h) { return h.tab.tw64; }
};
template <> struct Sel<float> {
    static const float* kin(const Ham& h) { return h.kinetic32.as<float>(); }
    static const float* vloc(const Ham& h) { return h.vloc32.as<float>(); }
    static const float2* a(const Ham& h) { return h.a32.as<float2>(); }
    static const float* dm(const Ham& h) { return h.dm32.as<float>(); }
    static const float2* tw(const Ham& h) { return h.tab.tw32; }
};

// Event set for the current column pass (nullptr when profiling is off).
static hipEvent_t* next_event_set(Ham& h) {
    if (!h.profile) return nullptr;
    if (h.ev_used == h.ev_sets.size()) {
        std::vector<hipEvent_t> set(DZHIP_PHASE_COUNT + 1);
        for (auto& e : set) DZ_HIP(hipEventCreate(&e));
        h.ev_sets.push_back(std::move(set));
    }
    return h.ev_sets[h.ev_used++].data();
}

static void mark(Ham& h, hipEvent_t* set, int i) {
    if (set) DZ_HIP(hipEventRecord(set[i], h.stream()));
}

void Ham::collect_profile() {
    for (size_t k = 0; k < ev_used; ++k) {
        hipEvent_t* set = ev_sets[k].data();
        DZ_HIP(hipEventSynchronize(set[DZHIP_PHASE_COUNT]));
        for (int i = 0; i < DZHIP_PHASE_COUNT; ++i) {
            float ms = 0.f;
            DZ_HIP(hipEventElapsedTime(&ms, set[i], set[i + 1]));
            phase_ms[i] += ms;
        }
    }
    ev_used = 0;
}


template <typename T>
void launch_xline(Ham& h, typename C2T<T>::type* tb, size_t nc, int split, hipStream_t s) {
    using C = typename C2T<T>::type;
    const int ylen = h.ny / split;
    const size_t lds = xline_lds_bytes<C>(h.nx, ylen);
    set_max_dynamic_lds((const void*)k_xline<T>, lds);
    hipLaunchKernelGGL(k_xline<T>, dim3(h.nz, (unsigned)nc, (unsigned)split), dim3(TileThreads<T>::value), lds, s, tb,
                       h.plane_xpos.as<int>(), h.nplanes, h.plane_xabsent.as<int>(), h.nxabsent, Sel<T>::vloc(h), h.nx,
                       h.ny, h.nz, ylen, h.tab.ax[0], Sel<T>::tw(h));
    DZ_HIP(hipGetLastError());
}

// x-line variant launched by code: d >= 1 is the LDS kernel with d y-slabs, -1 / -2 / -3 the
// register kernel, -4 / -5 / -6 the lane-pair and -7 / -8 / -9 the lane-triple register kernel in
// workgroups of 256 / 128 / 64 threads.

template <typename T>
void launch_xline_any(Ham& h, typename C2T<T>::type* tb, size_t nc, int code, hipStream_t s) {
    if (code == kXlineConv) {
        launch_xline_conv<T>(h, tb, nc, s);
        return;
    }
    if (code <= kXlineReg && code >= kXlineRegLast) {
        launch_xline_reg<T>(h, tb, nc, xline_reg_threads(code), xline_reg_kind(code), s);
        return;
    }
    launch_xline<T>(h, tb, nc, code, s);
}

// Process-wide x-line tuning results by (device, grid, sphere x-run, precision): contexts of the
// same shape (the k-points of one SCF, solved concurrently) share one tuning.  The mutex makes
// other contexts of that shape wait for the result instead of tuning too; it does not stop
// contexts that are already tuned from launching work meanwhile.  The trials start from an idle
// device (hipDeviceSynchronize inside the lock drains what is queued), and a caller that runs
// several contexts concurrently tunes each shape once before it starts them (the SCF driver applies
// H to a 16-column block per shape and precision first), so the timings are not taken under load.
std::mutex g_xline_tune_mu;
std::map<std::array<int, 8>, int> g_xline_tune;

// Below this many columns a pass is launch-bound and its timings would not carry over.
constexpr size_t kXlineTuneMinCols = 16;

// x-line variant, chosen once per shape and precision by timing every divisor d of ny with
// ny/d >= 8 on the real shapes and, where they can run, the lane-pair and one-lane register
// kernels in three workgroup sizes each.  The trial launches transform tbuf in place, so this
// runs before plane_inv writes the pass's data into it.  tune_fixed != 0 forces a variant (d, or
// -1..-6 for the register kernels).
template <typename T>
int xline_variant(Ham& h, typename C2T<T>::type* tb, size_t nc, hipStream_t s) {
    constexpr bool fp32 = std::is_same<T, float>::value;
    int& sp = h.xline_split[fp32 ? 1 : 0];
    if (sp != 0) return sp;
    if (h.tune_fixed) {
        // A register variant forced where it cannot run in this precision keeps one slab.
        sp = xline_code_usable(h, h.tune_fixed, !fp32) ? h.tune_fixed : 1;
        return sp;
    }
    std::vector<int> cand;
    // The first candidate is also what narrow passes run before a wide pass has tuned.
    if (xline_pair_usable(h))
        for (int c = -4; c >= -6; --c) cand.push_back(c);
    if (xline_seq3_usable(h))
        for (int c = -10; c >= -12; --c) cand.push_back(c);
    if (xline_tri_usable(h))
        for (int c = -7; c >= -9; --c) cand.push_back(c);
    if (xline_reg_usable(h, !fp32))
        for (int c = -1; c >= -3; --c) cand.push_back(c);
    if (xline_conv_usable(h)) cand.push_back(kXlineConv);
    for (int d = 1; d <= 12; ++d)
        if (h.ny % d == 0 && h.ny / d >= 8) cand.push_back(d);
    if (cand.empty()) cand.push_back(1);
    if (cand.size() == 1) {
        sp = cand[0];
        return sp;
    }
    // Narrow passes run the first candidate without caching it; the next wide pass tunes.
    if (nc < kXlineTuneMinCols) return cand[0];
    const std::array<int, 8> key = {h.ctx->device, h.nx, h.ny, h.nz, h.nplanes, h.plane_xs, 0, fp32 ? 1 : 0};
    std::lock_guard<std::mutex> lk(g_xline_tune_mu);
    const auto it = g_xline_tune.find(key);
    if (it != g_xline_tune.end()) {
        sp = it->second;
        return sp;
    }
    DZ_HIP(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    DZ_HIP(hipEventCreate(&e0));
    DZ_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int best_d = cand[0];
    for (int d : cand) {
        launch_xline_any<T>(h, tb, nc, d, s);
        DZ_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 5; ++r) launch_xline_any<T>(h, tb, nc, d, s);
        DZ_HIP(hipEventRecord(e1, s));
        DZ_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        DZ_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            best_d = d;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    sp = best_d;
    g_xline_tune[key] = best_d;
    return sp;
}

// Plane tables of one context or of a k-point set's concatenated planes, for the plane launchers.
struct PlaneArgs {
    int ny, nz, nplanes, zrow_max;
    const int *start, *pw, *pos, *ystart;
    const short* ycols;
    const int *pw0, *posr, *zstart;
    const short* zrows;
    FftAxis ay, az;
};

static PlaneArgs plane_args(const Ham& h) {
    return PlaneArgs{h.ny, h.nz, h.nplanes, h.zrow_max, h.plane_start.as<int>(), h.plane_pw.as<int>(),
                     h.plane_pos.as<int>(), h.plane_ystart.as<int>(), h.plane_ycols.as<short>(),
                     h.plane_pw0.as<int>(), h.plane_posr.as<int>(), h.plane_zstart.as<int>(),
                     h.plane_zrows.as<short>(), h.tab.ax[1], h.tab.ax[2]};
}

static bool plane_zreg_fits(int ny, int nz, int zrow_max) {
    return xline_tri_supported(nz) && ny <= kZrLinesPerWave * (kZrMaxThreads / 64) && zrow_max > 0 &&
           plane_zr_lds_bytes<double2>(ny, nz, zrow_max) <= kMaxLdsBytes;
}

bool plane_zreg_usable(const Ham& h) { return plane_zreg_fits(h.ny, h.nz, h.zrow_max); }

// Kernel attributes of the plane kernels of `kind` (before any launch a graph may capture).
template <typename T>
void plane_prepare(const PlaneArgs& pa, int kind) {
    using C = typename C2T<T>::type;
    if (kind == kPlaneZreg) {
        const size_t lds = plane_zr_lds_bytes<C>(pa.ny, pa.nz, pa.zrow_max);
        switch (pa.nz) {
#define DZ_ZR_CASE(N)                                                \
    case N:                                                          \
        set_max_dynamic_lds((const void*)k_plane_inv_zr<T, N>, lds); \
        set_max_dynamic_lds((const void*)k_plane_fwd_zr<T, N>, lds); \
        break;
            DZ_XTRI_SIZES(DZ_ZR_CASE)
#undef DZ_ZR_CASE
            default:
                break;
        }
    } else {
        set_max_dynamic_lds((const void*)k_plane_inv<T>, plane_lds_bytes<C>(pa.ny, pa.nz));
        set_max_dynamic_lds((const void*)k_plane_fwd<T>, plane_lds_bytes<C>(pa.ny, pa.nz));
    }
}

// Sphere -> [z][y] planes of every band (kind: kPlaneTile or kPlaneZreg).
template <typename T>
void launch_plane_inv(const PlaneArgs& pa, int kind, const typename C2T<T>::type* x, size_t ld,
                      typename C2T<T>::type* tb, size_t nc, const typename C2T<T>::type* tw, hipStream_t s) {
    using C = typename C2T<T>::type;
    const dim3 grid((unsigned)pa.nplanes, (unsigned)nc);
    if (kind == kPlaneZreg) {
        const size_t lds = plane_zr_lds_bytes<C>(pa.ny, pa.nz, pa.zrow_max);
        const dim3 block((unsigned)plane_zr_threads(pa.ny));
        bool done = false;
        switch (pa.nz) {
#define DZ_ZR_CASE(N)                                                                                         \
    case N:                                                                                                   \
        hipLaunchKernelGGL((k_plane_inv_zr<T, N>), grid, block, lds, s, x, ld, tb, pa.start, pa.pw, pa.pw0,   \
                           pa.posr, pa.zstart, pa.zrows, pa.ny, pa.nplanes, pa.ay, tw);                       \
        done = true;                                                                                          \
        break;
            DZ_XTRI_SIZES(DZ_ZR_CASE)
#undef DZ_ZR_CASE
            default:
                break;
        }
        DZ_REQUIRE(done, DZHIP_ERR_STATE, "no z-register plane kernel for this nz");
    } else {
        hipLaunchKernelGGL(k_plane_inv<T>, grid, dim3(TileThreads<T>::value), plane_lds_bytes<C>(pa.ny, pa.nz), s, x,
                           ld, tb, pa.start, pa.pw, pa.pos, pa.ystart, pa.ycols, pa.ny, pa.nz, pa.nplanes, pa.ay,
                           pa.az, tw);
    }
    DZ_HIP(hipGetLastError());
}

// [z][y] planes -> y (+)= forward transform at the sphere + kinetic * x.
template <typename T>
void launch_plane_fwd(const PlaneArgs& pa, int kind, const typename C2T<T>::type* tb,
                      const typename C2T<T>::type* x, size_t ld, typename C2T<T>::type* y, const T* kin,
                      int accumulate, size_t nc, const typename C2T<T>::type* tw, hipStream_t s) {
    using C = typename C2T<T>::type;
    const dim3 grid((unsigned)pa.nplanes, (unsigned)nc);
    if (kind == kPlaneZreg) {
        const size_t lds = plane_zr_lds_bytes<C>(pa.ny, pa.nz, pa.zrow_max);
        const dim3 block((unsigned)plane_zr_threads(pa.ny));
        bool done = false;
        switch (pa.nz) {
#define DZ_ZR_CASE(N)                                                                                          \
    case N:                                                                                                    \
        hipLaunchKernelGGL((k_plane_fwd_zr<T, N>), grid, block, lds, s, tb, x, ld, y, ld, kin, pa.start, pa.pw, \
                           pa.pw0, pa.posr, pa.zstart, pa.zrows, pa.ny, pa.nplanes, pa.ay, tw, accumulate);     \
        done = true;                                                                                           \
        break;
            DZ_XTRI_SIZES(DZ_ZR_CASE)
#undef DZ_ZR_CASE
            default:
                break;
        }
        DZ_REQUIRE(done, DZHIP_ERR_STATE, "no z-register plane kernel for this nz");
    } else {
        hipLaunchKernelGGL(k_plane_fwd<T>, grid, dim3(TileThreads<T>::value), plane_lds_bytes<C>(pa.ny, pa.nz), s, tb,
                           x, ld, y, ld, kin, pa.start, pa.pw, pa.pos, pa.ystart, pa.ycols, pa.ny, pa.nz, pa.nplanes,
                           pa.ay, pa.az, tw, accumulate);
    }
    DZ_HIP(hipGetLastError());
}

// Process-wide plane-kernel choices by (device, grid, sphere shape, precision), shared like the
// x-line tuning.
std::mutex g_plane_tune_mu;
std::map<std::array<int, 8>, int> g_plane_tune;

// Plane kernels, chosen once per shape and precision by timing plane_inv + plane_fwd of each kind
// that can run on the pass's real shape.  The trials write tbuf and y (not accumulating), so this
// runs before plane_inv and before the nonlocal chain writes y.
template <typename T>
int plane_variant(Ham& h, const typename C2T<T>::type* x, typename C2T<T>::type* y, typename C2T<T>::type* tb,
                  size_t nc, hipStream_t s) {
    constexpr bool fp32 = std::is_same<T, float>::value;
    int& kind = h.plane_kind[fp32 ? 1 : 0];
    if (kind != 0) return kind;
    const bool zr = plane_zreg_usable(h);
    if (h.plane_fixed != 0) {
        kind = h.plane_fixed == kPlaneZreg && !zr ? kPlaneTile : h.plane_fixed;
        return kind;
    }
    if (!zr) {
        kind = kPlaneTile;
        return kind;
    }
    if (nc < kXlineTuneMinCols) return kPlaneZreg;
    const std::array<int, 8> key = {h.ctx->device, h.nx, h.ny, h.nz, h.nplanes, h.zrow_max, (int)h.npw, fp32 ? 1 : 0};
    std::lock_guard<std::mutex> lk(g_plane_tune_mu);
    const auto it = g_plane_tune.find(key);
    if (it != g_plane_tune.end()) {
        kind = it->second;
        return kind;
    }
    DZ_HIP(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    DZ_HIP(hipEventCreate(&e0));
    DZ_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int best_k = kPlaneZreg;
    const PlaneArgs pa = plane_args(h);
    const typename C2T<T>::type* tw = Sel<T>::tw(h);
    for (int k : {kPlaneZreg, kPlaneTile}) {
        plane_prepare<T>(pa, k);
        launch_plane_inv<T>(pa, k, x, h.npw, tb, nc, tw, s);
        launch_plane_fwd<T>(pa, k, tb, x, h.npw, y, Sel<T>::kin(h), 0, nc, tw, s);
        DZ_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r) {
            launch_plane_inv<T>(pa, k, x, h.npw, tb, nc, tw, s);
            launch_plane_fwd<T>(pa, k, tb, x, h.npw, y, Sel<T>::kin(h), 0, nc, tw, s);
        }
        DZ_HIP(hipEventRecord(e1, s));
        DZ_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        DZ_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            best_k = k;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    kind = best_k;
    g_plane_tune[key] = best_k;
    return kind;
}

template <typename T>
void ham_apply_dev(Ham& h, const void* xv, void* yv, size_t ncols, bool local, bool nonlocal) {
    using C = typename C2T<T>::type;
    const C* x = reinterpret_cast<const C*>(xv);
    C* y = reinterpret_cast<C*>(yv);
    if (ncols == 0) return;
    DZ_REQUIRE(!local || h.has_local, DZHIP_ERR_STATE, "local potential not set");
    hipStream_t s = h.stream();
    const size_t ld = h.npw;
    const bool do_nl = nonlocal && h.nproj > 0;

    for (size_t c0 = 0; c0 < ncols; c0 += kMaxColsPerPass) {
        const size_t nc = std::min(kMaxColsPerPass, ncols - c0);
        const C* xc = x + c0 * ld;
        C* yc = y + c0 * ld;
        // Launch shapes and buffers first: nothing may reallocate once two streams run.
        constexpr int kNlP = nl_p_tile<T>();
        const int ptiles = (h.nproj + kNlP - 1) / kNlP;
        const int btiles = (int)((nc + kNlBands - 1) / kNlBands);
        const int ktiles = (int)((h.npw + kNlK - 1) / kNlK);
        // Split K so that about one workgroup per CU runs; each K-chunk a multiple of kNlK.
        int nks = std::max(1, std::min(ktiles, (256 + ptiles * btiles - 1) / std::max(1, ptiles * btiles)));
        const int kchunk = (int)round_up((h.npw + nks - 1) / nks, kNlK);
        nks = (int)((h.npw + kchunk - 1) / kchunk);
        if (do_nl) {
            h.partial.ensure((size_t)nks * nc * h.nproj * sizeof(C));
            h.qbuf.ensure(nc * h.nproj * sizeof(C));
        }
        C* tb = nullptr;
        int split = 1, pkind = kPlaneTile;
        if (local) {
            h.tbuf.ensure(nc * h.nplanes * (size_t)h.ny * h.nz * sizeof(C));
            tb = h.tbuf.as<C>();
            // Tuned before plane_inv fills tbuf (the trial launches transform tbuf in place) and
            // before the nonlocal kernels start, so the trials run alone.
            split = xline_variant<T>(h, tb, nc, s);
            pkind = plane_variant<T>(h, xc, yc, tb, nc, s);
        }
        // The nonlocal chain (matrix cores) and plane_inv + x-line (memory and latency bound) are
        // independent: they run on two streams and join before plane_fwd, which adds the local
        // and kinetic terms to the nonlocal y.  Profiling keeps them in order on one stream so
        // each phase is timed alone.
        const bool overlap = do_nl && local && h.overlap && !h.profile;
        hipStream_t snl = s;
        if (overlap) {
            ham_ensure_side_stream(h);
            snl = h.nl_stream;
            DZ_HIP(hipEventRecord(h.ev_fork, s));
            DZ_HIP(hipStreamWaitEvent(snl, h.ev_fork, 0));
        }
        hipEvent_t* evs = next_event_set(h);
        mark(h, evs, 0);
        if (do_nl) {
            const size_t lds = nl_proj_lds_bytes<T>();
            set_max_dynamic_lds((const void*)k_nl_proj<T>, lds);
            hipLaunchKernelGGL(k_nl_proj<T>, dim3(ptiles * nks, 1, btiles), dim3(kNlThreads), lds, snl,
                               Sel<T>::a(h), xc, ld, h.partial.as<C>(), h.nproj, (int)h.npw, (int)nc,
                               kchunk, ptiles, (const int2*)nullptr, (int)h.npw);
            DZ_HIP(hipGetLastError());
            const size_t lds_d = dij_lds_bytes(h.nproj, sizeof(C));
            set_max_dynamic_lds((const void*)k_nl_dij<T>, lds_d);
            hipLaunchKernelGGL(k_nl_dij<T>, dim3((unsigned)nc), dim3(kDijThreads), lds_d, snl, h.partial.as<C>(),
                               nks, h.qbuf.as<C>(), h.meta.as<ProjMeta>(), Sel<T>::dm(h), h.nproj, (int)nc,
                               (T)h.inv_volume, dij_split(h.nproj, sizeof(C)), (const int2*)nullptr);
            DZ_HIP(hipGetLastError());
        }
        mark(h, evs, 1);
        if (do_nl) {
            const size_t lds = nl_back_lds_bytes<T>();
            constexpr int kNlG = nl_g_tile<T>();
            set_max_dynamic_lds((const void*)k_nl_back<T>, lds);
            hipLaunchKernelGGL(k_nl_back<T>, dim3((unsigned)((h.npw + kNlG - 1) / kNlG), 1,
                                                  (unsigned)((nc + kNlBands - 1) / kNlBands)),
                               dim3(kNlThreads), lds, snl, Sel<T>::a(h), h.qbuf.as<C>(), yc, ld, h.nproj,
                               (int)h.npw, (int)nc, (const int4*)nullptr, (int)h.npw);
            DZ_HIP(hipGetLastError());
        }
        if (overlap) DZ_HIP(hipEventRecord(h.ev_join, snl));
        mark(h, evs, 2);
        if (local) {
            const PlaneArgs pa = plane_args(h);
            plane_prepare<T>(pa, pkind);
            launch_plane_inv<T>(pa, pkind, xc, ld, tb, nc, Sel<T>::tw(h), s);
            mark(h, evs, 3);
            launch_xline_any<T>(h, tb, nc, split, s);
            mark(h, evs, 4);
            if (overlap) DZ_HIP(hipStreamWaitEvent(s, h.ev_join, 0));
            launch_plane_fwd<T>(pa, pkind, tb, xc, ld, yc, Sel<T>::kin(h), do_nl ? 1 : 0, nc, Sel<T>::tw(h), s);
        } else {
            if (!do_nl) {
                const size_t n = nc * ld;
                hipLaunchKernelGGL(k_fill_zero<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, yc, n);
                DZ_HIP(hipGetLastError());
            }
            mark(h, evs, 3);
            mark(h, evs, 4);
        }
        mark(h, evs, 5);
        // Bound the number of outstanding event sets.
        if (h.profile && h.ev_used >= 4096) h.collect_profile();
    }
    if (h.profile) h.calls++;
}

void ham_density_dev(Ham& h, const double2* psi, size_t nbands, const double* wts_dev, double* rho_dev,
                     bool accumulate) {
    hipStream_t s = h.stream();
    DZ_REQUIRE((size_t)h.nx * h.ny <= (size_t)kDensThreads * kDensMaxPts, DZHIP_ERR_ARG,
               "x*y plane too large for the density kernel");
    const size_t plane = (size_t)h.ny * h.nz;
    const size_t lds_p = plane_lds_bytes<double2>(h.ny, h.nz);
    const size_t lds_x = xline_lds_bytes<double2>(h.nx, h.ny);
    set_max_dynamic_lds((const void*)k_plane_inv<double>, lds_p);
    set_max_dynamic_lds((const void*)k_xline_density, lds_x);
    // Band groups: enough workgroups to fill the device, deterministic sum afterwards.
    const int ngroups = (int)std::max<size_t>(1, std::min<size_t>(nbands, (512 + h.nz - 1) / h.nz));
    const int per_group = (int)((nbands + ngroups - 1) / ngroups);
    h.partial.ensure((size_t)ngroups * h.ngrid * sizeof(double));
    for (size_t c0 = 0; c0 < nbands; c0 += kMaxColsPerPass) {
        const size_t nc = std::min(kMaxColsPerPass, nbands - c0);
        h.tbuf.ensure(nc * h.nplanes * plane * sizeof(double2));
        hipLaunchKernelGGL(k_plane_inv<double>, dim3(h.nplanes, (unsigned)nc), dim3(TileThreads<double>::value),
                           lds_p, s, psi + c0 * h.npw, h.npw, h.tbuf.as<double2>(), h.plane_start.as<int>(),
                           h.plane_pw.as<int>(), h.plane_pos.as<int>(), h.plane_ystart.as<int>(),
                           h.plane_ycols.as<short>(), h.ny, h.nz, h.nplanes, h.tab.ax[1], h.tab.ax[2], h.tab.tw64);
        DZ_HIP(hipGetLastError());
        const int ng = (int)std::max<size_t>(1, std::min<size_t>(nc, (size_t)ngroups));
        const int pg = (int)((nc + ng - 1) / ng);
        hipLaunchKernelGGL(k_xline_density, dim3(h.nz, ng), dim3(kDensThreads), lds_x, s, h.tbuf.as<double2>(),
                           h.plane_xpos.as<int>(), h.nplanes, h.plane_xabsent.as<int>(), h.nxabsent, wts_dev + c0, (int)nc, pg, h.nx, h.ny, h.nz,
                           h.tab.ax[0], h.tab.tw64, h.inv_volume, h.partial.as<double>());
        DZ_HIP(hipGetLastError());
        const bool acc = accumulate || c0 > 0;
        hipLaunchKernelGGL(k_density_sum, dim3((unsigned)((h.ngrid + 255) / 256)), dim3(256), 0, s,
                           h.partial.as<double>(), ng, h.ngrid, rho_dev, acc ? 1 : 0);
        DZ_HIP(hipGetLastError());
    }
}

template void ham_apply_dev<double>(Ham&, const void*, void*, size_t, bool, bool);
template void ham_apply_dev<float>(Ham&, const void*, void*, size_t, bool, bool);

// ----------------------------------------------------------------- batched multi-k apply
// All k-points of a set as one sphere: per band, the plane waves of every k-point follow each
// other (k-point k at offset npw_off[k], band stride npw_total) and so do their x-planes, so the
// yz-plane kernels run unchanged over the concatenated plane tables; the x-line kernels take each
// k-point's plane range (blockIdx.z), and the nonlocal kernels each k-point's K-chunks, D block and
// G tiles from small tables.  Every per-k quantity keeps the per-context split (K-chunk boundaries,
// G tiles, x-line kernel), so the set's result equals the per-context applies bit for bit.

template <typename V>
static std::vector<V> download(const DevBuf& b, size_t n) {
    std::vector<V> v(n);
    if (n) DZ_HIP(hipMemcpy(v.data(), b.ptr, n * sizeof(V), hipMemcpyDeviceToHost));
    return v;
}

void hamset_build(HamSet& hs) {
    DZ_REQUIRE(!hs.ks.empty(), DZHIP_ERR_ARG, "empty k-point set");
    const Ham& h0 = *hs.ks[0];
    hs.nx = h0.nx;
    hs.ny = h0.ny;
    hs.nz = h0.nz;
    hs.nproj = h0.nproj;
    const int nk = (int)hs.ks.size();
    std::vector<int> start{0}, pw, pos, ystart{0}, kpl, pw0, posr, zstart{0};
    std::vector<short> ycols, zrows;
    hs.zrow_max = 0;
    hs.npw_off.assign(nk + 1, 0);
    hs.np_off.assign(nk + 1, 0);
    hs.max_planes = 0;
    for (int k = 0; k < nk; ++k) {
        const Ham& h = *hs.ks[k];
        DZ_REQUIRE(h.ctx->device == h0.ctx->device, DZHIP_ERR_ARG, "k-point contexts on different devices");
        DZ_REQUIRE(h.nx == h0.nx && h.ny == h0.ny && h.nz == h0.nz, DZHIP_ERR_SIZE, "k-points on different grids");
        DZ_REQUIRE(h.nproj == h0.nproj && h.nblocks == h0.nblocks, DZHIP_ERR_ARG,
                   "k-points with different projector sets");
        DZ_REQUIRE(h.has_local, DZHIP_ERR_STATE, "local potential not set on every k-point");
        const size_t off = hs.npw_off[k];
        const std::vector<int> st = download<int>(h.plane_start, h.nplanes + 1);
        const std::vector<int> pk = download<int>(h.plane_pw, h.npw);
        const std::vector<int> qk = download<int>(h.plane_pos, h.npw);
        const std::vector<int> ys = download<int>(h.plane_ystart, h.nplanes + 1);
        const std::vector<short> yc = download<short>(h.plane_ycols, (size_t)ys[h.nplanes]);
        const std::vector<int> w0 = download<int>(h.plane_pw0, h.nplanes);
        const std::vector<int> rk = download<int>(h.plane_posr, h.npw);
        const std::vector<int> zs = download<int>(h.plane_zstart, h.nplanes + 1);
        const std::vector<short> zr = download<short>(h.plane_zrows, (size_t)zs[h.nplanes]);
        const int e0 = start.back(), c0 = ystart.back(), r0 = zstart.back();
        for (int p = 1; p <= h.nplanes; ++p) {
            start.push_back(e0 + st[p]);
            ystart.push_back(c0 + ys[p]);
            zstart.push_back(r0 + zs[p]);
        }
        for (int p = 0; p < h.nplanes; ++p) pw0.push_back(w0[p] >= 0 ? (int)(off + w0[p]) : -1);
        for (size_t j = 0; j < h.npw; ++j) {
            pw.push_back((int)(off + pk[j]));
            pos.push_back(qk[j]);
            posr.push_back(rk[j]);
        }
        ycols.insert(ycols.end(), yc.begin(), yc.end());
        zrows.insert(zrows.end(), zr.begin(), zr.end());
        hs.zrow_max = std::max(hs.zrow_max, h.zrow_max);
        kpl.push_back(hs.np_off[k]);
        kpl.push_back(h.nplanes);
        hs.npw_off[k + 1] = off + h.npw;
        hs.np_off[k + 1] = hs.np_off[k] + h.nplanes;
        hs.max_planes = std::max(hs.max_planes, h.nplanes);
    }
    hs.npw_total = hs.npw_off[nk];
    hs.np_total = hs.np_off[nk];
    DZ_REQUIRE(hs.npw_total < ((size_t)1 << 31) && hs.npw_total * kMaxColsPerPass < ((size_t)1 << 31) * 8,
               DZHIP_ERR_SIZE, "k-point set too large");
    upload(hs.plane_start, start);
    upload(hs.plane_pw, pw);
    upload(hs.plane_pos, pos);
    upload(hs.plane_ystart, ystart);
    upload(hs.plane_ycols, ycols);
    upload(hs.plane_pw0, pw0);
    upload(hs.plane_posr, posr);
    upload(hs.plane_zstart, zstart);
    upload(hs.plane_zrows, zrows);
    upload(hs.kpl, kpl);
    hs.kinetic64.ensure(hs.npw_total * sizeof(double));
    hs.kinetic32.ensure(hs.npw_total * sizeof(float));
    if (hs.nproj > 0) {
        hs.a64.ensure((size_t)hs.nproj * hs.npw_total * sizeof(double2));
        hs.a32.ensure((size_t)hs.nproj * hs.npw_total * sizeof(float2));
    }
    for (int k = 0; k < nk; ++k) {
        const Ham& h = *hs.ks[k];
        const size_t off = hs.npw_off[k];
        DZ_HIP(hipMemcpy(hs.kinetic64.as<double>() + off, h.kinetic64.ptr, h.npw * sizeof(double),
                         hipMemcpyDeviceToDevice));
        DZ_HIP(hipMemcpy(hs.kinetic32.as<float>() + off, h.kinetic32.ptr, h.npw * sizeof(float),
                         hipMemcpyDeviceToDevice));
        if (hs.nproj > 0) {
            DZ_HIP(hipMemcpy2D(hs.a64.as<double2>() + off, hs.npw_total * sizeof(double2), h.a64.ptr,
                               h.npw * sizeof(double2), h.npw * sizeof(double2), hs.nproj, hipMemcpyDeviceToDevice));
            DZ_HIP(hipMemcpy2D(hs.a32.as<float2>() + off, hs.npw_total * sizeof(float2), h.a32.ptr,
                               h.npw * sizeof(float2), h.npw * sizeof(float2), hs.nproj, hipMemcpyDeviceToDevice));
        }
    }
    hs.seg_cols = 0;
}

// Nonlocal segment tables for ncols bands: per k-point the K-chunks a single context's apply of
// that many columns would use, their ranges in the concatenated G axis, and the G tiles.
template <typename T>
static void hamset_segments(HamSet& hs, size_t nc) {
    if (hs.seg_cols == nc && hs.seg_prec == (int)sizeof(T)) return;
    constexpr int kNlP = nl_p_tile<T>();
    constexpr int kNlG = nl_g_tile<T>();
    const int ptiles = (hs.nproj + kNlP - 1) / kNlP;
    const int btiles = (int)((nc + kNlBands - 1) / kNlBands);
    std::vector<int> chunks, kseg;
    std::vector<int> tiles;
    for (size_t k = 0; k < hs.ks.size(); ++k) {
        const int npw = (int)hs.ks[k]->npw;
        const int off = (int)hs.npw_off[k];
        // The same split as ham_apply_dev for this k-point alone.
        const int ktiles = (npw + kNlK - 1) / kNlK;
        int nks = std::max(1, std::min(ktiles, (256 + ptiles * btiles - 1) / std::max(1, ptiles * btiles)));
        const int kchunk = (int)round_up((npw + nks - 1) / nks, kNlK);
        nks = (npw + kchunk - 1) / kchunk;
        kseg.push_back((int)chunks.size() / 2);
        for (int c = 0; c < nks; ++c) {
            chunks.push_back(off + c * kchunk);
            chunks.push_back(off + std::min(npw, (c + 1) * kchunk));
        }
        kseg.push_back((int)chunks.size() / 2);
        for (int g = 0; g < npw; g += kNlG) {
            tiles.push_back(off + g);
            tiles.push_back(off + npw);
            tiles.push_back((int)k);
            tiles.push_back(0);
        }
    }
    upload(hs.chunks, chunks);
    upload(hs.kseg, kseg);
    upload(hs.tiles, tiles);
    hs.nks_total = (int)chunks.size() / 2;
    hs.ntiles = (int)tiles.size() / 4;
    hs.seg_cols = nc;
    hs.seg_prec = (int)sizeof(T);
}

// x-line kernel of the set: the register variant member 0 runs for this precision (tuned on its
// first wide apply), which must be usable on every member (same grid; their plane runs).
template <typename T>
static int hamset_xline_code(HamSet& hs, size_t nc, hipStream_t s) {
    using C = typename C2T<T>::type;
    constexpr bool fp32 = std::is_same<T, float>::value;
    Ham& h0 = *hs.ks[0];
    int code = h0.tune_fixed;
    if (code == 0) {
        code = h0.xline_split[fp32 ? 1 : 0];
        if (code == 0) {
            h0.tbuf.ensure(nc * h0.nplanes * (size_t)h0.ny * h0.nz * sizeof(C));
            code = xline_variant<T>(h0, h0.tbuf.as<C>(), nc, s);
        }
    }
    auto usable_all = [&](int c) {
        if (c > 0) return false;
        for (Ham* h : hs.ks)
            if (!xline_code_usable(*h, c, !fp32)) return false;
        return true;
    };
    if (!usable_all(code)) {
        code = 0;
        for (int c = kXlineReg; c >= kXlineRegLast && code == 0; --c)
            if (usable_all(c)) code = c;
    }
    DZ_REQUIRE(code != 0, DZHIP_ERR_ARG, "no register x-line kernel runs on every k-point of the set");
    return code;
}

static PlaneArgs plane_args(const HamSet& hs) {
    const Ham& h0 = *hs.ks[0];
    return PlaneArgs{hs.ny, hs.nz, hs.np_total, hs.zrow_max, hs.plane_start.as<int>(), hs.plane_pw.as<int>(),
                     hs.plane_pos.as<int>(), hs.plane_ystart.as<int>(), hs.plane_ycols.as<short>(),
                     hs.plane_pw0.as<int>(), hs.plane_posr.as<int>(), hs.plane_zstart.as<int>(),
                     hs.plane_zrows.as<short>(), h0.tab.ax[1], h0.tab.ax[2]};
}

// Plane kernels of the set: member 0's forced or tuned choice for this precision; untuned, the
// z-register kernels where they run.
template <typename T>
static int hamset_plane_kind(const HamSet& hs) {
    constexpr bool fp32 = std::is_same<T, float>::value;
    const Ham& h0 = *hs.ks[0];
    int kind = h0.plane_fixed != 0 ? h0.plane_fixed : h0.plane_kind[fp32 ? 1 : 0];
    const bool zr = plane_zreg_fits(hs.ny, hs.nz, hs.zrow_max);
    if (kind == 0) kind = zr ? kPlaneZreg : kPlaneTile;
    return kind == kPlaneZreg && !zr ? kPlaneTile : kind;
}

// The apply's launch sequence for the whole set (no host synchronisation, no allocation: it is
// what the graph captures).
template <typename T>
static void hamset_launch(HamSet& hs, const void* xv, void* yv, size_t nc, int code, int pkind, hipStream_t s) {
    using C = typename C2T<T>::type;
    const C* x = reinterpret_cast<const C*>(xv);
    C* y = reinterpret_cast<C*>(yv);
    Ham& h0 = *hs.ks[0];
    const size_t ld = hs.npw_total;
    const int nk = (int)hs.ks.size();
    const bool do_nl = hs.nproj > 0;
    const C* tw = Sel<T>::tw(h0);
    C* tb = hs.tbuf.as<C>();
    const C* a = std::is_same<T, float>::value ? (const C*)hs.a32.ptr : (const C*)hs.a64.ptr;
    const T* kin = std::is_same<T, float>::value ? (const T*)hs.kinetic32.ptr : (const T*)hs.kinetic64.ptr;
    if (do_nl) {
        constexpr int kNlP = nl_p_tile<T>();
        const int ptiles = (hs.nproj + kNlP - 1) / kNlP;
        const int btiles = (int)((nc + kNlBands - 1) / kNlBands);
        hipLaunchKernelGGL(k_nl_proj<T>, dim3(ptiles * hs.nks_total, 1, btiles), dim3(kNlThreads),
                           nl_proj_lds_bytes<T>(), s, a, x, ld, hs.partial.as<C>(), hs.nproj, (int)ld, (int)nc, 0,
                           ptiles, hs.chunks.as<int2>(), (int)ld);
        DZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_nl_dij<T>, dim3((unsigned)nc, (unsigned)nk), dim3(kDijThreads),
                           dij_lds_bytes(hs.nproj, sizeof(C)), s, hs.partial.as<C>(), hs.nks_total, hs.qbuf.as<C>(),
                           h0.meta.as<ProjMeta>(), Sel<T>::dm(h0), hs.nproj, (int)nc, (T)h0.inv_volume,
                           dij_split(hs.nproj, sizeof(C)), hs.kseg.as<int2>());
        DZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_nl_back<T>, dim3((unsigned)hs.ntiles, 1, (unsigned)((nc + kNlBands - 1) / kNlBands)),
                           dim3(kNlThreads), nl_back_lds_bytes<T>(), s, a, hs.qbuf.as<C>(), y, ld, hs.nproj, (int)ld,
                           (int)nc, hs.tiles.as<int4>(), (int)ld);
        DZ_HIP(hipGetLastError());
    }
    const PlaneArgs pa = plane_args(hs);
    launch_plane_inv<T>(pa, pkind, x, ld, tb, nc, tw, s);
    if (code == kXlineConv)
        launch_xline_conv<T>(h0, tb, nc, s, hs.kpl.as<int2>(), nk, hs.np_total, hs.max_planes);
    else
        launch_xline_reg<T>(h0, tb, nc, xline_reg_threads(code), xline_reg_kind(code), s, hs.kpl.as<int2>(), nk,
                            hs.np_total, hs.max_planes);
    launch_plane_fwd<T>(pa, pkind, tb, x, ld, y, kin, do_nl ? 1 : 0, nc, tw, s);
}

template <typename T>
void hamset_apply_dev(HamSet& hs, const void* x, void* y, size_t nc, bool use_graph) {
    using C = typename C2T<T>::type;
    if (nc == 0) return;
    DZ_REQUIRE(nc <= kMaxColsPerPass, DZHIP_ERR_ARG, "batched apply: at most 256 columns");
    hipStream_t s = hs.stream();
    Ham& h0 = *hs.ks[0];
    // Everything that allocates, synchronises or sets kernel attributes happens here, before
    // the launch sequence (which a graph may capture).
    const int code = hamset_xline_code<T>(hs, nc, s);
    hamset_segments<T>(hs, nc);
    hs.tbuf.ensure(nc * (size_t)hs.np_total * hs.ny * hs.nz * sizeof(C));
    if (hs.nproj > 0) {
        hs.partial.ensure((size_t)hs.nks_total * nc * hs.nproj * sizeof(C));
        hs.qbuf.ensure(hs.ks.size() * nc * hs.nproj * sizeof(C));
        set_max_dynamic_lds((const void*)k_nl_proj<T>, nl_proj_lds_bytes<T>());
        set_max_dynamic_lds((const void*)k_nl_dij<T>, dij_lds_bytes(hs.nproj, sizeof(C)));
        set_max_dynamic_lds((const void*)k_nl_back<T>, nl_back_lds_bytes<T>());
    }
    const int pkind = hamset_plane_kind<T>(hs);
    plane_prepare<T>(plane_args(hs), pkind);
    (void)h0;
    if (!use_graph) {
        hamset_launch<T>(hs, x, y, nc, code, pkind, s);
        return;
    }
    const HamSet::GraphKey key{x, y, nc, (int)sizeof(T), code, pkind, s, hs.tbuf.ptr, hs.partial.ptr, hs.qbuf.ptr};
    if (!(hs.graph_exec && hs.graph_key == key)) {
        if (hs.graph_exec) {
            DZ_HIP(hipGraphExecDestroy(hs.graph_exec));
            hs.graph_exec = nullptr;
        }
        hipGraph_t g = nullptr;
        DZ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            hamset_launch<T>(hs, x, y, nc, code, pkind, s);
        } catch (...) {
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        DZ_HIP(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&hs.graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        DZ_HIP(e);
        hs.graph_key = key;
        hs.graph_builds++;
    }
    DZ_HIP(hipGraphLaunch(hs.graph_exec, s));
}

void hamset_release(HamSet& hs) {
    if (hs.graph_exec) {
        (void)hipGraphExecDestroy(hs.graph_exec);
        hs.graph_exec = nullptr;
    }
    DevBuf* bufs[] = {&hs.plane_start, &hs.plane_pw, &hs.plane_pos, &hs.plane_ystart, &hs.plane_ycols, &hs.kpl,
                      &hs.plane_pw0, &hs.plane_posr, &hs.plane_zstart, &hs.plane_zrows,
                      &hs.kinetic64, &hs.kinetic32, &hs.a64, &hs.a32, &hs.chunks, &hs.kseg, &hs.tiles, &hs.tbuf,
                      &hs.partial, &hs.qbuf};
    for (DevBuf* b : bufs) b->release();
}

template void hamset_apply_dev<double>(HamSet&, const void*, void*, size_t, bool);
template void hamset_apply_dev<float>(HamSet&, const void*, void*, size_t, bool);

}  // namespace dzhip
