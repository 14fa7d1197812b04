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
//               write back only the rows of sphere x-planes.
//   plane_fwd : one workgroup per (plane, band): forward FFT along y and z, gather the
//               sphere points, add |k+G|^2 x.
// HBM traffic per band: psi + H psi on the sphere and two round trips of half a grid.
//
// Nonlocal part: complex projector matrix A[p][G] = phi_p(G) e^{+iG.tau_a} held on the
// device; proj = A x (split-K), q = D proj / Omega per atom block, y += A^H q.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fft_lds.h"
#include "ham.h"

namespace dzhip {

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
    for (int i = tid; i < nz * pitch; i += nthr) buf[i] = C{T(0), T(0)};
    load_twiddles(twy, tw + ay.tw_off, ny, tid, nthr);
    load_twiddles(twz, tw + az.tw_off, nz, tid, nthr);
    const int c0 = ystart[p], nact = ystart[p + 1] - c0;
    for (int i = tid; i < nact; i += nthr) cols[i] = ycols[c0 + i];
    __syncthreads();
    const C* xc = x + col * ldx;
    const int j1 = pstart[p + 1];
    for (int j = pstart[p] + tid; j < j1; j += nthr) buf[ppos[j]] = xc[ppw[j]];
    __syncthreads();
    // Inverse (e^{+i}) along z, only on the y-columns that hold plane waves.
    lds_fft_lines<T, C, true>(buf, az, twz, LineSet{nact, pitch, 0, cols}, T(1), tid, nthr);
    // Inverse along y on every z-row.
    lds_fft_lines<T, C, true>(buf, ay, twy, LineSet{nz, 1, pitch, nullptr}, T(1), tid, nthr);
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
    const T* __restrict__ vloc, int nx, int ny, int nz, FftAxis ax,
    const typename C2T<T>::type* __restrict__ tw) {
    using C = typename C2T<T>::type;
    extern __shared__ unsigned char smem_raw[];
    const int pitch = ny + 1;
    C* buf = reinterpret_cast<C*>(smem_raw);  // [x][y]: element (x, y) at y + pitch*x
    C* twx = buf + nx * pitch;
    const int z = blockIdx.x;
    const size_t col = blockIdx.y;
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int i = tid; i < nx * pitch; i += nthr) buf[i] = C{T(0), T(0)};
    load_twiddles(twx, tw + ax.tw_off, nx, tid, nthr);
    __syncthreads();
    C* tcol = tbuf + col * (size_t)nplanes * ny * nz;
    const unsigned my = div_magic(ny);
    for (int i = tid; i < nplanes * ny; i += nthr) {
        const int p = fdiv(i, ny, my), y = i - p * ny;
        buf[y + pitch * pxpos[p]] = tcol[((size_t)p * nz + z) * ny + y];
    }
    __syncthreads();
    const LineSet xl{ny, pitch, 1, nullptr};
    lds_fft_lines<T, C, true>(buf, ax, twx, xl, T(1), tid, nthr);
    const T* vz = vloc + (size_t)nx * ny * z;
    const unsigned mx = div_magic(nx);
    for (int i = tid; i < nx * ny; i += nthr) {
        const int y = fdiv(i, nx, mx), xx = i - y * nx;
        C& e = buf[y + pitch * xx];
        const T v = vz[i];
        e = C{e.x * v, e.y * v};
    }
    __syncthreads();
    lds_fft_lines<T, C, false>(buf, ax, twx, xl, T(-1), tid, nthr);
    for (int i = tid; i < nplanes * ny; i += nthr) {
        const int p = fdiv(i, ny, my), y = i - p * ny;
        tcol[((size_t)p * nz + z) * ny + y] = buf[y + pitch * pxpos[p]];
    }
}

template <typename T>
__global__ __launch_bounds__(TileThreads<T>::value) void k_plane_fwd(
    const typename C2T<T>::type* __restrict__ tbuf, const typename C2T<T>::type* __restrict__ x,
    size_t ldx, typename C2T<T>::type* __restrict__ y, size_t ldy, const T* __restrict__ kinetic,
    const int* __restrict__ pstart, const int* __restrict__ ppw, const int* __restrict__ ppos,
    const int* __restrict__ ystart, const short* __restrict__ ycols, int ny, int nz, int nplanes,
    FftAxis ay, FftAxis az, const typename C2T<T>::type* __restrict__ tw, int add_kinetic) {
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
    load_twiddles(twy, tw + ay.tw_off, ny, tid, nthr);
    load_twiddles(twz, tw + az.tw_off, nz, tid, nthr);
    const int c0 = ystart[p], nact = ystart[p + 1] - c0;
    for (int i = tid; i < nact; i += nthr) cols[i] = ycols[c0 + i];
    const C* in = tbuf + (col * nplanes + p) * (size_t)ny * nz;
    const unsigned my = div_magic(ny);
    for (int i = tid; i < ny * nz; i += nthr) {
        const int z = fdiv(i, ny, my), yy = i - z * ny;
        buf[yy + pitch * z] = in[i];
    }
    __syncthreads();
    // Forward along y on every z-row (frequency y lands at py[y]), then along z only on the
    // columns whose frequencies are gathered.
    lds_fft_lines<T, C, false>(buf, ay, twy, LineSet{nz, 1, pitch, nullptr}, T(-1), tid, nthr);
    lds_fft_lines<T, C, false>(buf, az, twz, LineSet{nact, pitch, 0, cols}, T(-1), tid, nthr);
    const C* xc = x + col * ldx;
    C* yc = y + col * ldy;
    const int j1 = pstart[p + 1];
    for (int j = pstart[p] + tid; j < j1; j += nthr) {
        const int g = ppw[j];
        C v = buf[ppos[j]];
        if (add_kinetic) {
            const T kin = kinetic[g];
            const C xv = xc[g];
            v = C{v.x + kin * xv.x, v.y + kin * xv.y};
        }
        yc[g] = v;
    }
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

constexpr int kProjTileP = 16, kProjTileB = 16, kProjTileK = 64;

template <typename T>
__global__ __launch_bounds__(256) void k_nl_proj(const typename C2T<T>::type* __restrict__ a,
                                                 const typename C2T<T>::type* __restrict__ x,
                                                 size_t ldx, typename C2T<T>::type* __restrict__ partial,
                                                 int P, int npw, int ncols, int kchunk) {
    using C = typename C2T<T>::type;
    __shared__ C as[kProjTileP][kProjTileK + 1];
    __shared__ C xs[kProjTileB][kProjTileK + 1];
    const int p0 = blockIdx.x * kProjTileP, b0 = blockIdx.y * kProjTileB, ks = blockIdx.z;
    const int k0 = ks * kchunk;
    const int k1 = min(npw, k0 + kchunk);
    const int tid = threadIdx.x;
    const int tp = tid & 15, tb = tid >> 4;
    C acc = C{T(0), T(0)};
    for (int kt = k0; kt < k1; kt += kProjTileK) {
        for (int e = tid; e < kProjTileP * kProjTileK; e += 256) {
            const int r = e / kProjTileK, kk = e - r * kProjTileK, k = kt + kk;
            as[r][kk] = (p0 + r < P && k < k1) ? a[(size_t)(p0 + r) * npw + k] : C{T(0), T(0)};
            xs[r][kk] = (b0 + r < ncols && k < k1) ? x[(size_t)(b0 + r) * ldx + k] : C{T(0), T(0)};
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < kProjTileK; ++kk) {
            const C av = as[tp][kk], xv = xs[tb][kk];
            acc.x += av.x * xv.x - av.y * xv.y;
            acc.y += av.x * xv.y + av.y * xv.x;
        }
        __syncthreads();
    }
    if (p0 + tp < P && b0 + tb < ncols) partial[((size_t)ks * ncols + b0 + tb) * P + p0 + tp] = acc;
}

template <typename T>
__global__ void k_nl_dij(const typename C2T<T>::type* __restrict__ partial, int nks,
                         typename C2T<T>::type* __restrict__ q, const ProjMeta* __restrict__ meta,
                         const T* __restrict__ dm, int P, int ncols, T inv_vol) {
    using C = typename C2T<T>::type;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P * ncols) return;
    const int b = i / P, p = i - b * P;
    const ProjMeta mt = meta[p];
    C acc = C{T(0), T(0)};
    for (int r = 0; r < mt.mt; ++r) {
        const T d = dm[mt.dm_off + mt.m * mt.mt + r];
        if (d == T(0)) continue;
        C s = C{T(0), T(0)};
        for (int ks = 0; ks < nks; ++ks) {
            const C v = partial[((size_t)ks * ncols + b) * P + mt.base + r];
            s.x += v.x;
            s.y += v.y;
        }
        acc.x += d * s.x;
        acc.y += d * s.y;
    }
    q[(size_t)b * P + p] = C{acc.x * inv_vol, acc.y * inv_vol};
}

template <typename T>
__global__ __launch_bounds__(256) void k_nl_back(const typename C2T<T>::type* __restrict__ a,
                                                 const typename C2T<T>::type* __restrict__ q,
                                                 typename C2T<T>::type* __restrict__ y, size_t ldy,
                                                 int P, int npw, int ncols) {
    using C = typename C2T<T>::type;
    __shared__ C qs[16][33];
    const int tid = threadIdx.x;
    const int g = blockIdx.x * 64 + (tid & 63);
    const int bq = tid >> 6;
    const int b0 = blockIdx.y * 16;
    C acc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = C{T(0), T(0)};
    for (int pt = 0; pt < P; pt += 32) {
        for (int e = tid; e < 16 * 32; e += 256) {
            const int bb = e >> 5, pp = e & 31;
            qs[bb][pp] = (b0 + bb < ncols && pt + pp < P) ? q[(size_t)(b0 + bb) * P + pt + pp]
                                                           : C{T(0), T(0)};
        }
        __syncthreads();
        if (g < npw) {
            const int pe = min(32, P - pt);
            for (int pp = 0; pp < pe; ++pp) {
                const C av = a[(size_t)(pt + pp) * npw + g];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const C qv = qs[bq + 4 * i][pp];
                    // conj(a) * q
                    acc[i].x += av.x * qv.x + av.y * qv.y;
                    acc[i].y += av.x * qv.y - av.y * qv.x;
                }
            }
        }
        __syncthreads();
    }
    if (g < npw) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int b = b0 + bq + 4 * i;
            if (b < ncols) {
                C& o = y[(size_t)b * ldy + g];
                o = C{o.x + acc[i].x, o.y + acc[i].y};
            }
        }
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
                            &plane_ycols, &kinetic64,
                            &kinetic32, &vloc64, &vloc32, &a64, &a32, &meta, &dm64, &dm32,
                            &xstage, &ystage, &x32, &y32, &tbuf, &partial, &qbuf};
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
    std::vector<int> plane_of_x(h.nx, -1), xval;
    for (int xx = 0; xx < h.nx; ++xx)
        if (count[xx] > 0) {
            plane_of_x[xx] = (int)xval.size();
            xval.push_back(xx);
        }
    h.nplanes = (int)xval.size();
    std::vector<int> start(h.nplanes + 1, 0);
    for (int p = 0; p < h.nplanes; ++p) start[p + 1] = start[p] + count[xval[p]];
    std::vector<int> fill(start.begin(), start.end() - 1);
    std::vector<int> pw(h.npw), pos(h.npw);
    const int pitch = h.ny + 1;  // LDS row pitch of the [z][y] plane tile
    std::vector<std::vector<char>> yused(h.nplanes, std::vector<char>(h.ny, 0));
    for (size_t i = 0; i < h.npw; ++i) {
        const int p = plane_of_x[xs[i]];
        const int j = fill[p]++;
        pw[j] = (int)i;
        pos[j] = py[ys[i]] + pitch * pz[zs[i]];
        yused[p][ys[i]] = 1;
    }
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

void ham_set_local(Ham& h, const double* local_r) {
    std::vector<double> v(h.ngrid);
    const double inv_n = 1.0 / (double)h.ngrid;
    for (size_t i = 0; i < h.ngrid; ++i) v[i] = local_r[i] * inv_n;
    std::vector<float> v32(v.begin(), v.end());
    upload(h.vloc64, v);
    upload(h.vloc32, v32);
    h.has_local = true;
}

void ham_release(Ham& h) {
    DevBuf* bufs[] = {&h.plane_start, &h.plane_pw, &h.plane_pos, &h.plane_xpos, &h.plane_ystart,
                      &h.plane_ycols, &h.kinetic64,
                      &h.kinetic32, &h.vloc64, &h.vloc32, &h.a64, &h.a32, &h.meta, &h.dm64,
                      &h.dm32, &h.xstage, &h.ystage, &h.x32, &h.y32, &h.tbuf, &h.partial, &h.qbuf};
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

constexpr size_t kMaxColsPerPass = 256;

template <typename T>
void ham_apply_dev(Ham& h, const void* xv, void* yv, size_t ncols, bool local, bool nonlocal) {
    using C = typename C2T<T>::type;
    const C* x = reinterpret_cast<const C*>(xv);
    C* y = reinterpret_cast<C*>(yv);
    if (ncols == 0) return;
    DZ_REQUIRE(!local || h.has_local, DZHIP_ERR_STATE, "local potential not set");
    hipStream_t s = h.stream();
    const C* tw = Sel<T>::tw(h);
    const size_t ld = h.npw;
    const bool do_nl = nonlocal && h.nproj > 0;

    for (size_t c0 = 0; c0 < ncols; c0 += kMaxColsPerPass) {
        const size_t nc = std::min(kMaxColsPerPass, ncols - c0);
        const C* xc = x + c0 * ld;
        C* yc = y + c0 * ld;
        hipEvent_t* evs = next_event_set(h);
        mark(h, evs, 0);
        // Nonlocal projection first (reads x only).
        int nks = 1;
        if (do_nl) {
            const int gp = (h.nproj + kProjTileP - 1) / kProjTileP;
            const int gb = (int)((nc + kProjTileB - 1) / kProjTileB);
            const int ktiles = (int)((h.npw + kProjTileK - 1) / kProjTileK);
            nks = std::max(1, std::min(ktiles, (2048 + gp * gb - 1) / (gp * gb)));
            const int kchunk = (int)round_up((h.npw + nks - 1) / nks, kProjTileK);
            nks = (int)((h.npw + kchunk - 1) / kchunk);
            h.partial.ensure((size_t)nks * nc * h.nproj * sizeof(C));
            h.qbuf.ensure(nc * h.nproj * sizeof(C));
            hipLaunchKernelGGL(k_nl_proj<T>, dim3(gp, gb, nks), dim3(256), 0, s, Sel<T>::a(h), xc, ld,
                               h.partial.as<C>(), h.nproj, (int)h.npw, (int)nc, kchunk);
            DZ_HIP(hipGetLastError());
            const int tot = (int)(nc * h.nproj);
            hipLaunchKernelGGL(k_nl_dij<T>, dim3((tot + 255) / 256), dim3(256), 0, s,
                               h.partial.as<C>(), nks, h.qbuf.as<C>(), h.meta.as<ProjMeta>(),
                               Sel<T>::dm(h), h.nproj, (int)nc, (T)h.inv_volume);
            DZ_HIP(hipGetLastError());
        }
        mark(h, evs, 1);
        if (local) {
            const size_t plane = (size_t)h.ny * h.nz;
            h.tbuf.ensure(nc * h.nplanes * plane * sizeof(C));
            C* tb = h.tbuf.as<C>();
            const size_t lds_p = plane_lds_bytes<C>(h.ny, h.nz);
            const size_t lds_x = xline_lds_bytes<C>(h.nx, h.ny);
            constexpr int nt = TileThreads<T>::value;
            set_max_dynamic_lds((const void*)k_plane_inv<T>, lds_p);
            set_max_dynamic_lds((const void*)k_plane_fwd<T>, lds_p);
            set_max_dynamic_lds((const void*)k_xline<T>, lds_x);
            hipLaunchKernelGGL(k_plane_inv<T>, dim3(h.nplanes, (unsigned)nc), dim3(nt), lds_p, s,
                               xc, ld, tb, h.plane_start.as<int>(), h.plane_pw.as<int>(),
                               h.plane_pos.as<int>(), h.plane_ystart.as<int>(), h.plane_ycols.as<short>(),
                               h.ny, h.nz, h.nplanes, h.tab.ax[1], h.tab.ax[2], tw);
            DZ_HIP(hipGetLastError());
            mark(h, evs, 2);
            hipLaunchKernelGGL(k_xline<T>, dim3(h.nz, (unsigned)nc), dim3(nt), lds_x, s, tb,
                               h.plane_xpos.as<int>(), h.nplanes, Sel<T>::vloc(h), h.nx, h.ny, h.nz,
                               h.tab.ax[0], tw);
            DZ_HIP(hipGetLastError());
            mark(h, evs, 3);
            hipLaunchKernelGGL(k_plane_fwd<T>, dim3(h.nplanes, (unsigned)nc), dim3(nt), lds_p, s,
                               tb, xc, ld, yc, ld, Sel<T>::kin(h), h.plane_start.as<int>(),
                               h.plane_pw.as<int>(), h.plane_pos.as<int>(), h.plane_ystart.as<int>(),
                               h.plane_ycols.as<short>(), h.ny, h.nz, h.nplanes, h.tab.ax[1],
                               h.tab.ax[2], tw, 1);
            DZ_HIP(hipGetLastError());
            mark(h, evs, 4);
        } else {
            const size_t n = nc * ld;
            hipLaunchKernelGGL(k_fill_zero<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, yc, n);
            DZ_HIP(hipGetLastError());
            mark(h, evs, 2);
            mark(h, evs, 3);
            mark(h, evs, 4);
        }
        if (do_nl) {
            hipLaunchKernelGGL(k_nl_back<T>, dim3((unsigned)((h.npw + 63) / 64), (unsigned)((nc + 15) / 16)),
                               dim3(256), 0, s, Sel<T>::a(h), h.qbuf.as<C>(), yc, ld, h.nproj,
                               (int)h.npw, (int)nc);
            DZ_HIP(hipGetLastError());
        }
        mark(h, evs, 5);
        // Bound the number of outstanding event sets.
        if (h.profile && h.ev_used >= 4096) h.collect_profile();
    }
    if (h.profile) h.calls++;
}

template void ham_apply_dev<double>(Ham&, const void*, void*, size_t, bool, bool);
template void ham_apply_dev<float>(Ham&, const void*, void*, size_t, bool, bool);

}  // namespace dzhip
