// Hermitian eigensolver for the Rayleigh-Ritz subspace problems of the device eigensolvers:
// the lowest nev eigenpairs of a dense Hermitian m x m matrix (FP64 complex), m <= kHeevMaxM,
// all on the device in a handful of launches.
//
//   1. Householder tridiagonalisation (zhetd2's reflectors, A = Q T Q^H with T real tridiagonal)
//      in one cooperative launch: the trailing matrix is distributed over the workgroups' LDS by
//      columns and each reflector costs one grid barrier.
//   2. The lowest nev eigenvalues of T by 64-way multisection with Sturm counts (one wavefront
//      per eigenvalue).
//   3. Eigenvectors of T by inverse iteration (tridiagonal LU with partial pivoting in LDS, one
//      wavefront per eigenvalue); the members of a cluster of eigenvalues closer than a threshold
//      are then orthonormalised together (two classical Gram-Schmidt passes), and two
//      Newton-Schulz steps Z <- Z (3 I - Z^T Z) / 2 orthonormalise all nev vectors.
//   4. Back-transformation Y = Q Z (one wavefront per eigenvector).
#pragma once

#include "dzhip_internal.h"

namespace dzhip {

constexpr int kHeevMaxM = 1024;

struct HeevWork {
    DevBuf vh, tau, d, e, pbuf, nbuf, sync, z, zt, gram;
    int err_offset = 0;  // int index of the error flag in sync
    void release();
};

// Largest matrix of the allocation-free variant (its buffers live in a 6 MiB device array).
constexpr int kHeevNoAllocMaxM = 256;

// Lowest nev eigenpairs of the Hermitian matrix A (column-major, leading dimension lda; only the
// lower triangle is read, the imaginary part of the diagonal is ignored; A is not modified).  Eigenvalues (ascending) -> w_dev[0:nev] (device); eigenvectors ->
// Y[:, 0:nev] (complex, leading dimension ldy).  Enqueued on stream s.  Throws on a size outside
// 1 <= nev <= m <= kHeevMaxM.  After the call the host may read w_dev once s is synchronised.
void heev_lowest(HeevWork& w, const double2* A, int lda, int m, int nev, double* w_dev, double2* Y, int ldy,
                 hipStream_t s);

// The tridiagonal stage only (for tests): d[0:m], e[0:m-1] of T on the device.
void heev_tridiag_only(HeevWork& w, const double2* A, int lda, int m, hipStream_t s);

// Error flag of the last tridiagonalisation (a grid barrier that did not complete); synchronises s.
int heev_error_flag(HeevWork& w, hipStream_t s);

// Allocation-free variant for m <= kHeevNoAllocMaxM: every buffer in the library's static device
// scratch array of `device` (heev_scratch_init must have run for it, with it current).  A failed
// tridiagonalisation leaves NaN eigenvalues.  Calls that overlap on one device race.
void heev_scratch_init(int device);
void heev_lowest_noalloc(int device, const double2* A, int lda, int m, int nev, double* w_dev, double2* Y, int ldy,
                         hipStream_t s);

}  // namespace dzhip
