// Hermitian eigensolver for the Rayleigh-Ritz subspace problems of the device eigensolvers:
// the lowest nev eigenpairs of a dense Hermitian m x m matrix (FP64 complex), m <= kHeevMaxM,
// all on the device in a handful of launches.
//
//   1. Householder tridiagonalisation (zhetd2's reflectors, A = Q T Q^H with T real tridiagonal)
//      in one grid-synchronised launch: the trailing matrix is distributed over the workgroups'
//      LDS by columns and each reflector costs one grid barrier.  The enqueue-only entry points
//      launch it cooperatively; the solvers' Rayleigh-Ritz step launches it as an ordinary
//      dispatch admitted by the device's tridiagonalisation gate (TrdAdmission), so that the
//      grids of concurrent solves on different streams run side by side.
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

// Share of the device's tridiagonalisation gate held by one eigensolver call.  A grid barrier
// needs all of its workgroups resident at once.  Workgroups are dealt to the XCDs round robin, so
// a grid of G groups needs at most ceil(G / XCDs) CUs on any XCD (one workgroup per CU at worst).
// The gate admits grids only while their sum stays within the CUs of one XCD; every other kernel
// the library launches runs to completion without waiting on another workgroup, so an admitted
// grid's missing workgroups are placed as those retire.  That can take long: with many streams
// issuing kernels, smaller workgroups keep refilling the CUs a tridiagonalisation workgroup (80-100
// KB of LDS) waits for (one barrier ran out of its spin limit in a concurrent-k SCF of some
// thousands of calls).  An admitted grid's barrier therefore gives up after 2^19 polls (an eighth
// of the cooperative launch's limit) and heev_lowest_shared repeats the call cooperatively.  The holder synchronises its stream
// (the grid has finished) and returns its share when it goes out of scope.  Cooperative launches (the enqueue-only entry
// points) are not counted: one that runs while concurrent solves hold the gate may find the
// device short of CUs, and its barrier's spin limit then reports an error instead of waiting.
struct TrdAdmission {
    int device = -1, units = 0;
    hipStream_t stream = nullptr;
    TrdAdmission() = default;
    TrdAdmission(const TrdAdmission&) = delete;
    TrdAdmission& operator=(const TrdAdmission&) = delete;
    ~TrdAdmission();
};

// Lowest nev eigenpairs of the Hermitian matrix A (column-major, leading dimension lda; only the
// lower triangle is read, the imaginary part of the diagonal is ignored; A is not modified).  Eigenvalues (ascending) -> w_dev[0:nev] (device); eigenvectors ->
// Y[:, 0:nev] (complex, leading dimension ldy).  Enqueued on stream s.  Throws on a size outside
// 1 <= nev <= m <= kHeevMaxM.  After the call the host may read w_dev once s is synchronised.
// With adm == nullptr the tridiagonalisation is a cooperative launch; otherwise it waits for a
// share of the gate, launches as an ordinary dispatch and leaves the share in *adm (which must
// be empty and must outlive the call's use of s).
void heev_lowest(HeevWork& w, const double2* A, int lda, int m, int nev, double* w_dev, double2* Y, int ldy,
                 hipStream_t s, TrdAdmission* adm = nullptr);

// The tridiagonal stage only (for tests): d[0:m], e[0:m-1] of T on the device.
void heev_tridiag_only(HeevWork& w, const double2* A, int lda, int m, hipStream_t s);

// Error flag of the last tridiagonalisation (a grid barrier that did not complete); synchronises s.
int heev_error_flag(HeevWork& w, hipStream_t s);

// heev_lowest through the gate, synchronised: the tridiagonalisation runs as an admitted ordinary
// dispatch whose barrier gives up after 2^19 polls without all its workgroups (other streams'
// kernels can keep the CUs it waits for), and is then repeated as a cooperative launch.  Throws if
// that fails too.  heev_shared_retries counts the repeats in the process.
void heev_lowest_shared(HeevWork& w, const double2* A, int lda, int m, int nev, double* w_dev, double2* Y, int ldy,
                        hipStream_t s);
unsigned long long heev_shared_retries();

// Allocation-free variant for m <= kHeevNoAllocMaxM: every buffer in the library's static device
// scratch array of `device` (heev_scratch_init must have run for it, with it current).  A failed
// tridiagonalisation leaves NaN eigenvalues.  Calls that overlap on one device race.
void heev_scratch_init(int device);
void heev_lowest_noalloc(int device, const double2* A, int lda, int m, int nev, double* w_dev, double2* Y, int ldy,
                         hipStream_t s);

}  // namespace dzhip
