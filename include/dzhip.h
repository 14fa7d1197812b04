/*
 * dzhip.h — C ABI of the MI355X (gfx950) plane-wave H·psi / eigensolver library.
 *
 * This is the `hip` backend that sits beside dft-zig's fftw/zig/vdsp/metal FFT
 * backends and behind the H·psi callback its SCF, band and DFPT solvers call.
 * Every entry point takes plain pointers and sizes; complex arrays are
 * interleaved (re, im) doubles, blocks of vectors are column-major
 * [n x ncols] with column b at [b*n, (b+1)*n)
 * (reference: src/features/linalg/lobpcg/common.zig:54-62; math/complex.zig:3-6).
 *
 * Status codes: 0 ok, <0 error; dzhip_last_error() returns a message for the
 * calling thread.  Reference errors map as error.InvalidMatrixSize ->
 * DZHIP_ERR_SIZE (src/features/scf/apply.zig:584, 1592).
 *
 * Units follow the reference: Rydberg energies, bohr lengths
 * (docs/adr/0001-rydberg-unit-system.md).
 */
#ifndef DZHIP_H
#define DZHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DZHIP_OK 0
#define DZHIP_ERR_SIZE (-1)   /* length mismatch (reference error.InvalidMatrixSize) */
#define DZHIP_ERR_HIP (-2)    /* HIP runtime / library failure */
#define DZHIP_ERR_NOMEM (-3)  /* device or host allocation failed */
#define DZHIP_ERR_STATE (-4)  /* NULL / uninitialised handle, local potential not set */
#define DZHIP_ERR_ARG (-5)    /* unsupported argument (grid not 2/3/5-smooth, l > 4, ...) */

#define DZHIP_FP64 0 /* FFTs, V_loc multiply and projector GEMMs in double */
#define DZHIP_FP32 1 /* same in float, FP32 accumulation; inputs/outputs stay double at the host ABI */

typedef struct dzhip_ctx dzhip_ctx;
typedef struct dzhip_fft_plan dzhip_fft_plan;
typedef struct dzhip_ham dzhip_ham;

/* Library version (major*10000 + minor*100 + patch). */
int dzhip_version(void);

/* Message describing the last failed call on this thread ("" if none). */
const char* dzhip_last_error(void);

/* ------------------------------------------------------------------ context
 * One context per (process, device).  Replaces metal_create_context
 * (src/lib/gpu/metal_bridge.h:16-30): returns NULL on failure, with the
 * reason in dzhip_last_error(). */
dzhip_ctx* dzhip_ctx_create(int device);
void dzhip_ctx_destroy(dzhip_ctx* ctx);

/* ------------------------------------------------------------------ 3D FFT
 * Replaces Fft3dPlan / FftwPlan3d (src/features/fft/fft.zig:44-263,
 * src/lib/fft/fftw_fft.zig:63-172) and the Metal bridge's plan calls
 * (src/lib/gpu/metal_bridge.h:32-72).  In place on host memory, complex
 * interleaved, index x + nx*(y + ny*z); forward = sum x e^{-2 pi i jk/n}
 * unnormalised, inverse = (1/N) sum x e^{+2 pi i jk/n}
 * (src/lib/fft/mixed_radix.zig:47-61, fftw_fft.zig:141-166).
 * Each of nx, ny, nz must be 2/3/5-smooth and <= 256.  count must equal
 * nx*ny*nz (the Metal bridge silently ignored a mismatch; here it is
 * DZHIP_ERR_SIZE). */
dzhip_fft_plan* dzhip_fft_plan_create(dzhip_ctx* ctx, uint64_t nx, uint64_t ny, uint64_t nz,
                                      int precision);
void dzhip_fft_plan_destroy(dzhip_fft_plan* plan);
int dzhip_fft_forward(dzhip_fft_plan* plan, double* data, uint64_t count);
int dzhip_fft_inverse(dzhip_fft_plan* plan, double* data, uint64_t count);

/* ------------------------------------------------------------------ dense Hermitian eigensolver
 * The lowest nev eigenpairs of a dense Hermitian m x m matrix: the subspace eigenproblem of the
 * eigensolvers.  Replaces hermitian_eigen_decomp_small and linalg.hermitian_eigen_decomp
 * (src/features/linalg/lobpcg/common.zig:440-537), which call LAPACK zheev there; the device
 * eigensolvers call it internally for their Rayleigh-Ritz step.  a: complex interleaved,
 * column-major, leading dimension lda >= m; only its lower triangle is read, as zheev with
 * uplo = 'L' (not modified).  values[0:nev]
 * ascending; vectors: nev columns of length m, leading dimension ldz >= m, orthonormal.
 * 1 <= nev <= m <= 1024, else DZHIP_ERR_ARG.  Blocking (host memory). */
int dzhip_heev(dzhip_ctx* ctx, const double* a, uint64_t lda, uint64_t m, uint64_t nev, double* values,
               double* vectors, uint64_t ldz);
/* The same on device memory (a, values, vectors device pointers), enqueued on `stream` (NULL = the
 * context's stream); returns once enqueued.  dzhip_heev_check(ctx, stream) synchronises the stream
 * and reports DZHIP_ERR_HIP if the last call's tridiagonalisation did not complete. */
int dzhip_heev_dev(dzhip_ctx* ctx, const void* a, uint64_t lda, uint64_t m, uint64_t nev, void* values,
                   void* vectors, uint64_t ldz, void* stream);
int dzhip_heev_check(dzhip_ctx* ctx, void* stream);
/* Diagnostic, no counterpart in the reference: how many times in this process a solver's
 * Rayleigh-Ritz tridiagonalisation, launched as an ordinary dispatch beside other k-points' work,
 * gave up waiting for its workgroups and was repeated as a cooperative launch. */
uint64_t dzhip_heev_retries(void);
/* dzhip_heev_dev without any allocation, for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid it); it has no counterpart in
 * the reference.  m <= 256: every buffer lives in the library's 6 MiB static device array, one per
 * device, whose address dzhip_ctx_create takes.  Enqueue only; a tridiagonalisation that did not
 * complete leaves NaN eigenvalues.  Calls that overlap on one device share the array and race. */
int dzhip_heev_dev_noalloc(dzhip_ctx* ctx, const void* a, uint64_t lda, uint64_t m, uint64_t nev, void* values,
                           void* vectors, uint64_t ldz, void* stream);

/* ------------------------------------------------------------------ H·psi
 * Per-k Hamiltonian context: the data ApplyContext + KpointApplyCache carry
 * (src/features/scf/apply.zig:45-574, nonlocal_context.zig:9-246).
 * All arrays are copied at create time; the caller keeps ownership. */
typedef struct dzhip_species_desc {
    uint64_t n_beta;       /* projectors (beta functions) of this species */
    const int32_t* l_list; /* [n_beta] angular momentum, 0..4 */
    const double* dij;     /* [n_beta*n_beta] row-major D_ij (Ry), UPF PP_DIJ; only same-l pairs act
                              (apply.zig:940-966) */
    const double* dij_m;   /* optional [m_total*m_total] dense m-resolved D (apply.zig:926-939); NULL = use dij */
    uint64_t m_total;      /* sum_beta (2 l + 1) */
    const double* phi;     /* [m_total*n_pw] phi[(off_beta+m)*n_pw + G] = 4 pi R_beta(|k+G|) Y_lm(k+G)
                              (nonlocal_context.zig:524-549), m = -l..l */
} dzhip_species_desc;

typedef struct dzhip_ham_desc {
    uint64_t nx, ny, nz;       /* FFT grid (pw_grid.zig:9-24) */
    uint64_t n_pw;             /* plane waves at this k */
    const int32_t* fft_index;  /* [n_pw] FFT-order position of each PW (PwGridMap.fft_indices,
                                  pw_grid_map.zig:58-79) */
    const double* kinetic;     /* [n_pw] |k+G|^2 in Ry (basis.zig:57) */
    const double* gcart;       /* [n_pw*3] Cartesian G (not k+G) for the e^{iG.tau} phase (apply.zig:893) */
    const double* local_r;     /* [nx*ny*nz] V_loc(r) in Ry, FFT order; may be NULL (set later) */
    double inv_volume;         /* 1/Omega (bohr^-3) */
    uint64_t n_species;
    const dzhip_species_desc* species; /* [n_species]; may be NULL when n_species == 0 (no nonlocal) */
    uint64_t n_atoms;
    const double* atom_pos;      /* [n_atoms*3] Cartesian, bohr */
    const int32_t* atom_species; /* [n_atoms] index into species */
} dzhip_ham_desc;

/* Build the per-k device context (index tables, projector matrix, FFT plans).
 * Replaces ApplyContext.init_* (apply.zig:336-532).  NULL on failure. */
dzhip_ham* dzhip_ham_create(dzhip_ctx* ctx, const dzhip_ham_desc* desc);
void dzhip_ham_destroy(dzhip_ham* ham);

/* Replace V_loc(r) (once per SCF iteration; build_local_potential_real,
 * src/features/scf/potential.zig:320-334).  count must be nx*ny*nz. */
int dzhip_ham_set_local(dzhip_ham* ham, const double* local_r, uint64_t count);

/* y = H x for ncols vectors, host memory, blocking.  Replaces
 * apply_hamiltonian (ncols == 1, apply.zig:643-658) and
 * apply_hamiltonian_batched (apply.zig:1584-1618).  y is overwritten, never
 * accumulated.  n must equal n_pw (else DZHIP_ERR_SIZE). */
int dzhip_ham_apply(dzhip_ham* ham, const double* x, double* y, uint64_t n, uint64_t ncols,
                    int precision);

/* Device-resident variant used by the eigensolver and by callers that keep
 * psi in HBM.  x_dev/y_dev hold double2 (DZHIP_FP64) or float2 (DZHIP_FP32)
 * elements, column-major with leading dimension n.  Enqueued on the
 * context's stream (see dzhip_ham_set_stream); returns without waiting. */
int dzhip_ham_apply_dev(dzhip_ham* ham, const void* x_dev, void* y_dev, uint64_t n, uint64_t ncols,
                        int precision);

/* ------------------------------------------------------------------ batched multi-k apply
 * All bands of several k-points in one launch sequence (and, with use_graph, one captured
 * hipGraph launch): replaces the per-k-point loop of apply_hamiltonian_batched
 * (src/features/scf/apply.zig:1584-1618) over the k-point work queue
 * (src/features/scf/kpoint/workers.zig:133-180).  The members are existing contexts of one grid
 * and projector set (one per k-point, uneven n_pw allowed) with a register x-line kernel that runs
 * on all of them; they must outlive the set and hold the same local potential (the set applies
 * member 0's).  Each k-point's result equals its own dzhip_ham_apply_dev bit for bit when the
 * members use the same x-line variant and plane kernels (dzhip_ham_set_tuning,
 * dzhip_ham_set_plane_kernel); the set runs member 0's plane kernels (untuned: the z-register
 * kernels where they run). */
typedef struct dzhip_hamset dzhip_hamset;
dzhip_hamset* dzhip_hamset_create(dzhip_ham* const* hams, uint64_t count);
void dzhip_hamset_destroy(dzhip_hamset* set);
/* Offsets of the k-points in the concatenated layout: off[k] = sum_{k' < k} n_pw(k'), k <= count
 * (off[count] = the total). */
int dzhip_hamset_offsets(dzhip_hamset* set, uint64_t* off, uint64_t n_off);
/* Device apply.  x_dev, y_dev: band-major over the concatenated spheres, element (k, b, G) at
 * b * total + off[k] + G (double2 or float2 by precision); ncols <= 256.  Enqueued on the set's
 * stream (dzhip_hamset_set_stream; default the context of member 0).  use_graph != 0 captures the
 * launch sequence once per (x, y, ncols, precision, stream) and replays it. */
int dzhip_hamset_apply_dev(dzhip_hamset* set, const void* x_dev, void* y_dev, uint64_t ncols, int precision,
                           int use_graph);
/* Host blocks per k-point, column-major [n_pw(k) x ncols] as the reference's apply_batch; blocking. */
int dzhip_hamset_apply(dzhip_hamset* set, const double* const* x, double* const* y, uint64_t ncols, int precision);
int dzhip_hamset_set_stream(dzhip_hamset* set, void* stream);
/* out[0] = k-points, [1] = total n_pw, [2] = total x-planes, [3] = graphs captured so far. */
int dzhip_hamset_info(dzhip_hamset* set, uint64_t* out, uint64_t n_out);

/* Nonlocal part only, y = V_nl x (apply_nonlocal_potential, apply.zig:793-808 and
 * apply_nonlocal_batched, apply.zig:1449-1482); host memory, blocking. */
int dzhip_ham_apply_nonlocal(dzhip_ham* ham, const double* x, double* y, uint64_t n, uint64_t ncols,
                             int precision);

/* y = S x.  Norm-conserving species carry no q_ij, so S = I and this is a
 * copy (apply_overlap, apply.zig:1041-1111; the callback's apply_s = null). */
int dzhip_ham_apply_overlap(dzhip_ham* ham, const double* x, double* y, uint64_t n, uint64_t ncols);

/* Launch subsequent work of this context on `stream` (a hipStream_t, or NULL
 * for the context's own stream).  Lets callers interleave with their own
 * streams and time with their own events. */
int dzhip_ham_set_stream(dzhip_ham* ham, void* stream);

/* Wait for all work enqueued by this context. */
int dzhip_ham_sync(dzhip_ham* ham);

/* Phase timers (the names of ScfProfile, src/features/scf/logging.zig:114-131):
 * when enabled, each apply records HIP events around its phases on the
 * context's stream without waiting; dzhip_ham_phase_times() waits for them
 * and returns the accumulated milliseconds in out[DZHIP_PHASE_COUNT] and the
 * number of applies in *calls.  Enabling or disabling resets the totals. */
#define DZHIP_PHASE_NL_PROJECT 0 /* <beta|psi> projection + D_ij */
#define DZHIP_PHASE_NL_BACK 1    /* back-projection, written into y */
#define DZHIP_PHASE_PLANE_INV 2  /* scatter + inverse FFT along z, y (per x-plane) */
#define DZHIP_PHASE_XLINE 3      /* inverse FFT along x, V_loc multiply, forward FFT along x */
#define DZHIP_PHASE_PLANE_FWD 4  /* forward FFT along y, z + gather + kinetic, added to y */
#define DZHIP_PHASE_COUNT 5
int dzhip_ham_profile(dzhip_ham* ham, int enable);
int dzhip_ham_phase_times(dzhip_ham* ham, double* out_ms, uint64_t n_out, uint64_t* calls);

/* ------------------------------------------------------------------ eigensolver
 * Lowest nbands eigenpairs of H on this k-point, S = I.  Replaces
 * hermitian_eigen_decomp_iterative (src/features/linalg/iterative.zig:67-76) /
 * lobpcg_serial.solve (src/features/linalg/lobpcg/serial.zig:55-307) with the same outer
 * semantics: block = nbands + 4, subspace up to 4 nbands + 8, converged when
 * max_b ||H x_b - lambda_b x_b|| < tol (Ry), kinetic-shift preconditioner, LCG start with seed
 * 0x6f3d3cbe.  Orthonormalisation is block CGS2 + Cholesky QR, Rayleigh-Ritz is a device
 * zheevd, restarts keep [Ritz vectors, new block].  All subspace algebra is FP64; with
 * precision = DZHIP_FP32 the H apply runs in FP32; when tol < mixed_switch_tol the basis
 * restarts with an FP64 apply once the largest residual is below mixed_switch_tol (looser
 * tolerances are met in FP32). */
#define DZHIP_PRECOND_KINETIC 0 /* r / (|k+G|^2 - lambda), |den| >= 1e-8 (common.zig:214-226) */
#define DZHIP_PRECOND_TPA 1     /* Teter-Payne-Allan K(x) r, x = |k+G|^2 / <x_b|T|x_b> */
#define DZHIP_PRECOND_MLP 2     /* learned: B(G, b) exp(MLP(features)) r with B the model's base
                                   (kinetic shift or TPA), model set with
                                   dzhip_ham_set_precond_model; falls back to the base within a
                                   solve when its progress trails the base's mean progress at the
                                   same iteration on this context (stats.precond_fallback) */

typedef struct dzhip_eig_opts {
    uint64_t max_iter;       /* Options.max_iter (reference default 40) */
    double tol;              /* Options.tol: max residual 2-norm, Ry */
    uint64_t max_subspace;   /* 0 = min(n, 4 nbands + 8) */
    uint64_t block_size;     /* 0 = min(max_subspace, nbands + 4) */
    int init_diagonal;       /* Options.init_diagonal: unit vectors at the smallest diag entries */
    int precision;           /* DZHIP_FP64 or DZHIP_FP32 (mixed: FP32 apply, FP64 at the end) */
    double mixed_switch_tol; /* residual below which the FP32 apply switches to FP64; 0 = 1e-5 */
    int preconditioner;      /* DZHIP_PRECOND_* */
    uint64_t seed;           /* LCG seed; 0 = 0x6f3d3cbe */
    int reuse_vectors;       /* with init_vecs NULL: start from the Ritz vectors of this context's
                                previous solve, kept on the device (the warm-start cache of
                                kpoint/solve.zig:542-557, iterative_reuse_vectors) */
    uint64_t collect_snapshots; /* record (Ritz vectors, residuals) of the first this-many FP64
                                   iterations for dzhip_ham_training_samples; 0 = off */
    uint64_t nconv;          /* converged when the largest residual of the lowest nconv bands is
                                below tol (the rest are a buffer that still gets new directions);
                                0 = all nbands, the reference's test (serial.zig:238-283) */
} dzhip_eig_opts;

typedef struct dzhip_eig_stats {
    uint64_t iterations;            /* Rayleigh-Ritz steps */
    uint64_t apply_calls;           /* H applied to this many vectors in total */
    uint64_t fp32_applies, fp64_applies;
    uint64_t restarts;
    uint64_t precision_switch_iter; /* iteration at which the FP32 apply switched to FP64 (0: never) */
    uint64_t converged;             /* 1 if max residual < tol at return */
    double max_residual;
    double initial_residual;        /* largest residual of the start block */
    uint64_t precond_fallback;      /* DZHIP_PRECOND_MLP: iteration at which the solve fell back
                                       to the model's base preconditioner (0: never; 1 also when
                                       the model sits out this solve: after two solves in a row
                                       that fell back it sits out the next 8 on this context) */
    /* GPU time (ms, HIP events on the context's stream) per section of the solve:
     * Rayleigh-Ritz (T copy + zheevd), Ritz vectors + residuals + norms, preconditioner,
     * orthonormalisation (CGS2 + Cholesky QR), H apply of new vectors, T extension/restart.
     * Host waits between sections (small downloads) fall into the section that follows. */
    double t_rr_ms, t_ritz_ms, t_precond_ms, t_ortho_ms, t_apply_ms, t_extend_ms;
    uint64_t ortho_fallbacks;       /* new blocks whose Cholesky QR failed and that were
                                       orthonormalised through the Gram eigen-decomposition */
} dzhip_eig_stats;

/* diag: [n] preconditioner diagonal (|k+G|^2, kpoint/solve.zig:362-367).  init_vecs: NULL or
 * [n * init_cols] start vectors for the first columns (warm start, solve.zig:542-557).
 * values_out: [nbands] ascending; vectors_out: NULL or [n * nbands] column-major.  stats may
 * be NULL.  When not converged after max_iter the best estimate is returned (serial.zig:298-306). */
int dzhip_eig_solve(dzhip_ham* ham, const double* diag, uint64_t n, uint64_t nbands,
                    const dzhip_eig_opts* opts, const double* init_vecs, uint64_t init_cols,
                    double* values_out, double* vectors_out, dzhip_eig_stats* stats);

/* Band-by-band conjugate-gradient eigensolver, the `solver = "cg"` path: replaces cg.solve
 * (src/features/linalg/cg/cg.zig:79-116): per band, preconditioned (kinetic shift) CG with a
 * three-term Rayleigh-Ritz on {psi, d, psi_prev} (two-term closed form when psi_prev adds
 * nothing), band converged when |r| < tol or |delta lambda| < tol; then a subspace rotation of
 * the whole block (cg.zig:450-482).  Uses opts->max_iter (0 = 20), tol (0 = 1e-6), seed and
 * reuse_vectors; FP64 apply.  stats->iterations counts CG steps over all bands. */
int dzhip_cg_solve(dzhip_ham* ham, const double* diag, uint64_t n, uint64_t nbands,
                   const dzhip_eig_opts* opts, const double* init_vecs, uint64_t init_cols,
                   double* values_out, double* vectors_out, dzhip_eig_stats* stats);

/* ------------------------------------------------------------------ density
 * rho(r) (+)= (1/Omega) sum_b w_b |psi_b(r)|^2 with psi_b(r) = sum_G c_bG e^{iG.r}, FP64, rho in
 * FFT order [nx*ny*nz].  Replaces the per-band loop of accumulate_band_contributions /
 * accumulate_band_density_fft (src/features/scf/kpoint/density.zig:90-143, 444-474); the
 * caller supplies the weights w_b = w_k * occ_b * 2 (density.zig:98-107), zero weights skip
 * the band.  accumulate = 0 overwrites rho.  Host pointers, blocking. */
int dzhip_density_accumulate(dzhip_ham* ham, const double* psi, uint64_t n, uint64_t nbands,
                             const double* weights, double* rho, uint64_t count, int accumulate);

/* The same accumulation for the vectors the last dzhip_eig_solve on this context returned
 * (still on the device; the first nbands of them), so an SCF iteration moves no wave
 * functions over PCIe.  DZHIP_ERR_STATE if there is no solution with nbands columns. */
int dzhip_density_accumulate_solution(dzhip_ham* ham, uint64_t nbands, const double* weights,
                                      double* rho, uint64_t count, int accumulate);

/* ------------------------------------------------------------------ learned preconditioner
 * No counterpart in the reference (its Davidson uses the kinetic-shift diagonal,
 * src/features/linalg/lobpcg/common.zig:214-226).  A 3 -> 32 -> 32 -> 1 SiLU network (FP32)
 * rescales a base diagonal preconditioner B per coefficient:
 *   s(G, b) = B(G, b) exp(z),
 *   z = MLP(|k+G|^2/2 / e_s, lambda_b/2 / e_s, log10(|r_b|) / 4)     (energies in Ha)
 *   B = 1 / (|k+G|^2 - lambda_b), |den| >= 1e-8            base DZHIP_PRECOND_KINETIC
 *   B = K_tpa(|k+G|^2 / <x_b|T|x_b>)                         base DZHIP_PRECOND_TPA
 * Parameters: DZHIP_MLP_PARAMS floats, W1[32][3], b1[32], W2[32][32], b2[32], w3[32], b3.
 * Samples: structure of arrays [6][n] floats: x1, x2, x3, B, target, weight;
 * loss = sum_i w_i (s_i - t_i)^2 / n over a minibatch of n samples. */
#define DZHIP_MLP_PARAMS 1217

int dzhip_ham_set_precond_model(dzhip_ham* ham, const float* params, uint64_t n_params, double energy_scale,
                                int base);

/* Residual/correction pairs of the snapshots the last solve recorded (collect_snapshots > 0,
 * FP64, converged): correction e = X X^H x_i - x_i against the converged Ritz vectors X,
 * target t = Re(e conj r) / (alpha |r|^2) and weight alpha^2 |r|^2 / |e|^2 with alpha the
 * band's least-squares scale of base preconditioner `base` (DZHIP_PRECOND_KINETIC or
 * DZHIP_PRECOND_TPA); every stride-th coefficient.  Writes min(capacity, available) samples to
 * out ([6][capacity]) and their count.  Weight-0 samples mark unusable bands. */
int dzhip_ham_training_samples(dzhip_ham* ham, double energy_scale, int base, uint64_t stride, float* out,
                               uint64_t capacity, uint64_t* count);

/* Full-batch loss and gradient (FP64 out) of the network on n samples ([6][n]). */
int dzhip_mlp_loss_grad(dzhip_ctx* ctx, const float* samples, uint64_t n, const float* params,
                        double* grad_out, double* loss_out);

/* Adam training on the device: `epochs` passes over n samples in minibatches of `batch`, each
 * pass in a different pseudo-random order (seeded); params in/out; epoch_loss (may be NULL)
 * receives the full-batch loss after each epoch. */
int dzhip_mlp_train(dzhip_ctx* ctx, const float* samples, uint64_t n, float* params, uint64_t epochs,
                    uint64_t batch, double lr, uint64_t seed, double* epoch_loss);

/* ------------------------------------------------------------------ SCF grid
 * Device-resident density, potential and mixing state of one SCF run, so an iteration moves
 * no grid or wave-function data over PCIe.  Replaces, on the device: build_potential_grid /
 * build_local_potential_real (src/features/scf/potential.zig:144-203, 320-334) with Hartree
 * (potential.zig:34-44) and LDA-PZ (src/features/xc/xc.zig:195-210, 344-384); the potential-mixing
 * branch (src/features/scf/scf.zig:147-188) with linear (mixing.zig:10-18) and Pulay
 * (mixing.zig:183-234) mixing; the density sum over k and its RMS change (util.zig:17-24).
 * recip: reciprocal lattice rows b1, b2, b3 (bohr^-1, 9 doubles); volume in bohr^3.
 * Pulay overlaps are real-space sums / N (the reference's G-space norm by Parseval), scaled by
 * the largest diagonal; a singular system restarts the history with a linear step. */
typedef struct dzhip_scf_grid dzhip_scf_grid;

#define DZHIP_GRID_RHO 0   /* density of the last iteration (rho_in of the next) */
#define DZHIP_GRID_VIN 1   /* V_Hxc input of the next iteration */
#define DZHIP_GRID_VOUT 2  /* V_Hxc[rho] of the last iteration */
#define DZHIP_GRID_LOCAL 3 /* V_ion + V_in */
#define DZHIP_GRID_VION 4

dzhip_scf_grid* dzhip_scf_grid_create(dzhip_ctx* ctx, uint64_t nx, uint64_t ny, uint64_t nz,
                                      const double* recip, double volume, uint64_t pulay_history);
void dzhip_scf_grid_destroy(dzhip_scf_grid* grid);
/* V_ion(r) (Ry, FFT order, host). */
int dzhip_scf_grid_set_ion(dzhip_scf_grid* grid, const double* v_ion, uint64_t count);
/* rho = nelec / volume, V_in = V_Hxc[rho], Pulay history cleared, density sum cleared. */
int dzhip_scf_grid_init(dzhip_scf_grid* grid, double nelec);
/* The same from a given start density (host, FFT order), e.g. the superposition of atomic
 * charges of core_density.zig:128-178. */
int dzhip_scf_grid_init_rho(dzhip_scf_grid* grid, const double* rho, uint64_t count);
/* local_r of ham = V_ion + V_in (dzhip_ham_set_local without the host copy). */
int dzhip_ham_set_local_grid(dzhip_ham* ham, dzhip_scf_grid* grid);
/* Add (1/Omega) sum_b w_b |psi_b(r)|^2 of ham's last eigensolver solution to the density sum. */
int dzhip_scf_grid_add_solution_density(dzhip_scf_grid* grid, dzhip_ham* ham, const double* weights,
                                        uint64_t nbands);
/* From the density sum: V_out = V_H + V_xc; out[0] = E_H, out[1] = E_xc, out[2] = int V_in rho,
 * out[3] = RMS(rho - rho_previous) (energies in Ry); then the sum becomes rho and is cleared. */
int dzhip_scf_grid_hxc(dzhip_scf_grid* grid, double* out, uint64_t n_out);
/* Model dielectric preconditioner of the Pulay residual, P(G) = (dielng^2 G^2 + 1/diemac) /
 * (dielng^2 G^2 + 1) (src/features/scf/mixing.zig:309-338); diemac = 1 (default) disables it. */
int dzhip_scf_grid_set_dielectric(dzhip_scf_grid* grid, double diemac, double dielng);
/* Exchange-correlation functional of the grid's V_xc and E_xc (the reference's [scf] xc):
 * LDA-PZ point-wise (src/features/xc/xc.zig:195-210) or PBE with the gradient and the
 * divergence term taken spectrally on the grid (xc.zig:213-342,
 * src/features/scf/xc_fields.zig:87-140, 306-375).  Default DZHIP_XC_LDA_PZ; takes effect at the
 * next dzhip_scf_grid_init / _init_rho / _hxc. */
#define DZHIP_XC_LDA_PZ 0
#define DZHIP_XC_PBE 1
int dzhip_scf_grid_set_xc(dzhip_scf_grid* grid, int functional);
/* V_in <- linear (pulay = 0) or Pulay (pulay != 0) mix of V_in and V_out. */
int dzhip_scf_grid_mix(dzhip_scf_grid* grid, double beta, int pulay);
int dzhip_scf_grid_read(dzhip_scf_grid* grid, int which, double* out, uint64_t count);

/* Launch-shape tuning (no reference counterpart): the LDS x-line kernel processes the y-lines
 * of one (z, band) in xline_split slabs; -1 / -2 / -3 select the register-resident x-line
 * kernel with one x-line per thread (nx = 2^a 3^b 5^c, 8 <= nx <= 100 in FP32, <= 50 in FP64)
 * in workgroups of 256 / 128 / 64 threads, -4 / -5 / -6 the lane-pair register kernel (two
 * threads per x-line; even nx <= 100 whose half is 2^a 3^b 5^c, both precisions, when the
 * G-sphere spans at most half the line), -7 / -8 / -9 the lane-triple register kernel (three
 * threads per x-line; nx in 12..96 divisible by 3 whose third is 2^a 3^b 5^c, e.g. 75 = 3 x 25,
 * both precisions, any sphere width), -10 / -11 / -12 the residue-sequential register kernel (one
 * thread per x-line doing the three residues x mod 3 in turn; the same line lengths, both
 * precisions, when the G-sphere spans at most two thirds of the line); a precision where the
 * chosen kernel cannot run keeps one slab.  -1..-3 need the one-lane kernel in FP32, -4..-6 the
 * lane-pair kernel, -7..-9 the lane-triple kernel, -10..-12 the residue-sequential kernel.
 * -13 selects the convolution x-line kernel (the x-part of H on a line's present x-frequencies as
 * a Toeplitz matrix product over all bands on the matrix cores; the sphere's x-planes one cyclic
 * run of at most 64 planes, both precisions).  0 (default) times every divisor of ny with at
 * least 8 lines per slab, and
 * the register and convolution kernels' variants where they can run, on the first apply of each precision with at
 * least 16 columns and keeps the fastest; the result is shared by every context of the same
 * device, grid and sphere shape in the process. */
int dzhip_ham_set_tuning(dzhip_ham* ham, int xline_split);

/* Plane-kernel choice (no reference counterpart): the sphere <-> [z][y] plane transforms run either
 * as full-tile LDS kernels (kind 1: the whole ny x nz plane of one band in LDS, z- and y-transforms
 * as LDS stages; any grid) or as z-register kernels (kind 2: only the plane's active z-frequency rows
 * in LDS for the y-transforms, the z-transform of every y-line in registers, three lanes per line;
 * nz in 12..96 divisible by 3 with nz / 3 = 2^a 3^b 5^c and ny <= 126).  0 (default) times both on
 * the first apply of each precision with at least 16 columns and keeps the faster; the result is
 * shared like the x-line tuning.  Results agree to rounding either way. */
int dzhip_ham_set_plane_kernel(dzhip_ham* ham, int kind);

/* Overlap (no reference counterpart; default off): with on != 0 the nonlocal projection / D /
 * back-projection kernels run on a second stream of the context, concurrently with the local
 * plane_inv and x-line kernels, and join before plane_fwd.  Results are identical either way;
 * profiling (dzhip_ham_profile) runs the phases in order.  Measured on the bench block, the
 * co-running kernels slow each other down (back-projection 105 -> 290 us) to no net gain. */
int dzhip_ham_set_overlap(dzhip_ham* ham, int on);

/* Sizes the library uses for this context (for roofline accounting):
 * out[0] = n_pw, [1] = grid points, [2] = x-planes holding the G-sphere,
 * [3] = projectors P, [4] = device bytes held, [5] = LDS x-line slabs chosen for FP32,
 * [6] = for FP64 (0 = not tuned yet, or the register kernel), [7] = workgroup size when FP32
 * runs the register x-line kernel (0 = the LDS kernel or not tuned yet), [8] = the same for
 * FP64, [9] / [10] = threads per x-line of that register kernel, 1, 2 or 3 (FP32 / FP64; 0 = the
 * LDS or convolution kernel, or not tuned yet), [11] / [12] = its kind (FP32 / FP64): 1 one-lane,
 * 2 lane-pair, 3 lane-triple, 4 residue-sequential, 5 convolution (0 = the LDS kernel or not
 * tuned yet), [13] / [14] = the plane kernels (FP32 / FP64): 1 full-tile, 2 z-register (0 = not
 * tuned yet). */
int dzhip_ham_info(dzhip_ham* ham, uint64_t* out, uint64_t n_out);

#ifdef __cplusplus
}
#endif

#endif /* DZHIP_H */
