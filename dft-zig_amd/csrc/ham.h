// Per-k Hamiltonian context held on the device (host-side struct).
#pragma once

#include "dzhip_internal.h"

namespace dzhip {

template <typename T>
void fft3d_device(typename C2T<T>::type* data, const AxisTables& tab, bool inverse, hipStream_t s);
void convert_to_float(const double2* in, float2* out, size_t n, hipStream_t s);
void convert_to_double(const float2* in, double2* out, size_t n, hipStream_t s);

struct FftPlan {
    Context* ctx = nullptr;
    int nx = 0, ny = 0, nz = 0;
    int precision = DZHIP_FP64;
    AxisTables tab;
    DevBuf buf64, buf32;
    std::mutex mu;
};

// Projector metadata: row p of the complex projector matrix belongs to the atom
// block starting at row `base`, is its m-row `m`, the block has `mt` rows and
// the species' dense m-resolved D starts at `dm_off` in the D table.
struct ProjMeta {
    int base, m, mt, dm_off;
};

struct Ham {
    Context* ctx = nullptr;
    hipStream_t user_stream = nullptr;
    hipStream_t stream() const { return user_stream ? user_stream : ctx->stream; }

    int nx = 0, ny = 0, nz = 0;
    size_t npw = 0, ngrid = 0;
    double inv_volume = 1.0;
    AxisTables tab;

    // G-sphere split into x-planes (one plane per distinct FFT x of a plane wave):
    //   plane_start[p]..plane_start[p+1]: entries of plane p
    //   plane_pw[j]  : plane-wave index of entry j
    //   plane_pos[j] : LDS position py[y] + (ny+1)*pz[z] of that plane wave in the yz tile
    //   plane_xpos[p]: LDS row px[x] of plane p in the x-line tile
    //   plane_ystart[p]..plane_ystart[p+1]: plane_ycols entries = LDS columns py[y] of the
    //                  y values present in plane p (the z-transforms that are not all zero)
    //   plane_xabsent: LDS rows px[x] of the x values with no plane wave (zeroed in the x-line tile)
    int nplanes = 0, nxabsent = 0;
    DevBuf plane_start, plane_pw, plane_pos, plane_xpos, plane_ystart, plane_ycols, plane_xabsent;
    // Active z-rows of the z-register plane kernels:
    //   plane_zstart[p]..plane_zstart[p+1]: plane_zrows entries = the z-frequencies (natural
    //                  index) present in plane p, ascending; row b of the plane is plane_zrows[b]
    //   plane_posr[j]: LDS position py[y] + (ny|1)*b of entry j in the plane's active rows
    //   zrow_max:      most active rows of any plane (sizes the LDS)
    //   plane_pw0[p]:  first plane-wave index of plane p when its entries are one run of the
    //                  basis (plane_pw[j] = plane_pw0[p] + j - plane_start[p]), else -1
    int zrow_max = 0;
    DevBuf plane_zstart, plane_zrows, plane_posr, plane_pw0;

    DevBuf kinetic64, kinetic32;
    DevBuf vloc64, vloc32;  // V_loc(r)/N in Ry
    // Register x-line kernel: plane p is x = plane_xs + p (mod nx) when the sphere's x-planes form
    // one cyclic run (plane_xs = -1 otherwise), and V_loc/N in x-major order [x][z][y] (FP32, FP64).
    int plane_xs = -1;
    DevBuf vloc32_t, vloc64_t;
    // Convolution x-line kernel: the forward x-transform of every line of V_loc/N, [z][y][kx]
    // complex (FP64, FP32).
    DevBuf chat64, chat32;
    bool has_local = false;

    // Nonlocal: complex projector matrix A[p][G] = phi(G) e^{iG.tau_a} (P x n_pw), D tables.
    int nproj = 0;
    int nblocks = 0;  // atom blocks (first projector row of each atom in block_rows)
    DevBuf a64, a32, meta, block_rows, dm64, dm32;

    // Work buffers.
    DevBuf xstage, ystage, x32, y32, tbuf, partial, qbuf;
    // Tuned launch shapes: x-line variant per precision ([0] FP64, [1] FP32): 0 not yet tuned,
    // d >= 1 the LDS kernel with d y-slabs, -1..-12 a register kernel (xline_reg_threads,
    // xline_reg_kind), kXlineConv the convolution kernel; tune_fixed != 0 forces a variant (sweeps).
    int xline_split[2] = {0, 0};
    int tune_fixed = 0;
    // Plane kernels per precision ([0] FP64, [1] FP32): 0 not yet tuned, kPlaneTile the full-tile
    // LDS kernels, kPlaneZreg the z-register kernels; plane_fixed != 0 forces one.
    int plane_kind[2] = {0, 0};
    int plane_fixed = 0;

    // Profiling: one set of phase-boundary events per column pass, recorded on the stream
    // without waiting; read (and recycled) by collect_profile().
    bool profile = false;
    std::vector<std::vector<hipEvent_t>> ev_sets;  // each: DZHIP_PHASE_COUNT + 1 events
    size_t ev_used = 0;
    double phase_ms[DZHIP_PHASE_COUNT] = {};
    uint64_t calls = 0;
    void collect_profile();

    // Eigensolver workspace (created on first solve); it keeps the last solve's Ritz vectors.
    void* eig_work = nullptr;
    // Learned preconditioner model (kMlpParams floats), its feature energy scale (Ha) and the
    // base preconditioner it rescales (DZHIP_PRECOND_KINETIC or DZHIP_PRECOND_TPA).
    DevBuf mlp_params;
    double mlp_energy_scale = 0.0;
    int mlp_base = DZHIP_PRECOND_TPA;
    bool has_mlp = false;
    // Mean drop of the log residual since the first iteration, per iteration index, of the solves
    // with the kinetic-shift [0] and TPA [1] preconditioners on this context (with sample counts).
    std::vector<double> precond_curve[2];
    std::vector<int> precond_curve_n[2];
    // Learned-model solves that fell back in a row, and solves the model still sits out.
    int mlp_fail_streak = 0, mlp_rest = 0;
    // Density staging (grow-only).
    DevBuf dens_rho, dens_w;
    // Overlapped apply: the nonlocal chain on a second stream (created on first use), joined
    // before plane_fwd; only when enabled by dzhip_ham_set_overlap and not profiling.  Off by
    // default: at the bench size the co-running kernels slow each other down to no net gain.
    bool overlap = false;
    hipStream_t nl_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // Stream switch (dzhip_ham_set_stream): recorded on the old stream, waited on by the new one,
    // so work enqueued before the switch (which shares the workspaces) completes first.
    hipEvent_t ev_switch = nullptr;

    size_t device_bytes() const;
    std::mutex mu;
};

// ----------------------------------------------------------------- learned preconditioner
constexpr int kMlpHidden = 32;
constexpr int kMlpParams = 3 * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden + kMlpHidden + 1;

// Q[:, j] = s(G, sel_j) R[:, sel_j] with the network's scaling (device pointers).
void precondition_mlp(const double2* r, const int* sel, int nsel, const double* lam, const double* ekin,
                      const double* rnorm, const double* diag, const float* prm, double energy_scale, int base,
                      double2* q, size_t ld, int n, hipStream_t s);
// Training samples (SoA [6][cap] floats: x1 x2 x3 base-factor target weight) of `cols`
// residual/correction column pairs for base preconditioner pbase, every `stride`-th coefficient,
// written from index base; returns how many.  bandsum: 3 * cols doubles of device scratch.
size_t mlp_write_samples(const double2* e, const double2* r, const double* lam, const double* ekin,
                         const double* rnorm, const double* diag, size_t ld, int n, int cols, int stride,
                         double energy_scale, int pbase, double* bandsum, float* out, size_t cap, size_t base,
                         hipStream_t s);

// Device state of one training run: samples (SoA [6][cap]), parameters, Adam moments, gradient.
struct MlpTrainer {
    size_t cap;
    DevBuf samples, prm, m, v, g, part;
    explicit MlpTrainer(size_t cap);
    ~MlpTrainer();
    // g[0..kMlpParams) = gradient, g[kMlpParams] = loss of the minibatch of `count` samples
    // idx(j) = (mul (start + j) + add) mod n.
    double loss_grad(size_t n, unsigned long long mul, unsigned long long add, size_t start, size_t count,
                     hipStream_t s);
    void adam(int step, double lr, hipStream_t s);
};

// rho[r] (+)= inv_volume * sum_b w_b |psi_b(r)|^2 (FP64), psi and rho on the device.
void ham_density_dev(Ham& h, const double2* psi, size_t nbands, const double* wts_dev, double* rho_dev,
                     bool accumulate);
void eig_solve(Ham& h, const double* diag, uint64_t nbands, const dzhip_eig_opts& o, const double* init,
               uint64_t init_cols, double* values, double* vectors, dzhip_eig_stats* stats);
void eig_release(Ham& h);
// Band-by-band CG with three-term Rayleigh-Ritz and a final subspace rotation (FP64).
void cg_solve(Ham& h, const double* diag, uint64_t nbands, const dzhip_eig_opts& o, const double* init,
              uint64_t init_cols, double* values, double* vectors, dzhip_eig_stats* stats);
// Ritz vectors [npw x cols] (FP64, device) the last eig_solve on h left, or nullptr.
const double2* eig_solution(const Ham& h, uint64_t* cols);
// Training samples from the snapshots the last solve recorded (opts.collect_snapshots) against
// its converged Ritz vectors, relative to base preconditioner `base`; host output SoA [6][cap];
// returns the count written.
size_t eig_training_samples(Ham& h, double energy_scale, int base, int stride, float* out_host, size_t cap);

void ham_build(Ham& h, const dzhip_ham_desc& d);
void ham_ensure_side_stream(Ham& h);
// True when the register-resident x-line kernel is built for line length nx in the given
// precision, and when it can run on this context (its x-planes form a cyclic run).
bool xline_reg_supported(int nx, bool fp64);
bool xline_reg_usable(const Ham& h, bool fp64);
// ... and when the lane-pair register kernel (two lanes per x-line, half-length transforms) is
// built for nx and can run here: the sphere spans at most half the line.
bool xline_pair_supported(int nx);
bool xline_pair_usable(const Ham& h);
// ... and the lane-triple kernel (three lanes per x-line, third-length transforms, any sphere
// width): nx divisible by 3 with nx / 3 = 2^a 3^b 5^c.
bool xline_tri_supported(int nx);
bool xline_tri_usable(const Ham& h);
// Register x-line variants: -1 / -2 / -3 = one lane per x-line in workgroups of 256 / 128 / 64
// threads, -4 / -5 / -6 = the lane-pair kernel, -7 / -8 / -9 = the lane-triple kernel, in the same
// workgroup sizes.
constexpr int kXlineReg = -1, kXlineRegLast = -12;
// The convolution x-line kernel on the matrix cores (one variant; kind 4 in xline_reg_kind's
// numbering).
constexpr int kXlineConv = -13;
// Register x-line kernel kinds (codes -1..-3, -4..-6, -7..-9, -10..-12 in that order).
constexpr int kXlineOneLane = 0, kXlinePair = 1, kXlineTriple = 2, kXlineSeq3 = 3;
constexpr int xline_reg_threads(int code) { return 256 >> ((-code - 1) % 3); }
constexpr int xline_reg_kind(int code) { return (-code - 1) / 3; }
// Threads per x-line: one (one-lane and residue-sequential kernels), two or three.
constexpr int xline_reg_lanes(int code) {
    return code == kXlineConv ? 0 : xline_reg_kind(code) == kXlineSeq3 ? 1 : 1 + xline_reg_kind(code);
}
constexpr bool xline_reg_pair(int code) { return xline_reg_kind(code) == kXlinePair; }
// ... and the convolution kernel: the sphere's x-planes form one cyclic run of at most 64 planes.
bool xline_conv_usable(const Ham& h);
// Whether variant `code` (d >= 1 LDS slabs, a register variant or the convolution kernel) can run
// on h in a precision.
bool xline_code_usable(const Ham& h, int code, bool fp64);
// Plane kernels: the full-tile LDS kernels (any grid), or the z-register kernels (nz in 12..96
// divisible by 3 with nz / 3 = 2^a 3^b 5^c, ny <= 126).
constexpr int kPlaneTile = 1, kPlaneZreg = 2;
bool plane_zreg_usable(const Ham& h);
void ham_set_local(Ham& h, const double* local_r);
// local_r already on the device (read on h's stream).
void ham_set_local_dev(Ham& h, const double* local_dev);

// ----------------------------------------------------------------- batched multi-k apply
// Several k-point contexts of one grid and projector set applied in one launch sequence
// (optionally one graph launch).  Device layout of x and y: band-major over the concatenated
// spheres, element (k, b, G) at b * npw_total + npw_off[k] + G.  The members' local potential is
// taken from member 0 (all members hold the same one in an SCF iteration).
struct HamSet {
    Context* ctx = nullptr;
    std::vector<Ham*> ks;
    hipStream_t user_stream = nullptr;
    hipStream_t stream() const { return user_stream ? user_stream : ctx->stream; }
    int nx = 0, ny = 0, nz = 0, nproj = 0;
    size_t npw_total = 0;
    int np_total = 0, max_planes = 0;
    std::vector<size_t> npw_off;
    std::vector<int> np_off;
    DevBuf plane_start, plane_pw, plane_pos, plane_ystart, plane_ycols, kpl, kinetic64, kinetic32, a64, a32;
    // Active z-row tables of the z-register plane kernels over the concatenated planes.
    int zrow_max = 0;
    DevBuf plane_pw0, plane_posr, plane_zstart, plane_zrows;
    // Nonlocal segment tables for seg_cols columns in precision seg_prec (element size).
    size_t seg_cols = 0;
    int seg_prec = 0, nks_total = 0, ntiles = 0;
    DevBuf chunks, kseg, tiles;
    DevBuf tbuf, partial, qbuf;
    // Captured launch sequence and what it was captured for.
    struct GraphKey {
        const void* x;
        void* y;
        size_t nc;
        int prec, code, pkind;
        hipStream_t s;
        void *tbuf, *partial, *qbuf;
        bool operator==(const GraphKey& o) const {
            return x == o.x && y == o.y && nc == o.nc && prec == o.prec && code == o.code && pkind == o.pkind &&
                   s == o.s && tbuf == o.tbuf && partial == o.partial && qbuf == o.qbuf;
        }
    };
    hipGraphExec_t graph_exec = nullptr;
    GraphKey graph_key{};
    uint64_t graph_builds = 0;
    hipEvent_t ev_switch = nullptr;
    std::mutex mu;
};
void hamset_build(HamSet& hs);
template <typename T>
void hamset_apply_dev(HamSet& hs, const void* x, void* y, size_t ncols, bool use_graph);
void hamset_release(HamSet& hs);

// ----------------------------------------------------------------- SCF grid (potential.hip)
struct ScfGrid;
ScfGrid* scf_grid_create(Context* ctx, int nx, int ny, int nz, const double* recip, double volume, int history);
void scf_grid_destroy(ScfGrid* g);
void scf_grid_init(ScfGrid& g, double nelec, const double* rho_host);  // rho_host NULL: uniform
void scf_grid_local(ScfGrid& g);
void scf_grid_add(ScfGrid& g, const double* rho_k);
void scf_grid_hxc(ScfGrid& g, double out[4]);
void scf_grid_mix(ScfGrid& g, double beta, int pulay);
void scf_grid_set_xc(ScfGrid& g, int functional);  // DZHIP_XC_LDA_PZ or DZHIP_XC_PBE
void ham_release(Ham& h);
// Device apply; x/y hold C2T<T> elements, column-major with leading dimension npw.
template <typename T>
void ham_apply_dev(Ham& h, const void* x, void* y, size_t ncols, bool local, bool nonlocal);

}  // namespace dzhip
