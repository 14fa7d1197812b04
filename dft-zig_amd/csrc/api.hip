// C ABI of libdzhip (declared in include/dzhip.h).
#include <cstring>
#include <exception>
#include <new>
#include <vector>

#include "ham.h"
#include "heev.h"
#include "potential.h"

namespace dzhip {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

template <typename F>
static int guarded(F&& f) {
    try {
        f();
        return DZHIP_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return DZHIP_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return DZHIP_ERR_HIP;
    }
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) DZ_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace dzhip

using namespace dzhip;

struct dzhip_ctx : Context {};
struct dzhip_fft_plan : FftPlan {};
struct dzhip_ham : Ham {};
struct dzhip_hamset : HamSet {};
struct dzhip_scf_grid : ScfGrid {};

namespace dzhip {
// rho (host, ngrid) (+)= density of the nbands device columns psi, through the context's
// staging buffers; blocking.
static void density_host_rho(Ham& h, const double2* psi, uint64_t nbands, const double* weights, double* rho,
                             bool accumulate) {
    hipStream_t s = h.stream();
    const size_t count = h.ngrid;
    h.dens_rho.ensure(count * sizeof(double));
    if (accumulate) DZ_HIP(hipMemcpyAsync(h.dens_rho.ptr, rho, count * sizeof(double), hipMemcpyHostToDevice, s));
    if (nbands > 0) {
        h.dens_w.ensure(nbands * sizeof(double));
        DZ_HIP(hipMemcpyAsync(h.dens_w.ptr, weights, nbands * sizeof(double), hipMemcpyHostToDevice, s));
        ham_density_dev(h, psi, nbands, h.dens_w.as<double>(), h.dens_rho.as<double>(), accumulate);
    } else if (!accumulate) {
        DZ_HIP(hipMemsetAsync(h.dens_rho.ptr, 0, count * sizeof(double), s));
    }
    DZ_HIP(hipMemcpyAsync(rho, h.dens_rho.ptr, count * sizeof(double), hipMemcpyDeviceToHost, s));
    DZ_HIP(hipStreamSynchronize(s));
}
}  // namespace dzhip

extern "C" {

int dzhip_version(void) { return 100; }

const char* dzhip_last_error(void) { return g_last_error.c_str(); }

dzhip_ctx* dzhip_ctx_create(int device) {
    dzhip_ctx* out = nullptr;
    int rc = guarded([&] {
        int n = 0;
        DZ_HIP(hipGetDeviceCount(&n));
        DZ_REQUIRE(device >= 0 && device < n, DZHIP_ERR_ARG,
                   "device " + std::to_string(device) + " not present (" + std::to_string(n) + " visible)");
        DeviceGuard g(device);
        // Set-up of this device for the allocation-free eigensolver (its scratch array's address).
        heev_scratch_init(device);
        auto* c = new dzhip_ctx();
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw Error{DZHIP_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)};
        }
        out = c;
    });
    return rc == DZHIP_OK ? out : nullptr;
}

void dzhip_ctx_destroy(dzhip_ctx* ctx) {
    if (!ctx) return;
    (void)guarded([&] {
        DeviceGuard g(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        if (ctx->heev) {
            ctx->heev->release();
            delete ctx->heev;
            ctx->heev = nullptr;
        }
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    });
    delete ctx;
}

// ----------------------------------------------------------------- dense Hermitian eigensolver
int dzhip_heev(dzhip_ctx* ctx, const double* a, uint64_t lda, uint64_t m, uint64_t nev, double* values,
               double* vectors, uint64_t ldz) {
    return guarded([&] {
        DZ_REQUIRE(ctx, DZHIP_ERR_STATE, "NULL context");
        DZ_REQUIRE(a && values && vectors, DZHIP_ERR_ARG, "NULL array");
        DZ_REQUIRE(m >= 1 && m <= (uint64_t)kHeevMaxM && nev >= 1 && nev <= m && lda >= m && ldz >= m, DZHIP_ERR_ARG,
                   "need 1 <= nev <= m <= 1024 and lda, ldz >= m");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        if (!ctx->heev) ctx->heev = new HeevWork();
        hipStream_t s = ctx->stream;
        double2 *da = nullptr, *dz = nullptr;
        double* dw = nullptr;
        auto cleanup = [&] {
            if (da) (void)hipFree(da);
            if (dz) (void)hipFree(dz);
            if (dw) (void)hipFree(dw);
        };
        try {
            DZ_HIP(hipMalloc(&da, (size_t)lda * m * sizeof(double2)));
            DZ_HIP(hipMalloc(&dz, (size_t)m * nev * sizeof(double2)));
            DZ_HIP(hipMalloc(&dw, (size_t)nev * sizeof(double)));
            DZ_HIP(hipMemcpyAsync(da, a, (size_t)lda * m * sizeof(double2), hipMemcpyHostToDevice, s));
            heev_lowest_shared(*ctx->heev, da, (int)lda, (int)m, (int)nev, dw, dz, (int)m, s);
            DZ_HIP(hipMemcpyAsync(values, dw, nev * sizeof(double), hipMemcpyDeviceToHost, s));
            DZ_HIP(hipMemcpy2DAsync(vectors, ldz * sizeof(double2), dz, m * sizeof(double2), m * sizeof(double2), nev,
                                    hipMemcpyDeviceToHost, s));
            DZ_HIP(hipStreamSynchronize(s));
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

int dzhip_heev_dev(dzhip_ctx* ctx, const void* a, uint64_t lda, uint64_t m, uint64_t nev, void* values,
                   void* vectors, uint64_t ldz, void* stream) {
    return guarded([&] {
        DZ_REQUIRE(ctx, DZHIP_ERR_STATE, "NULL context");
        DZ_REQUIRE(a && values && vectors, DZHIP_ERR_ARG, "NULL array");
        DZ_REQUIRE(m >= 1 && m <= (uint64_t)kHeevMaxM && nev >= 1 && nev <= m && lda >= m && ldz >= m, DZHIP_ERR_ARG,
                   "need 1 <= nev <= m <= 1024 and lda, ldz >= m");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        if (!ctx->heev) ctx->heev = new HeevWork();
        hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : ctx->stream;
        heev_lowest(*ctx->heev, static_cast<const double2*>(a), (int)lda, (int)m, (int)nev, static_cast<double*>(values),
                    static_cast<double2*>(vectors), (int)ldz, s);
    });
}

int dzhip_heev_dev_noalloc(dzhip_ctx* ctx, const void* a, uint64_t lda, uint64_t m, uint64_t nev, void* values,
                           void* vectors, uint64_t ldz, void* stream) {
    return guarded([&] {
        DZ_REQUIRE(ctx, DZHIP_ERR_STATE, "NULL context");
        DZ_REQUIRE(a && values && vectors, DZHIP_ERR_ARG, "NULL array");
        DZ_REQUIRE(m >= 1 && m <= (uint64_t)kHeevNoAllocMaxM && nev >= 1 && nev <= m && lda >= m && ldz >= m,
                   DZHIP_ERR_ARG, "allocation-free eigensolver: need 1 <= nev <= m <= 256 and lda, ldz >= m");
        DeviceGuard g(ctx->device);
        hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : ctx->stream;
        heev_lowest_noalloc(ctx->device, static_cast<const double2*>(a), (int)lda, (int)m, (int)nev,
                            static_cast<double*>(values), static_cast<double2*>(vectors), (int)ldz, s);
    });
}

uint64_t dzhip_heev_retries(void) { return (uint64_t)heev_shared_retries(); }

int dzhip_heev_check(dzhip_ctx* ctx, void* stream) {
    return guarded([&] {
        DZ_REQUIRE(ctx, DZHIP_ERR_STATE, "NULL context");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : ctx->stream;
        if (!ctx->heev) {
            DZ_HIP(hipStreamSynchronize(s));
            return;
        }
        DZ_REQUIRE(heev_error_flag(*ctx->heev, s) == 0, DZHIP_ERR_HIP,
                   "Hermitian eigensolver: tridiagonalisation barrier did not complete");
    });
}

// ----------------------------------------------------------------- FFT
dzhip_fft_plan* dzhip_fft_plan_create(dzhip_ctx* ctx, uint64_t nx, uint64_t ny, uint64_t nz,
                                      int precision) {
    dzhip_fft_plan* out = nullptr;
    int rc = guarded([&] {
        DZ_REQUIRE(ctx, DZHIP_ERR_STATE, "NULL context");
        DZ_REQUIRE(precision == DZHIP_FP64 || precision == DZHIP_FP32, DZHIP_ERR_ARG, "bad precision");
        DZ_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nx <= 256 && ny <= 256 && nz <= 256, DZHIP_ERR_ARG,
                   "FFT dimensions must be in 1..256");
        const size_t esz = precision == DZHIP_FP64 ? sizeof(double2) : sizeof(float2);
        DZ_REQUIRE(fft_xy_lds_bytes((int)nx, (int)ny, esz) <= kMaxLdsBytes &&
                       fft_z_lds_bytes((int)nx, (int)nz, esz) <= kMaxLdsBytes,
                   DZHIP_ERR_ARG, "FFT plane exceeds one LDS tile");
        DeviceGuard g(ctx->device);
        auto* p = new dzhip_fft_plan();
        p->ctx = ctx;
        p->nx = (int)nx;
        p->ny = (int)ny;
        p->nz = (int)nz;
        p->precision = precision;
        try {
            p->tab.build(p->nx, p->ny, p->nz);
        } catch (...) {
            p->tab.release();
            delete p;
            throw;
        }
        out = p;
    });
    return rc == DZHIP_OK ? out : nullptr;
}

void dzhip_fft_plan_destroy(dzhip_fft_plan* plan) {
    if (!plan) return;
    (void)guarded([&] {
        DeviceGuard g(plan->ctx->device);
        plan->tab.release();
        plan->buf64.release();
        plan->buf32.release();
    });
    delete plan;
}

static int fft_exec(dzhip_fft_plan* plan, double* data, uint64_t count, bool inverse) {
    return guarded([&] {
        DZ_REQUIRE(plan, DZHIP_ERR_STATE, "NULL plan");
        const size_t n = (size_t)plan->nx * plan->ny * plan->nz;
        DZ_REQUIRE(count == n, DZHIP_ERR_SIZE, "count != nx*ny*nz");
        DZ_REQUIRE(data, DZHIP_ERR_ARG, "NULL data");
        std::lock_guard<std::mutex> lk(plan->mu);
        DeviceGuard g(plan->ctx->device);
        hipStream_t s = plan->ctx->stream;
        plan->buf64.ensure(n * sizeof(double2));
        DZ_HIP(hipMemcpyAsync(plan->buf64.ptr, data, n * sizeof(double2), hipMemcpyHostToDevice, s));
        if (plan->precision == DZHIP_FP64) {
            fft3d_device<double>(plan->buf64.as<double2>(), plan->tab, inverse, s);
        } else {
            plan->buf32.ensure(n * sizeof(float2));
            convert_to_float(plan->buf64.as<double2>(), plan->buf32.as<float2>(), n, s);
            fft3d_device<float>(plan->buf32.as<float2>(), plan->tab, inverse, s);
            convert_to_double(plan->buf32.as<float2>(), plan->buf64.as<double2>(), n, s);
        }
        DZ_HIP(hipMemcpyAsync(data, plan->buf64.ptr, n * sizeof(double2), hipMemcpyDeviceToHost, s));
        DZ_HIP(hipStreamSynchronize(s));
    });
}

int dzhip_fft_forward(dzhip_fft_plan* plan, double* data, uint64_t count) {
    return fft_exec(plan, data, count, false);
}
int dzhip_fft_inverse(dzhip_fft_plan* plan, double* data, uint64_t count) {
    return fft_exec(plan, data, count, true);
}

// ----------------------------------------------------------------- Hamiltonian
dzhip_ham* dzhip_ham_create(dzhip_ctx* ctx, const dzhip_ham_desc* desc) {
    dzhip_ham* out = nullptr;
    int rc = guarded([&] {
        DZ_REQUIRE(ctx, DZHIP_ERR_STATE, "NULL context");
        DZ_REQUIRE(desc, DZHIP_ERR_ARG, "NULL descriptor");
        DeviceGuard g(ctx->device);
        auto* h = new dzhip_ham();
        h->ctx = ctx;
        try {
            ham_build(*h, *desc);
        } catch (...) {
            ham_release(*h);
            delete h;
            throw;
        }
        out = h;
    });
    return rc == DZHIP_OK ? out : nullptr;
}

void dzhip_ham_destroy(dzhip_ham* ham) {
    if (!ham) return;
    (void)guarded([&] {
        DeviceGuard g(ham->ctx->device);
        (void)hipStreamSynchronize(ham->stream());
        ham_release(*ham);
    });
    delete ham;
}

int dzhip_ham_set_local(dzhip_ham* ham, const double* local_r, uint64_t count) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(local_r, DZHIP_ERR_ARG, "NULL local_r");
        DZ_REQUIRE(count == ham->ngrid, DZHIP_ERR_SIZE, "count != nx*ny*nz");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        DZ_HIP(hipStreamSynchronize(ham->stream()));
        ham_set_local(*ham, local_r);
    });
}

static int ham_apply_host(dzhip_ham* ham, const double* x, double* y, uint64_t n, uint64_t ncols,
                          int precision, bool local) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(n == ham->npw, DZHIP_ERR_SIZE, "n != n_pw (InvalidMatrixSize)");
        DZ_REQUIRE(precision == DZHIP_FP64 || precision == DZHIP_FP32, DZHIP_ERR_ARG, "bad precision");
        if (ncols == 0) return;
        DZ_REQUIRE(x && y, DZHIP_ERR_ARG, "NULL x or y");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        hipStream_t s = ham->stream();
        const size_t cnt = n * ncols;
        ham->xstage.ensure(cnt * sizeof(double2));
        ham->ystage.ensure(cnt * sizeof(double2));
        DZ_HIP(hipMemcpyAsync(ham->xstage.ptr, x, cnt * sizeof(double2), hipMemcpyHostToDevice, s));
        if (precision == DZHIP_FP64) {
            ham_apply_dev<double>(*ham, ham->xstage.ptr, ham->ystage.ptr, ncols, local, true);
        } else {
            ham->x32.ensure(cnt * sizeof(float2));
            ham->y32.ensure(cnt * sizeof(float2));
            convert_to_float(ham->xstage.as<double2>(), ham->x32.as<float2>(), cnt, s);
            ham_apply_dev<float>(*ham, ham->x32.ptr, ham->y32.ptr, ncols, local, true);
            convert_to_double(ham->y32.as<float2>(), ham->ystage.as<double2>(), cnt, s);
        }
        DZ_HIP(hipMemcpyAsync(y, ham->ystage.ptr, cnt * sizeof(double2), hipMemcpyDeviceToHost, s));
        DZ_HIP(hipStreamSynchronize(s));
    });
}

int dzhip_ham_apply(dzhip_ham* ham, const double* x, double* y, uint64_t n, uint64_t ncols,
                    int precision) {
    return ham_apply_host(ham, x, y, n, ncols, precision, true);
}

int dzhip_ham_apply_nonlocal(dzhip_ham* ham, const double* x, double* y, uint64_t n, uint64_t ncols,
                             int precision) {
    return ham_apply_host(ham, x, y, n, ncols, precision, false);
}

int dzhip_ham_apply_dev(dzhip_ham* ham, const void* x_dev, void* y_dev, uint64_t n, uint64_t ncols,
                        int precision) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(n == ham->npw, DZHIP_ERR_SIZE, "n != n_pw (InvalidMatrixSize)");
        if (ncols == 0) return;
        DZ_REQUIRE(x_dev && y_dev, DZHIP_ERR_ARG, "NULL x or y");
        DZ_REQUIRE(x_dev != y_dev, DZHIP_ERR_ARG, "x and y must not alias");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        if (precision == DZHIP_FP64)
            ham_apply_dev<double>(*ham, x_dev, y_dev, ncols, true, true);
        else if (precision == DZHIP_FP32)
            ham_apply_dev<float>(*ham, x_dev, y_dev, ncols, true, true);
        else
            throw Error{DZHIP_ERR_ARG, "bad precision"};
    });
}

// ----------------------------------------------------------------- batched multi-k apply
dzhip_hamset* dzhip_hamset_create(dzhip_ham* const* hams, uint64_t count) {
    dzhip_hamset* out = nullptr;
    int rc = guarded([&] {
        DZ_REQUIRE(hams && count > 0, DZHIP_ERR_ARG, "no k-point contexts");
        for (uint64_t k = 0; k < count; ++k) DZ_REQUIRE(hams[k], DZHIP_ERR_STATE, "NULL ham in the set");
        auto* hs = new dzhip_hamset();
        hs->ctx = hams[0]->ctx;
        for (uint64_t k = 0; k < count; ++k) hs->ks.push_back(hams[k]);
        try {
            DeviceGuard g(hs->ctx->device);
            // The members' tables are read while each member is locked in turn.
            for (uint64_t k = 0; k < count; ++k) {
                for (uint64_t j = 0; j < k; ++j)
                    DZ_REQUIRE(hams[j] != hams[k], DZHIP_ERR_ARG, "a context appears twice in the set");
            }
            hamset_build(*hs);
        } catch (...) {
            hamset_release(*hs);
            delete hs;
            throw;
        }
        out = hs;
    });
    return rc == DZHIP_OK ? out : nullptr;
}

void dzhip_hamset_destroy(dzhip_hamset* set) {
    if (!set) return;
    (void)guarded([&] {
        DeviceGuard g(set->ctx->device);
        (void)hipStreamSynchronize(set->stream());
        if (set->ev_switch) (void)hipEventDestroy(set->ev_switch);
        hamset_release(*set);
    });
    delete set;
}

int dzhip_hamset_offsets(dzhip_hamset* set, uint64_t* off, uint64_t n_off) {
    return guarded([&] {
        DZ_REQUIRE(set, DZHIP_ERR_STATE, "NULL set");
        DZ_REQUIRE(off, DZHIP_ERR_ARG, "NULL output");
        for (uint64_t k = 0; k < n_off && k < set->npw_off.size(); ++k) off[k] = set->npw_off[k];
    });
}

int dzhip_hamset_apply_dev(dzhip_hamset* set, const void* x_dev, void* y_dev, uint64_t ncols, int precision,
                           int use_graph) {
    return guarded([&] {
        DZ_REQUIRE(set, DZHIP_ERR_STATE, "NULL set");
        if (ncols == 0) return;
        DZ_REQUIRE(x_dev && y_dev, DZHIP_ERR_ARG, "NULL x or y");
        DZ_REQUIRE(x_dev != y_dev, DZHIP_ERR_ARG, "x and y must not alias");
        DZ_REQUIRE(precision == DZHIP_FP64 || precision == DZHIP_FP32, DZHIP_ERR_ARG, "bad precision");
        std::lock_guard<std::mutex> lk(set->mu);
        std::lock_guard<std::mutex> l0(set->ks[0]->mu);
        DeviceGuard g(set->ctx->device);
        if (precision == DZHIP_FP64)
            hamset_apply_dev<double>(*set, x_dev, y_dev, ncols, use_graph != 0);
        else
            hamset_apply_dev<float>(*set, x_dev, y_dev, ncols, use_graph != 0);
    });
}

int dzhip_hamset_apply(dzhip_hamset* set, const double* const* x, double* const* y, uint64_t ncols, int precision) {
    return guarded([&] {
        DZ_REQUIRE(set, DZHIP_ERR_STATE, "NULL set");
        if (ncols == 0) return;
        DZ_REQUIRE(x && y, DZHIP_ERR_ARG, "NULL x or y");
        DZ_REQUIRE(precision == DZHIP_FP64 || precision == DZHIP_FP32, DZHIP_ERR_ARG, "bad precision");
        DZ_REQUIRE(ncols <= 256, DZHIP_ERR_ARG, "batched apply: at most 256 columns");
        const size_t nk = set->ks.size();
        for (size_t k = 0; k < nk; ++k) DZ_REQUIRE(x[k] && y[k], DZHIP_ERR_ARG, "NULL x or y block");
        std::lock_guard<std::mutex> lk(set->mu);
        std::lock_guard<std::mutex> l0(set->ks[0]->mu);
        DeviceGuard g(set->ctx->device);
        hipStream_t s = set->stream();
        const size_t tot = set->npw_total;
        const size_t esz = precision == DZHIP_FP64 ? sizeof(double2) : sizeof(float2);
        DevBuf dx, dy, d64;
        dx.ensure(tot * ncols * esz);
        dy.ensure(tot * ncols * esz);
        if (precision == DZHIP_FP32) d64.ensure(tot * ncols * sizeof(double2));
        double2* stage = precision == DZHIP_FP64 ? dx.as<double2>() : d64.as<double2>();
        for (size_t k = 0; k < nk; ++k) {
            const size_t n = set->ks[k]->npw;
            DZ_HIP(hipMemcpy2DAsync(stage + set->npw_off[k], tot * sizeof(double2), x[k], n * sizeof(double2),
                                    n * sizeof(double2), ncols, hipMemcpyHostToDevice, s));
        }
        if (precision == DZHIP_FP64) {
            hamset_apply_dev<double>(*set, dx.ptr, dy.ptr, ncols, false);
        } else {
            convert_to_float(d64.as<double2>(), dx.as<float2>(), tot * ncols, s);
            hamset_apply_dev<float>(*set, dx.ptr, dy.ptr, ncols, false);
            convert_to_double(dy.as<float2>(), d64.as<double2>(), tot * ncols, s);
        }
        const double2* res = precision == DZHIP_FP64 ? dy.as<double2>() : d64.as<double2>();
        for (size_t k = 0; k < nk; ++k) {
            const size_t n = set->ks[k]->npw;
            DZ_HIP(hipMemcpy2DAsync(y[k], n * sizeof(double2), res + set->npw_off[k], tot * sizeof(double2),
                                    n * sizeof(double2), ncols, hipMemcpyDeviceToHost, s));
        }
        DZ_HIP(hipStreamSynchronize(s));
    });
}

int dzhip_hamset_set_stream(dzhip_hamset* set, void* stream) {
    return guarded([&] {
        DZ_REQUIRE(set, DZHIP_ERR_STATE, "NULL set");
        std::lock_guard<std::mutex> lk(set->mu);
        DeviceGuard g(set->ctx->device);
        const hipStream_t prev = set->stream();
        set->user_stream = reinterpret_cast<hipStream_t>(stream);
        const hipStream_t next = set->stream();
        if (next == prev) return;
        // The set's workspaces are shared by every apply: the new stream waits for the old one.
        if (!set->ev_switch) DZ_HIP(hipEventCreateWithFlags(&set->ev_switch, hipEventDisableTiming));
        DZ_HIP(hipEventRecord(set->ev_switch, prev));
        DZ_HIP(hipStreamWaitEvent(next, set->ev_switch, 0));
    });
}

int dzhip_hamset_info(dzhip_hamset* set, uint64_t* out, uint64_t n_out) {
    return guarded([&] {
        DZ_REQUIRE(set, DZHIP_ERR_STATE, "NULL set");
        const uint64_t v[4] = {set->ks.size(), set->npw_total, (uint64_t)set->np_total, set->graph_builds};
        for (uint64_t i = 0; i < n_out && i < 4; ++i) out[i] = v[i];
    });
}

int dzhip_ham_apply_overlap(dzhip_ham* ham, const double* x, double* y, uint64_t n, uint64_t ncols) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(n == ham->npw, DZHIP_ERR_SIZE, "n != n_pw (InvalidMatrixSize)");
        if (ncols == 0) return;
        DZ_REQUIRE(x && y, DZHIP_ERR_ARG, "NULL x or y");
        if (x != y) std::memmove(y, x, n * ncols * 2 * sizeof(double));
    });
}

int dzhip_ham_set_stream(dzhip_ham* ham, void* stream) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        const hipStream_t prev = ham->stream();
        ham->user_stream = reinterpret_cast<hipStream_t>(stream);
        const hipStream_t next = ham->stream();
        if (next == prev) return;
        // Every apply shares this context's workspaces: the new stream waits for what the old
        // one still has queued, so the two streams never run this context's kernels at once.
        if (!ham->ev_switch) DZ_HIP(hipEventCreateWithFlags(&ham->ev_switch, hipEventDisableTiming));
        DZ_HIP(hipEventRecord(ham->ev_switch, prev));
        DZ_HIP(hipStreamWaitEvent(next, ham->ev_switch, 0));
    });
}

int dzhip_ham_sync(dzhip_ham* ham) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DeviceGuard g(ham->ctx->device);
        DZ_HIP(hipStreamSynchronize(ham->stream()));
    });
}

int dzhip_ham_profile(dzhip_ham* ham, int enable) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        ham->collect_profile();
        ham->profile = enable != 0;
        for (double& v : ham->phase_ms) v = 0.0;
        ham->calls = 0;
    });
}

int dzhip_ham_phase_times(dzhip_ham* ham, double* out_ms, uint64_t n_out, uint64_t* calls) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        ham->collect_profile();
        for (uint64_t i = 0; i < n_out && i < DZHIP_PHASE_COUNT; ++i) out_ms[i] = ham->phase_ms[i];
        if (calls) *calls = ham->calls;
    });
}

int dzhip_eig_solve(dzhip_ham* ham, const double* diag, uint64_t n, uint64_t nbands,
                    const dzhip_eig_opts* opts, const double* init_vecs, uint64_t init_cols,
                    double* values_out, double* vectors_out, dzhip_eig_stats* stats) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(n == ham->npw, DZHIP_ERR_SIZE, "n != n_pw (InvalidMatrixSize)");
        DZ_REQUIRE(diag && opts && values_out, DZHIP_ERR_ARG, "NULL diag, opts or values_out");
        DZ_REQUIRE(ham->has_local, DZHIP_ERR_STATE, "local potential not set");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        eig_solve(*ham, diag, nbands, *opts, init_vecs, init_cols, values_out, vectors_out, stats);
    });
}

int dzhip_cg_solve(dzhip_ham* ham, const double* diag, uint64_t n, uint64_t nbands, const dzhip_eig_opts* opts,
                   const double* init_vecs, uint64_t init_cols, double* values_out, double* vectors_out,
                   dzhip_eig_stats* stats) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(n == ham->npw, DZHIP_ERR_SIZE, "n != n_pw (InvalidMatrixSize)");
        DZ_REQUIRE(diag && opts && values_out, DZHIP_ERR_ARG, "NULL diag, opts or values_out");
        DZ_REQUIRE(ham->has_local, DZHIP_ERR_STATE, "local potential not set");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        cg_solve(*ham, diag, nbands, *opts, init_vecs, init_cols, values_out, vectors_out, stats);
    });
}

int dzhip_density_accumulate(dzhip_ham* ham, const double* psi, uint64_t n, uint64_t nbands,
                             const double* weights, double* rho, uint64_t count, int accumulate) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(n == ham->npw, DZHIP_ERR_SIZE, "n != n_pw (InvalidMatrixSize)");
        DZ_REQUIRE(count == ham->ngrid, DZHIP_ERR_SIZE, "count != nx*ny*nz");
        DZ_REQUIRE(rho && (nbands == 0 || (psi && weights)), DZHIP_ERR_ARG, "NULL psi, weights or rho");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        hipStream_t s = ham->stream();
        DevBuf dpsi;
        if (nbands > 0) {
            dpsi.ensure(n * nbands * sizeof(double2));
            DZ_HIP(hipMemcpyAsync(dpsi.ptr, psi, n * nbands * sizeof(double2), hipMemcpyHostToDevice, s));
        }
        density_host_rho(*ham, dpsi.as<double2>(), nbands, weights, rho, accumulate != 0);
        dpsi.release();
    });
}

int dzhip_density_accumulate_solution(dzhip_ham* ham, uint64_t nbands, const double* weights, double* rho,
                                      uint64_t count, int accumulate) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(count == ham->ngrid, DZHIP_ERR_SIZE, "count != nx*ny*nz");
        DZ_REQUIRE(rho && (nbands == 0 || weights), DZHIP_ERR_ARG, "NULL weights or rho");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        uint64_t cols = 0;
        const double2* x = eig_solution(*ham, &cols);
        DZ_REQUIRE(nbands == 0 || (x && nbands <= cols), DZHIP_ERR_STATE,
                   "no eigensolver solution with that many bands on this context");
        density_host_rho(*ham, x, nbands, weights, rho, accumulate != 0);
    });
}

int dzhip_ham_set_precond_model(dzhip_ham* ham, const float* params, uint64_t n_params, double energy_scale,
                                int base) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(params && n_params == (uint64_t)kMlpParams, DZHIP_ERR_SIZE, "need DZHIP_MLP_PARAMS parameters");
        DZ_REQUIRE(energy_scale > 0.0, DZHIP_ERR_ARG, "energy_scale must be > 0");
        DZ_REQUIRE(base == DZHIP_PRECOND_KINETIC || base == DZHIP_PRECOND_TPA, DZHIP_ERR_ARG,
                   "base must be DZHIP_PRECOND_KINETIC or DZHIP_PRECOND_TPA");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        ham->mlp_params.ensure(kMlpParams * sizeof(float));
        DZ_HIP(hipMemcpyAsync(ham->mlp_params.ptr, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice,
                              ham->stream()));
        DZ_HIP(hipStreamSynchronize(ham->stream()));
        ham->mlp_energy_scale = energy_scale;
        ham->mlp_base = base;
        ham->has_mlp = true;
        ham->mlp_fail_streak = 0;
        ham->mlp_rest = 0;
    });
}

int dzhip_ham_training_samples(dzhip_ham* ham, double energy_scale, int base, uint64_t stride, float* out,
                               uint64_t capacity, uint64_t* count) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(out && count, DZHIP_ERR_ARG, "NULL out or count");
        DZ_REQUIRE(stride >= 1 && stride < (1u << 30), DZHIP_ERR_ARG, "bad stride");
        std::lock_guard<std::mutex> lk(ham->mu);
        DeviceGuard g(ham->ctx->device);
        *count = eig_training_samples(*ham, energy_scale, base, (int)stride, out, capacity);
    });
}

namespace {
void upload_samples(MlpTrainer& t, const float* samples, uint64_t n, hipStream_t s) {
    DZ_HIP(hipMemcpyAsync(t.samples.ptr, samples, 6 * n * sizeof(float), hipMemcpyHostToDevice, s));
}
}  // namespace

int dzhip_mlp_loss_grad(dzhip_ctx* ctx, const float* samples, uint64_t n, const float* params, double* grad_out,
                        double* loss_out) {
    return guarded([&] {
        DZ_REQUIRE(ctx, DZHIP_ERR_STATE, "NULL ctx");
        DZ_REQUIRE(samples && params && n > 0, DZHIP_ERR_ARG, "NULL samples/params or n = 0");
        DZ_REQUIRE(n < (1ull << 31), DZHIP_ERR_SIZE, "n must be below 2^31");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        hipStream_t s = ctx->stream;
        MlpTrainer t(n);
        upload_samples(t, samples, n, s);
        DZ_HIP(hipMemcpyAsync(t.prm.ptr, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, s));
        t.loss_grad(n, 1, 0, 0, n, s);
        std::vector<double> gh(kMlpParams + 1);
        DZ_HIP(hipMemcpyAsync(gh.data(), t.g.ptr, gh.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        DZ_HIP(hipStreamSynchronize(s));
        if (grad_out) std::memcpy(grad_out, gh.data(), kMlpParams * sizeof(double));
        if (loss_out) *loss_out = gh[kMlpParams];
    });
}

int dzhip_mlp_train(dzhip_ctx* ctx, const float* samples, uint64_t n, float* params, uint64_t epochs, uint64_t batch,
                    double lr, uint64_t seed, double* epoch_loss) {
    return guarded([&] {
        DZ_REQUIRE(ctx, DZHIP_ERR_STATE, "NULL ctx");
        DZ_REQUIRE(samples && params && n > 0 && batch > 0, DZHIP_ERR_ARG, "NULL samples/params, n = 0 or batch = 0");
        DZ_REQUIRE(n < (1ull << 31), DZHIP_ERR_SIZE, "n must be below 2^31");
        DZ_REQUIRE(lr > 0.0, DZHIP_ERR_ARG, "lr must be > 0");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard g(ctx->device);
        hipStream_t s = ctx->stream;
        MlpTrainer t(n);
        upload_samples(t, samples, n, s);
        DZ_HIP(hipMemcpyAsync(t.prm.ptr, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, s));
        DZ_HIP(hipMemsetAsync(t.m.ptr, 0, kMlpParams * sizeof(float), s));
        DZ_HIP(hipMemsetAsync(t.v.ptr, 0, kMlpParams * sizeof(float), s));
        unsigned long long x = seed ? seed : 0x6f3d3cbeULL;
        auto next = [&]() {
            x = x * 6364136223846793005ULL + 1442695040888963407ULL;
            return x >> 17;
        };
        auto gcd = [](unsigned long long a, unsigned long long b) {
            while (b) {
                const unsigned long long r = a % b;
                a = b;
                b = r;
            }
            return a;
        };
        int step = 0;
        for (uint64_t e = 0; e < epochs; ++e) {
            // Visit order idx(j) = (mul j + add) mod n, a permutation since gcd(mul, n) = 1.
            unsigned long long mul = 1 + next() % n;
            while (gcd(mul, n) != 1) mul = 1 + next() % n;
            const unsigned long long add = next() % n;
            for (uint64_t start = 0; start < n; start += batch) {
                const uint64_t cnt = std::min<uint64_t>(batch, n - start);
                t.loss_grad(n, mul, add, start, cnt, s);
                t.adam(++step, lr, s);
            }
            if (epoch_loss) {
                t.loss_grad(n, 1, 0, 0, n, s);
                double l = 0.0;
                DZ_HIP(hipMemcpyAsync(&l, t.g.as<double>() + kMlpParams, sizeof(double), hipMemcpyDeviceToHost, s));
                DZ_HIP(hipStreamSynchronize(s));
                epoch_loss[e] = l;
            }
        }
        DZ_HIP(hipMemcpyAsync(params, t.prm.ptr, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, s));
        DZ_HIP(hipStreamSynchronize(s));
    });
}

dzhip_scf_grid* dzhip_scf_grid_create(dzhip_ctx* ctx, uint64_t nx, uint64_t ny, uint64_t nz, const double* recip,
                                      double volume, uint64_t pulay_history) {
    dzhip_scf_grid* out = nullptr;
    const int rc = guarded([&] {
        DZ_REQUIRE(ctx, DZHIP_ERR_STATE, "NULL ctx");
        DZ_REQUIRE(recip && volume > 0.0, DZHIP_ERR_ARG, "NULL recip or volume <= 0");
        DZ_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && nx <= 1024 && ny <= 1024 && nz <= 1024, DZHIP_ERR_SIZE,
                   "grid dimensions must be in 1..1024");
        DZ_REQUIRE(pulay_history <= 64, DZHIP_ERR_ARG, "pulay_history <= 64");
        DeviceGuard g(ctx->device);
        std::lock_guard<std::mutex> lk(ctx->mu);
        out = static_cast<dzhip_scf_grid*>(
            scf_grid_create(ctx, (int)nx, (int)ny, (int)nz, recip, volume, (int)pulay_history));
    });
    return rc == DZHIP_OK ? out : nullptr;
}

void dzhip_scf_grid_destroy(dzhip_scf_grid* grid) {
    if (!grid) return;
    DeviceGuard g(grid->ctx->device);
    scf_grid_destroy(grid);
}

int dzhip_scf_grid_set_ion(dzhip_scf_grid* grid, const double* v_ion, uint64_t count) {
    return guarded([&] {
        DZ_REQUIRE(grid, DZHIP_ERR_STATE, "NULL grid");
        DZ_REQUIRE(v_ion && count == grid->n, DZHIP_ERR_SIZE, "count != nx*ny*nz");
        std::lock_guard<std::mutex> lk(grid->mu);
        DeviceGuard g(grid->ctx->device);
        DZ_HIP(hipMemcpyAsync(grid->v_ion.ptr, v_ion, count * sizeof(double), hipMemcpyHostToDevice, grid->stream()));
        DZ_HIP(hipStreamSynchronize(grid->stream()));
    });
}

int dzhip_scf_grid_init(dzhip_scf_grid* grid, double nelec) {
    return guarded([&] {
        DZ_REQUIRE(grid, DZHIP_ERR_STATE, "NULL grid");
        std::lock_guard<std::mutex> lk(grid->mu);
        DeviceGuard g(grid->ctx->device);
        scf_grid_init(*grid, nelec, nullptr);
    });
}

int dzhip_scf_grid_init_rho(dzhip_scf_grid* grid, const double* rho, uint64_t count) {
    return guarded([&] {
        DZ_REQUIRE(grid, DZHIP_ERR_STATE, "NULL grid");
        DZ_REQUIRE(rho && count == grid->n, DZHIP_ERR_SIZE, "count != nx*ny*nz");
        std::lock_guard<std::mutex> lk(grid->mu);
        DeviceGuard g(grid->ctx->device);
        scf_grid_init(*grid, 0.0, rho);
    });
}

int dzhip_ham_set_local_grid(dzhip_ham* ham, dzhip_scf_grid* grid) {
    return guarded([&] {
        DZ_REQUIRE(ham && grid, DZHIP_ERR_STATE, "NULL ham or grid");
        DZ_REQUIRE(ham->ngrid == grid->n && ham->nx == grid->nx && ham->ny == grid->ny, DZHIP_ERR_SIZE,
                   "grid dimensions differ");
        DZ_REQUIRE(ham->ctx->device == grid->ctx->device, DZHIP_ERR_ARG, "ham and grid on different devices");
        // Both calls that take a grid and a ham lock them together (std::scoped_lock orders the
        // acquisition), so two threads calling them on the same pair cannot deadlock.
        std::scoped_lock lk(grid->mu, ham->mu);
        DeviceGuard g(ham->ctx->device);
        scf_grid_local(*grid);
        DZ_HIP(hipStreamSynchronize(grid->stream()));
        ham_set_local_dev(*ham, grid->local.as<double>());
        DZ_HIP(hipStreamSynchronize(ham->stream()));
    });
}

int dzhip_scf_grid_add_solution_density(dzhip_scf_grid* grid, dzhip_ham* ham, const double* weights,
                                        uint64_t nbands) {
    return guarded([&] {
        DZ_REQUIRE(ham && grid, DZHIP_ERR_STATE, "NULL ham or grid");
        DZ_REQUIRE(ham->ngrid == grid->n, DZHIP_ERR_SIZE, "grid dimensions differ");
        DZ_REQUIRE(nbands == 0 || weights, DZHIP_ERR_ARG, "NULL weights");
        if (nbands == 0) return;
        std::scoped_lock lk(grid->mu, ham->mu);
        DeviceGuard g(ham->ctx->device);
        uint64_t cols = 0;
        const double2* x = eig_solution(*ham, &cols);
        DZ_REQUIRE(x && nbands <= cols, DZHIP_ERR_STATE, "no eigensolver solution with that many bands");
        hipStream_t s = ham->stream();
        ham->dens_rho.ensure(ham->ngrid * sizeof(double));
        ham->dens_w.ensure(nbands * sizeof(double));
        DZ_HIP(hipMemcpyAsync(ham->dens_w.ptr, weights, nbands * sizeof(double), hipMemcpyHostToDevice, s));
        ham_density_dev(*ham, x, nbands, ham->dens_w.as<double>(), ham->dens_rho.as<double>(), false);
        DZ_HIP(hipStreamSynchronize(s));
        scf_grid_add(*grid, ham->dens_rho.as<double>());
        DZ_HIP(hipStreamSynchronize(grid->stream()));
    });
}

int dzhip_scf_grid_hxc(dzhip_scf_grid* grid, double* out, uint64_t n_out) {
    return guarded([&] {
        DZ_REQUIRE(grid, DZHIP_ERR_STATE, "NULL grid");
        DZ_REQUIRE(out && n_out >= 4, DZHIP_ERR_ARG, "out needs 4 entries");
        std::lock_guard<std::mutex> lk(grid->mu);
        DeviceGuard g(grid->ctx->device);
        scf_grid_hxc(*grid, out);
    });
}

int dzhip_scf_grid_set_dielectric(dzhip_scf_grid* grid, double diemac, double dielng) {
    return guarded([&] {
        DZ_REQUIRE(grid, DZHIP_ERR_STATE, "NULL grid");
        DZ_REQUIRE(diemac >= 1.0 && dielng > 0.0, DZHIP_ERR_ARG, "diemac >= 1 and dielng > 0");
        std::lock_guard<std::mutex> lk(grid->mu);
        grid->diemac = diemac;
        grid->dielng = dielng;
    });
}

int dzhip_scf_grid_set_xc(dzhip_scf_grid* grid, int functional) {
    return guarded([&] {
        DZ_REQUIRE(grid, DZHIP_ERR_STATE, "NULL grid");
        DZ_REQUIRE(functional == DZHIP_XC_LDA_PZ || functional == DZHIP_XC_PBE, DZHIP_ERR_ARG,
                   "functional must be DZHIP_XC_LDA_PZ or DZHIP_XC_PBE");
        std::lock_guard<std::mutex> lk(grid->mu);
        DeviceGuard g(grid->ctx->device);
        scf_grid_set_xc(*grid, functional);
    });
}

int dzhip_scf_grid_mix(dzhip_scf_grid* grid, double beta, int pulay) {
    return guarded([&] {
        DZ_REQUIRE(grid, DZHIP_ERR_STATE, "NULL grid");
        std::lock_guard<std::mutex> lk(grid->mu);
        DeviceGuard g(grid->ctx->device);
        scf_grid_mix(*grid, beta, pulay);
        DZ_HIP(hipStreamSynchronize(grid->stream()));
    });
}

int dzhip_scf_grid_read(dzhip_scf_grid* grid, int which, double* out, uint64_t count) {
    return guarded([&] {
        DZ_REQUIRE(grid, DZHIP_ERR_STATE, "NULL grid");
        DZ_REQUIRE(out && count == grid->n, DZHIP_ERR_SIZE, "count != nx*ny*nz");
        DevBuf* src[] = {&grid->rho_in, &grid->v_in, &grid->v_out, &grid->local, &grid->v_ion};
        DZ_REQUIRE(which >= 0 && which < 5, DZHIP_ERR_ARG, "which must be DZHIP_GRID_*");
        std::lock_guard<std::mutex> lk(grid->mu);
        DeviceGuard g(grid->ctx->device);
        if (which == DZHIP_GRID_LOCAL) scf_grid_local(*grid);
        DZ_HIP(hipMemcpyAsync(out, src[which]->ptr, count * sizeof(double), hipMemcpyDeviceToHost, grid->stream()));
        DZ_HIP(hipStreamSynchronize(grid->stream()));
    });
}

int dzhip_ham_set_tuning(dzhip_ham* ham, int xline_split) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(xline_split == 0 || xline_code_usable(*ham, xline_split, false), DZHIP_ERR_ARG,
                   "xline_split must divide ny, or be -1..-3 (-4..-6, -7..-9, -10..-12) with a (lane-pair, "
                   "lane-triple, residue-sequential) register kernel for this context, or -13 where the "
                   "convolution kernel runs (0 = autotune)");
        std::lock_guard<std::mutex> lk(ham->mu);
        ham->tune_fixed = xline_split;
        ham->xline_split[0] = ham->xline_split[1] = 0;
    });
}

int dzhip_ham_set_plane_kernel(dzhip_ham* ham, int kind) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        DZ_REQUIRE(kind == 0 || kind == kPlaneTile || (kind == kPlaneZreg && plane_zreg_usable(*ham)),
                   DZHIP_ERR_ARG,
                   "plane kernel must be 0 (autotune), 1 (full-tile LDS kernels) or 2 (z-register kernels, "
                   "nz in 12..96 divisible by 3 with nz / 3 = 2^a 3^b 5^c and ny <= 126)");
        std::lock_guard<std::mutex> lk(ham->mu);
        ham->plane_fixed = kind;
        ham->plane_kind[0] = ham->plane_kind[1] = 0;
    });
}

int dzhip_ham_set_overlap(dzhip_ham* ham, int on) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        std::lock_guard<std::mutex> lk(ham->mu);
        ham->overlap = on != 0;
    });
}

int dzhip_ham_info(dzhip_ham* ham, uint64_t* out, uint64_t n_out) {
    return guarded([&] {
        DZ_REQUIRE(ham, DZHIP_ERR_STATE, "NULL ham");
        auto slabs = [](int code) { return code > 0 ? (uint64_t)code : 0; };
        auto reg = [](int code) { return code < 0 ? (uint64_t)xline_reg_threads(code) : 0u; };
        auto waves = [](int code) { return code < 0 ? (uint64_t)xline_reg_lanes(code) : 0u; };
        auto kind = [](int code) { return code < 0 ? (uint64_t)(1 + xline_reg_kind(code)) : 0u; };
        const uint64_t v[15] = {ham->npw, ham->ngrid, (uint64_t)ham->nplanes, (uint64_t)ham->nproj,
                                ham->device_bytes(), slabs(ham->xline_split[1]), slabs(ham->xline_split[0]),
                                reg(ham->xline_split[1]), reg(ham->xline_split[0]), waves(ham->xline_split[1]),
                                waves(ham->xline_split[0]), kind(ham->xline_split[1]), kind(ham->xline_split[0]),
                                (uint64_t)ham->plane_kind[1], (uint64_t)ham->plane_kind[0]};
        for (uint64_t i = 0; i < n_out && i < 15; ++i) out[i] = v[i];
    });
}

}  // extern "C"
