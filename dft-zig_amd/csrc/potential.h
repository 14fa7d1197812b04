// Device-resident SCF grid state (density, potentials, Pulay history); see potential.hip.
#pragma once

#include <mutex>
#include <vector>

#include "ham.h"

namespace dzhip {

struct ScfGrid {
    Context* ctx = nullptr;
    int nx = 0, ny = 0, nz = 0;
    size_t n = 0;
    double volume = 0.0;
    AxisTables tab;
    DevBuf g2, work, rho_in, rho_out, v_in, v_out, v_ion, local, part, hv, hr, slots, coef;
    // Exchange-correlation functional (DZHIP_XC_LDA_PZ or DZHIP_XC_PBE), reciprocal rows, and the
    // PBE workspaces: second complex grid, gradient/b components [3][N], f_xc and v_xc [N].
    int xc = DZHIP_XC_LDA_PZ;
    double3 recip_rows[3] = {};
    DevBuf work2, grad, gga_f, gga_v;
    int hist_cap = 0;
    double diemac = 1.0, dielng = 1.0;  // model dielectric preconditioner of the Pulay residual
    std::vector<int> order;              // Pulay ring slots, oldest first
    std::vector<std::vector<double>> b;  // overlaps <R_i|R_j>/N by slot
    std::mutex mu;

    hipStream_t stream() const { return ctx->stream; }
    double reduce(const double* part_dev, int nblocks, int stride, int comp);
    double dot(const double* a, const double* c);
    // v_out = V_H[rho] + v_xc[rho]; out = E_H, E_xc, int v_in rho (Ry), RMS(rho - rho_ref).
    void hxc(const double* rho, const double* rho_ref, const double* vin, double* vout, double out[4]);
    // PBE f_xc -> gga_f and v_xc -> gga_v for density rho (uses work, work2, grad).
    void gga(const double* rho);
};

}  // namespace dzhip
