// Lowest eigenpair of a 3x3 Hermitian matrix by cyclic complex Jacobi rotations (host code).
// Used by the CG eigensolver's three-term Rayleigh-Ritz step, where the reference calls LAPACK
// zheev on the 3x3 projected matrix (src/features/linalg/cg/cg.zig:346-362).
#pragma once

#include <cmath>
#include <complex>

namespace dzhip {

// a_in row-major; returns the smallest eigenvalue and a unit eigenvector.
using cplx = std::complex<double>;

inline void herm3_lowest(const cplx a_in[9], double* lam, cplx vec[3]) {
    cplx a[3][3], v[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            a[i][j] = a_in[3 * i + j];
            v[i][j] = (i == j) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = i + 1; j < 3; ++j) off += std::norm(a[i][j]);
        if (off < 1e-30) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double apq = std::abs(a[p][q]);
                if (apq < 1e-300) continue;
                // Unitary rotation zeroing a[p][q]: phase e = a_pq/|a_pq|, angle from the real 2x2.
                const cplx e = a[p][q] / apq;
                const double app = a[p][p].real(), aqq = a[q][q].real();
                const double theta = 0.5 * std::atan2(2.0 * apq, aqq - app);
                const double c = std::cos(theta), sn = std::sin(theta);
                // Columns p, q: new_p = c col_p - sn conj(e) col_q... applied as R with
                // R_pp = c, R_pq = sn e, R_qp = -sn conj(e), R_qq = c;  A <- R^H A R, V <- V R.
                const cplx rpp = c, rpq = sn * e, rqp = -sn * std::conj(e), rqq = c;
                for (int k = 0; k < 3; ++k) {  // A R
                    const cplx akp = a[k][p], akq = a[k][q];
                    a[k][p] = akp * rpp + akq * rqp;
                    a[k][q] = akp * rpq + akq * rqq;
                }
                for (int k = 0; k < 3; ++k) {  // R^H (A R)
                    const cplx apk = a[p][k], aqk = a[q][k];
                    a[p][k] = std::conj(rpp) * apk + std::conj(rqp) * aqk;
                    a[q][k] = std::conj(rpq) * apk + std::conj(rqq) * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const cplx vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = vkp * rpp + vkq * rqp;
                    v[k][q] = vkp * rpq + vkq * rqq;
                }
            }
    }
    int lo = 0;
    for (int i = 1; i < 3; ++i)
        if (a[i][i].real() < a[lo][lo].real()) lo = i;
    *lam = a[lo][lo].real();
    for (int k = 0; k < 3; ++k) vec[k] = v[k][lo];
}


}  // namespace dzhip
