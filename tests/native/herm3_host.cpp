// Host driver for herm3_lowest: reads 3x3 Hermitian matrices (18 doubles each, row-major
// re/im pairs) from stdin, prints the lowest eigenvalue and eigenvector (7 doubles per line).
#include <cstdio>

#include "herm3.h"

int main() {
    double buf[18];
    while (true) {
        for (int i = 0; i < 18; ++i)
            if (std::scanf("%lf", &buf[i]) != 1) return 0;
        dzhip::cplx a[9];
        for (int i = 0; i < 9; ++i) a[i] = dzhip::cplx(buf[2 * i], buf[2 * i + 1]);
        double lam = 0.0;
        dzhip::cplx v[3];
        dzhip::herm3_lowest(a, &lam, v);
        std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", lam, v[0].real(), v[0].imag(), v[1].real(),
                    v[1].imag(), v[2].real(), v[2].imag());
    }
}
