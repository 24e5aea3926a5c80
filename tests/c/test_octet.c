/* The reference's known-answer test (eelregit/fftgal src/tests/octet.c)
 * rewritten for this library: the same calls through the unchanged C API
 * (include/fft.h, gal.h), built with gcc against libfftgal_amd.so.
 * One particle at (-8,-8,-8) in a 2^3 grid of side 2 with offset (7,7,7):
 * paint, forward transform with the offset phase, inverse transform, each
 * checked against its closed form at eps = 1e-15. */
#include <stdio.h>
#include <math.h>
#include "fft.h"
#include "gal.h"

static int check(const char *what, double got, double exp, double eps) {
    double err = got - exp;
    if (fabs(err) >= eps) {
        printf("FAIL %s: got %.17g expected %.17g\n", what, got, exp);
        return 1;
    }
    return 0;
}

int main(void) {
    const double eps = 1e-15;
    int fails = 0;
    int Ng = 2;
    double L = 2;
    fft_t *grid = fft_init(Ng, L, "FFTW_ESTIMATE");
    gal_t *part = gal_init(1, L * L * L);
    part->x[0] = part->y[0] = part->z[0] = -8;
    double offset[3] = {7, 7, 7};
    fft_p2g(grid, part, offset);
    for (int i = 0; i < 3; ++i) fails += check("offset", offset[i], -1, eps);
    for (int i = 0; i < Ng; ++i)
        for (int j = 0; j < Ng; ++j)
            for (int k = 0; k < Ng; ++k)
                fails += check("paint", F(grid, i, j, k), 8.0 / 27 * (i + 1) * (j + 1) * (k + 1), eps);
    fft_x2k(grid, 1);
    for (int i = 0; i < Ng; ++i)
        for (int j = 0; j < Ng; ++j)
            for (int k = 0; k < Ng / 2; ++k) {
                fails += check("x2k re", F_Re(grid, i, j, k), 8.0 / (1 + 2 * i) / (1 + 2 * j) / (1 + 2 * k), eps);
                fails += check("x2k im", F_Im(grid, i, j, k), 0, eps);
            }
    fft_k2x(grid, 0);
    for (int i = 0; i < Ng; ++i)
        for (int j = 0; j < Ng; ++j)
            for (int k = 0; k < Ng; ++k)
                fails += check("k2x", F(grid, i, j, k), 64.0 / 27 / (i + 1) / (j + 1) / (k + 1), eps);
    fft_free(grid);
    gal_free(part);
    if (fails)
        printf("* test FAILED (%d checks)\n", fails);
    else
        printf("* test PASSED on eps = %e\n", eps);
    return fails != 0;
}
