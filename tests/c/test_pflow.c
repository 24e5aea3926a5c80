/* The power-spectrum driver flow of the reference (eelregit/fftgal
 * src/app/P.c:19-97) as a test program of this library, written against the
 * unchanged C API (include/fft.h, gal.h, io.h, power.h) and linked with
 * libfftgal_amd.so:
 *   one QPM catalog (gal_loadqpm_cubic), a grid of box L/fold;
 *   redshift space: for each of the six lines of sight with components in
 *     {0, 1} (all but (0,0,0) and (1,1,1)): gal_rsd with aH = a H(a),
 *     fft_p2k, P, Pl and Pmu into text files;
 *   real space: fft_p2k once, fft_exportf, then per line of sight
 *     fft_importf, P, Pl and Pmu.
 * Output files are named as the reference names them.
 * usage: test_pflow Ng L fold dK qpm_file a outdir */
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include "fft.h"
#include "gal.h"
#include "io.h"
#include "power.h"

/* H(a) in km/s/(Mpc/h) for Omega_m = 0.29, flat (src/app/P.c:13-17) */
static double hubble(double a) {
    const double om = 0.29;
    return 100 * sqrt(om / (a * a * a) + (1 - om));
}

/* the lines of sight (i, j, k) in {0,1}^3 without (0,0,0) and (1,1,1), in
 * the reference's loop order */
static int lines_of_sight(int los[6][3]) {
    int n = 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                if ((i + j + k) == 0 || (i + j + k) == 3) continue;
                los[n][0] = i;
                los[n][1] = j;
                los[n][2] = k;
                ++n;
            }
    return n;
}

int main(int argc, char **argv) {
    if (argc != 8) {
        fprintf(stderr, "usage: %s Ng L fold dK qpm_file a outdir\n", argv[0]);
        return 2;
    }
    const int Ng = atoi(argv[1]);
    const double L = atof(argv[2]);
    const int fold = atoi(argv[3]);
    const double dK = atof(argv[4]);
    char *infile = argv[5];
    const double a = atof(argv[6]);
    const char *outdir = argv[7];
    char name[4096];

    fft_t *grid = fft_init(Ng, L / fold, "FFTW_ESTIMATE");
    gal_t *part = gal_loadqpm_cubic(infile, L);
    int los[6][3];
    const int nlos = lines_of_sight(los);

    for (int r = 0; r < nlos; ++r) {       /* redshift space */
        double l[3] = {los[r][0], los[r][1], los[r][2]};
        gal_t *rsd = gal_rsd(part, l, a * hubble(a));
        fft_p2k(grid, rsd);
        P(grid, rsd);
        snprintf(name, sizeof name, "%s/Pl_rsd1_los%d%d%d_fd%d.txt", outdir, los[r][0],
                 los[r][1], los[r][2], fold);
        Pl(grid, rsd, l, dK, name);
        snprintf(name, sizeof name, "%s/Pmu_rsd1_los%d%d%d_fd%d.txt", outdir, los[r][0],
                 los[r][1], los[r][2], fold);
        Pmu(grid, rsd, l, dK, name);
        gal_free(rsd);
    }

    fft_p2k(grid, part);                   /* real space: one transform, reused */
    double *fk = fft_exportf(grid);
    for (int r = 0; r < nlos; ++r) {
        double l[3] = {los[r][0], los[r][1], los[r][2]};
        fft_importf(grid, fk);
        P(grid, part);
        snprintf(name, sizeof name, "%s/Pl_rsd0_los%d%d%d_fd%d.txt", outdir, los[r][0],
                 los[r][1], los[r][2], fold);
        Pl(grid, part, l, dK, name);
        snprintf(name, sizeof name, "%s/Pmu_rsd0_los%d%d%d_fd%d.txt", outdir, los[r][0],
                 los[r][1], los[r][2], fold);
        Pmu(grid, part, l, dK, name);
    }
    free(fk);
    fft_free(grid);
    gal_free(part);
    printf("test_pflow done: %d lines of sight\n", nlos);
    return 0;
}
