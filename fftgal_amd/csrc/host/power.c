/* Drop-in estimators (include/power.h) and their array forms
 * (include/fftgal_ext.h).  The per-mode work runs in the binning kernel
 * (fgk_bin_pgrid_f64); this file sets up the bins, normalises the sums and
 * writes the reference's text format (src/power.c:78-99, 155-176). */
#define _GNU_SOURCE
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include "internal.h"
#include "../../../include/power.h"
#include "../../../include/fftgal_ext.h"

/* sqrt(3) to double precision (gsl_math.h's M_SQRT3, used at power.c:36) */
#define FFTGAL_SQRT3 1.73205080756887729352744634151

/* src/power.c:34-36 */
int fftgal_nk(int Ng, double L, double dK) {
    double KF = 2 * M_PI / L;
    double dKinv = KF / dK;
    return (int)floor(FFTGAL_SQRT3 * (Ng / 2) * dKinv) + 1;
}

int fftgal_NK(int Ng, double L, double dK) { return fftgal_nk(Ng, L, dK); }

/* src/power.c:12-27 */
void P_params(fft_t *grid, long Np, double V, long Nrand) {
    double t = fftgal_now();
    fft_dev *d = fft_dev_of(grid);
    double Vinv = 1 / V;
    double alpha = Nrand <= 0 ? 0 : (double)Np / Nrand;
    double Pshot = (1 + alpha) * V / Np;
    fftgal_dev_pull(grid);
    DI_CHECK(fgk_P_f64(d->d_f, grid->Ng, Vinv, Pshot, d->s));
    fftgal_dev_push(grid);
    if (fftgal_verbose()) {
        fgk_stream_sync(d->s);
        fprintf(stderr, "P() %.2fs\n", fftgal_now() - t);
    }
}

void P(fft_t *grid, gal_t *part) {
    P_params(grid, part->Np, part->V, part->rand ? part->rand->Np : 0);
}

/* run the binning kernel over the P(k) grid; returns NK and fills the raw
 * sums (bins: layout of fgk_binspec) and counts */
static int run_bins(fft_t *grid, const double los[3], double dK, int Nmu, int do_pl,
                    double **bins_out, long **counts_out, fgk_binspec *spec) {
    fft_dev *d = fft_dev_of(grid);
    const int Ng = grid->Ng;
    double KF = 2 * M_PI / grid->L;
    spec->Ng = Ng;
    spec->NK = fftgal_nk(Ng, grid->L, dK);
    spec->Nmu = Nmu;
    spec->dKinv = KF / dK;
    spec->los[0] = los[0];
    spec->los[1] = los[1];
    spec->los[2] = los[2];
    spec->do_pl = do_pl;
    spec->geometry = 1;
    spec->bk = 0;
    spec->fs_n1 = 0;
    size_t nb = fgk_bins_doubles(spec), nc = fgk_bins_counts(spec);
    double *dbins = NULL;
    int64_t *dcnt = NULL;
    DI_CHECK(fgk_malloc((void **)&dbins, nb * sizeof(double)));
    DI_CHECK(fgk_malloc((void **)&dcnt, nc * sizeof(int64_t)));
    DI_CHECK(fgk_fill_zero(dbins, nb * sizeof(double), d->s));
    DI_CHECK(fgk_fill_zero(dcnt, nc * sizeof(int64_t), d->s));
    fftgal_dev_pull(grid);
    DI_CHECK(fgk_bin_pgrid_f64(d->d_f, spec, dbins, dcnt, d->s));
    double *hb = (double *)malloc(nb * sizeof(double));
    long *hc = (long *)malloc(nc * sizeof(long));
    DI_ASSERT(hb && hc, "Pl/Pmu: out of memory");
    DI_CHECK(fgk_d2h(hb, dbins, nb * sizeof(double), d->s));
    DI_CHECK(fgk_d2h(hc, dcnt, nc * sizeof(int64_t), d->s));
    DI_CHECK(fgk_stream_sync(d->s));
    fgk_free(dbins);
    fgk_free(dcnt);
    *bins_out = hb;
    *counts_out = hc;
    return spec->NK;
}

/* src/power.c:30-76 as arrays */
int Pl_bins(fft_t *grid, double los[3], double dK, double *K, long *N, double *P) {
    hat(los);
    fgk_binspec spec;
    double *b;
    long *c;
    int NK = run_bins(grid, los, dK, 0, 1, &b, &c, &spec);
    const double KF = 2 * M_PI / grid->L;
    const double *Ksum = b + (size_t)NK * 4;   /* Nmu = 0: K sums follow the Pl sums */
    long Ntot = 0;
    for (int iK = 0; iK < NK; ++iK) {
        N[iK] = c[iK];
        K[iK] = Ksum[iK] * (KF / N[iK]);
        for (int l = 0; l < 4; ++l) P[iK * 4 + l] = b[iK * 4 + l] * ((4 * l + 1.0) / N[iK]);
        Ntot += N[iK];
    }
    free(b);
    free(c);
    DI_ASSERT(Ntot == (long)grid->Ng * grid->Ng * grid->Ng - 1, "Pl: sum of mode counts != Ng^3-1");
    return NK;
}

/* Pl_bins with its device sums in the library's fixed bin scratch instead
 * of per-call allocations: no hipMalloc/hipFree and no host malloc.  The
 * raw sums are copied straight into the caller's K, N and P and normalised
 * there.  Errors are returned (-1, fftgal_last_error()), not aborted on. */
int Pl_bins_noalloc(fft_t *grid, double los[3], double dK, double *K, long *N, double *P) {
    if (!grid || !K || !N || !P || !(dK > 0)) {
        fgk_set_error("Pl_bins_noalloc: null argument or dK <= 0");
        return -1;
    }
    fft_dev *d = fft_dev_of(grid);
    const int Ng = grid->Ng;
    const double KF = 2 * M_PI / grid->L;
    fgk_binspec spec;
    memset(&spec, 0, sizeof spec);
    spec.Ng = Ng;
    spec.NK = fftgal_nk(Ng, grid->L, dK);
    spec.Nmu = 0;
    spec.dKinv = KF / dK;
    spec.do_pl = 1;
    spec.geometry = 1;
    const int NK = spec.NK;
    /* layout with Nmu = 0: Pl sums [NK][4], then K sums [NK]; counts [NK] */
    const size_t nb = fgk_bins_doubles(&spec), nc = fgk_bins_counts(&spec);
    const size_t cnt_off = (nb * sizeof(double) + 255) & ~(size_t)255;
    void *base;
    size_t cap;
    if (fgk_bin_scratch(&base, &cap)) return -1;
    /* 5 doubles and one count per bin, plus at most 255 bytes of alignment */
    const long nk_max = (long)((cap - 256) / (5 * sizeof(double) + sizeof(int64_t)));
    if (NK > nk_max) {
        fgk_set_error("Pl_bins_noalloc: NK = %d bins exceed the %ld that its %zu-byte bin "
                      "scratch holds", NK, nk_max, cap);
        return -1;
    }
    hat(los);
    spec.los[0] = los[0];
    spec.los[1] = los[1];
    spec.los[2] = los[2];
    double *dbins = (double *)base;
    int64_t *dcnt = (int64_t *)((char *)base + cnt_off);
    if (fgk_fill_zero(base, cnt_off + nc * sizeof(int64_t), d->s)) return -1;
    fftgal_dev_pull(grid);
    if (fgk_bin_pgrid_f64(d->d_f, &spec, dbins, dcnt, d->s)) return -1;
    if (fgk_d2h(P, dbins, (size_t)NK * 4 * sizeof(double), d->s)) return -1;
    if (fgk_d2h(K, dbins + (size_t)NK * 4, (size_t)NK * sizeof(double), d->s)) return -1;
    if (fgk_d2h(N, dcnt, (size_t)NK * sizeof(int64_t), d->s)) return -1;
    if (fgk_stream_sync(d->s)) return -1;
    long Ntot = 0;
    for (int iK = 0; iK < NK; ++iK) {
        K[iK] *= KF / N[iK];
        for (int l = 0; l < 4; ++l) P[iK * 4 + l] *= (4 * l + 1.0) / N[iK];
        Ntot += N[iK];
    }
    if (Ntot != (long)Ng * Ng * Ng - 1) {
        fgk_set_error("Pl_bins_noalloc: sum of mode counts %ld != Ng^3-1", Ntot);
        return -1;
    }
    return NK;
}

/* src/power.c:104-153 as arrays, Nmu wedges */
int Pmu_bins(fft_t *grid, double los[3], double dK, int Nmu, double *K, long *N, double *P) {
    DI_ASSERT(Nmu >= 1, "Pmu: Nmu must be >= 1");
    hat(los);
    fgk_binspec spec;
    double *b;
    long *c;
    int NK = run_bins(grid, los, dK, Nmu, 0, &b, &c, &spec);
    const int W = Nmu + 1;
    const double KF = 2 * M_PI / grid->L;
    /* layout of fgk_binspec: Pl sums [NK][4], Pmu sums [NK][W], K [NK], Kmu [NK][W] */
    const double *Psum = b + (size_t)NK * 4;
    const double *Kmu = b + (size_t)NK * (4 + W) + NK;
    const long *Nmu_c = c + NK;                               /* [NK][W] */
    long Ntot = 0;
    for (int iK = 0; iK < NK; ++iK) {
        for (int m = 0; m < Nmu; ++m) {
            double ks = Kmu[iK * W + m], ps = Psum[iK * W + m];
            long n = Nmu_c[iK * W + m];
            if (m == Nmu - 1) {              /* corner case mu = 1 (power.c:139-143) */
                ks += Kmu[iK * W + Nmu];
                ps += Psum[iK * W + Nmu];
                n += Nmu_c[iK * W + Nmu];
            }
            if (n) {
                ks *= KF / n;
                ps /= n;
                Ntot += n;
            }
            K[iK * Nmu + m] = ks;
            N[iK * Nmu + m] = n;
            P[iK * Nmu + m] = ps;
        }
    }
    free(b);
    free(c);
    DI_ASSERT(Ntot == (long)grid->Ng * grid->Ng * grid->Ng - 1, "Pmu: sum of mode counts != Ng^3-1");
    return NK;
}

/* src/power.c:30-101 */
void Pl(fft_t *grid, gal_t *part, double los[3], double dK, char *file) {
    double t = fftgal_now();
    const int NK = fftgal_nk(grid->Ng, grid->L, dK);
    const int Nl = 4;
    double *K = (double *)malloc((size_t)NK * sizeof(double));
    long *N = (long *)malloc((size_t)NK * sizeof(long));
    double *P = (double *)malloc((size_t)NK * Nl * sizeof(double));
    DI_ASSERT(K && N && P, "Pl: out of memory");
    Pl_bins(grid, los, dK, K, N, P);
    double alpha = part->rand == NULL ? 0 : (double)part->Np / part->rand->Np;
    FILE *fp = fopen(file, "w");
    DI_ASSERT(fp != NULL, "Pl: cannot open output file");
    fprintf(fp, "# NK %d\n", NK);
    fprintf(fp, "# dK %g\n", dK);
    fprintf(fp, "# Nl %d\n", Nl);
    fprintf(fp, "# Ng %d\n", grid->Ng);
    fprintf(fp, "# L %g\n", grid->L);
    fprintf(fp, "# Np %ld\n", part->Np);
    fprintf(fp, "# V %g\n", part->V);
    fprintf(fp, "# alpha %g\n", alpha);
    fprintf(fp, "# los[3] %g %g %g\n", los[0], los[1], los[2]);
    fprintf(fp, "# K N");
    for (int l = 0; l < Nl; ++l) fprintf(fp, " P%d", 2 * l);
    fprintf(fp, "\n");
    for (int iK = 0; iK < NK; ++iK) {
        fprintf(fp, "%e %9ld", K[iK], N[iK]);
        for (int l = 0; l < Nl; ++l) fprintf(fp, " % e", P[iK * Nl + l]);
        fprintf(fp, "\n");
    }
    fclose(fp);
    free(K);
    free(N);
    free(P);
    if (fftgal_verbose()) fprintf(stderr, "Pl() %.2fs, saved to %s\n", fftgal_now() - t, file);
}

/* src/power.c:104-178 (Nmu = 3 as in power.c:111) */
void Pmu(fft_t *grid, gal_t *part, double los[3], double dK, char *file) {
    double t = fftgal_now();
    const int NK = fftgal_nk(grid->Ng, grid->L, dK);
    const int Nmu = 3;
    double *K = (double *)malloc((size_t)NK * Nmu * sizeof(double));
    long *N = (long *)malloc((size_t)NK * Nmu * sizeof(long));
    double *P = (double *)malloc((size_t)NK * Nmu * sizeof(double));
    DI_ASSERT(K && N && P, "Pmu: out of memory");
    Pmu_bins(grid, los, dK, Nmu, K, N, P);
    double alpha = part->rand == NULL ? 0 : (double)part->Np / part->rand->Np;
    FILE *fp = fopen(file, "w");
    DI_ASSERT(fp != NULL, "Pmu: cannot open output file");
    fprintf(fp, "# NK %d\n", NK);
    fprintf(fp, "# dK %g\n", dK);
    fprintf(fp, "# Nmu %d\n", Nmu);
    fprintf(fp, "# Ng %d\n", grid->Ng);
    fprintf(fp, "# L %g\n", grid->L);
    fprintf(fp, "# Np %ld\n", part->Np);
    fprintf(fp, "# V %g\n", part->V);
    fprintf(fp, "# alpha %g\n", alpha);
    fprintf(fp, "# los[3] %g %g %g\n", los[0], los[1], los[2]);
    fprintf(fp, "#");
    for (int m = 0; m < Nmu; ++m) fprintf(fp, " K%d N%d P%d", m, m, m);
    fprintf(fp, "\n");
    for (int iK = 0; iK < NK; ++iK) {
        fprintf(fp, "%e %9ld % e", K[iK * Nmu], N[iK * Nmu], P[iK * Nmu]);
        for (int m = 1; m < Nmu; ++m)
            fprintf(fp, " %e %9ld % e", K[iK * Nmu + m], N[iK * Nmu + m], P[iK * Nmu + m]);
        fprintf(fp, "\n");
    }
    fclose(fp);
    free(K);
    free(N);
    free(P);
    if (fftgal_verbose()) fprintf(stderr, "Pmu() %.2fs, saved to %s\n", fftgal_now() - t, file);
}
