/* Mixed-precision power-spectrum pipeline (include/fftgal_batch.h).
 *
 * One fftgal_power_t holds everything a catalog needs on the device:
 *   int64 fixed-point grids  gA (plain) and gB (dual)     Ng^3 each, gB = gA + Ng^3
 *   (and, when device memory allows, a second such set: catalogs alternate
 *   between the two, and each paint clears the other set while it runs, so
 *   that the row FFT need not clear the grid it reads)
 *   FP32 spectra             fA, fB  (packed half spectrum) Ng^3 floats each, fB after fA
 *   (one allocation each, so that the FFT passes run over both grids in one launch)
 *   paint plan (sort buffers), FFT plan (FP32 twiddles from FP64),
 *   offset-phase and deconvolution tables (FP64, built as src/fft.c:204-245),
 *   bin sums (FP64) and the geometry of the bins (K sums, mode counts),
 *   computed once at creation because they depend only on Ng, L, dK, los
 *   and Nmu.
 * An enqueue runs: zero bins -> fixed-point paint -> rows r2c from int64,
 * cols j, cols i (each over both grids) -> fused bin kernel -> copy sums to
 * pinned host memory.  The sequence is captured in a hipGraph keyed by the
 * catalog pointers, Np and V, and replayed while they stay the same.
 */
#define _GNU_SOURCE
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include "../fgk.h"
#include "../../../include/fftgal_batch.h"
#include "tune.h"

/* paint stages of the fgk layer (fgk_paint_nstages) */
#define NPAINT 7

int fftgal_nk(int Ng, double L, double dK);
void fftgal_winv_table(int Ng, double *winv);
void fftgal_phase_table(int Ng, const double offset[3], int onoff, double *phase);

struct fftgal_power {
    fftgal_power_desc d;
    int NK, W;
    double KF;
    double los[3];
    fgk_stream_t s;
    fgk_fftplan *plan;
    fgk_paintplan *paint;
    int64_t *gA, *gB;
    int64_t *gA2;              /* second grid set (its dual grid at gA2 + Ng^3), or NULL */
    int parity;                /* grid set of the next enqueue (0: gA, 1: gA2) */
    float *fA, *fB;
    double *d_phase, *d_winv;
    fgk_binspec spec;          /* per-run sums */
    int64_t *d_bins;           /* fixed-point pairs (fgk.h), 2 * nbins */
    size_t nbins;              /* values */
    int64_t *h_bins;           /* pinned copy */
    double *h_sums;            /* the values in FP64 */
    /* geometry */
    double *gK, *gKmu;
    long *gN, *gNmu;
    /* last enqueue */
    long Np;
    double V;
    int grids_dirty;
    /* graphs, one per grid set */
    int use_graph;
    fgk_graph_t graph[2];
    const float *kx[2], *ky[2], *kz[2];
    long kNp[2];
    double kV[2];
    /* tuning */
    fftgal_tuning tune;
};

const char *fftgal_last_error(void) { return fgk_last_error(); }

#define TRY(call)                 \
    do {                          \
        int rc_ = (call);         \
        if (rc_) return rc_;      \
    } while (0)

/* K sums and mode counts of a binning (they depend only on Ng, L, dK, los
 * and Nmu), normalised by nothing yet: gK[NK], gN[NK], gKmu[NK*W],
 * gNmu[NK*W] */
static int bin_geometry(const fgk_binspec *spec, int NK, int W, fgk_stream_t s, double *gK,
                        long *gN, double *gKmu, long *gNmu) {
    fgk_binspec g = *spec;
    g.geometry = 1;
    g.do_pl = 1;
    size_t nb = fgk_bins_doubles(&g), nc = fgk_bins_counts(&g);
    int64_t *db = NULL;
    int64_t *dc = NULL;
    TRY(fgk_malloc((void **)&db, 2 * nb * sizeof(int64_t)));
    if (fgk_malloc((void **)&dc, nc * sizeof(int64_t))) { fgk_free(db); return -1; }
    int64_t *hb = (int64_t *)malloc(2 * nb * sizeof(int64_t));
    long *hc = (long *)malloc(nc * sizeof(long));
    int rc = (!hb || !hc) ? -1 : 0;
    if (!rc) rc = fgk_fill_zero(db, 2 * nb * sizeof(int64_t), s);
    if (!rc) rc = fgk_fill_zero(dc, nc * sizeof(int64_t), s);
    if (!rc) rc = fgk_bin_geometry(&g, db, dc, s);
    if (!rc) rc = fgk_d2h(hb, db, 2 * nb * sizeof(int64_t), s);
    if (!rc) rc = fgk_d2h(hc, dc, nc * sizeof(int64_t), s);
    if (!rc) rc = fgk_stream_sync(s);
    if (!rc) {
        /* K sums: fixed point with exponent FGK_BIN_K_EXP */
        const int64_t *Ks = hb + 2 * (size_t)NK * (4 + W);
        const int64_t *Kmu = Ks + 2 * (size_t)NK;
        for (int b = 0; b < NK; ++b) {
            gK[b] = fgk_bin_fixed_value(Ks + 2 * b, FGK_BIN_K_EXP);
            gN[b] = hc[b];
            for (int m = 0; m < W; ++m) {
                gKmu[(size_t)b * W + m] =
                    fgk_bin_fixed_value(Kmu + 2 * ((size_t)b * W + m), FGK_BIN_K_EXP);
                gNmu[(size_t)b * W + m] = hc[NK + (size_t)b * W + m];
            }
        }
        long tot = 0;
        for (int b = 0; b < NK; ++b) tot += gN[b];
        if (tot != (long)spec->Ng * spec->Ng * spec->Ng - 1) {
            fgk_set_error("geometry: sum of mode counts %ld != Ng^3-1", tot);
            rc = -3;
        }
    } else if (!hb || !hc) {
        fgk_set_error("out of host memory");
    }
    free(hb);
    free(hc);
    fgk_free(db);
    fgk_free(dc);
    return rc;
}

static int compute_geometry(fftgal_power_t *p) {
    return bin_geometry(&p->spec, p->NK, p->W, p->s, p->gK, p->gN, p->gKmu, p->gNmu);
}

/* n fixed-point power sums (pairs, fgk.h) of a catalog with shot noise
 * Pshot, the scale the bin kernel used, to FP64 */
static void fixed_sums(const int64_t *hl, size_t n, double Pshot, double *out) {
    const int e = fgk_bin_fixed_exponent(Pshot);
    for (size_t i = 0; i < n; ++i) out[i] = fgk_bin_fixed_value(hl + 2 * i, e);
}

/* normalised Pl and Pmu of one catalog from its bin sums hb ([NK][4] then
 * [NK][W]) and the geometry: src/power.c:70-74 (no guard: empty bins give
 * NaN like the reference) and src/power.c:139-151 */
static void normalise_bins(int NK, int W, int Nmu, double KF, const double *hb, const double *gK,
                           const long *gN, const double *gKmu, const long *gNmu, double *K,
                           long *N, double *Pl, double *Kmu, long *Nmu_count, double *Pmu) {
    const double *Psum = hb;                   /* [NK][4] */
    const double *Pw = hb + (size_t)NK * 4;    /* [NK][W] */
    for (int b = 0; b < NK; ++b) {
        long n = gN[b];
        if (K) K[b] = gK[b] * (KF / n);
        if (N) N[b] = n;
        if (Pl)
            for (int l = 0; l < 4; ++l) Pl[(size_t)b * 4 + l] = Psum[(size_t)b * 4 + l] * ((4 * l + 1.0) / n);
        for (int m = 0; m < Nmu; ++m) {
            double ks = gKmu[(size_t)b * W + m], ps = Pw[(size_t)b * W + m];
            long c = gNmu[(size_t)b * W + m];
            if (m == Nmu - 1) {
                ks += gKmu[(size_t)b * W + Nmu];
                ps += Pw[(size_t)b * W + Nmu];
                c += gNmu[(size_t)b * W + Nmu];
            }
            if (c) {
                ks *= KF / c;
                ps /= c;
            }
            if (Kmu) Kmu[(size_t)b * Nmu + m] = ks;
            if (Nmu_count) Nmu_count[(size_t)b * Nmu + m] = c;
            if (Pmu) Pmu[(size_t)b * Nmu + m] = ps;
        }
    }
}

int fftgal_power_destroy(fftgal_power_t *p) {
    if (!p) return 0;
    if (p->s) fgk_stream_sync(p->s);
    fgk_graph_destroy(p->graph[0]);
    fgk_graph_destroy(p->graph[1]);
    fgk_fft_plan_destroy(p->plan);
    fgk_paint_plan_destroy(p->paint);
    fgk_free(p->gA);   /* gB, fB share these allocations */
    fgk_free(p->gA2);
    fgk_free(p->fA);
    fgk_free(p->d_phase);
    fgk_free(p->d_winv);
    fgk_free(p->d_bins);
    fgk_host_free(p->h_bins);
    free(p->h_sums);
    free(p->gK);
    free(p->gKmu);
    free(p->gN);
    free(p->gNmu);
    if (p->s) fgk_stream_destroy(p->s);
    free(p);
    return 0;
}

/* a second int64 grid set (catalogs alternate between the sets, each paint
 * clears the other set): unless FFTGAL_GRID_SETS=1, when it leaves at least
 * 4 GiB of device memory */
static int want_second_set(size_t bytes) {
    const char *e = getenv("FFTGAL_GRID_SETS");
    if (e && atoi(e) == 1) return 0;
    return fgk_mem_free() > bytes + ((size_t)4 << 30);
}

int fftgal_power_create(fftgal_power_t **out, const fftgal_power_desc *desc) {
    *out = NULL;
    if (!desc) { fgk_set_error("NULL descriptor"); return -2; }
    const int Ng = desc->Ng;
    if (!fgk_fft_supported(Ng)) {
        fgk_set_error("Ng=%d unsupported: need even Ng with prime factors <= 13", Ng);
        return -2;
    }
    if (!(desc->L > 0) || !(desc->dK > 0) || desc->Nmu < 0 || desc->Np_max < 1) {
        fgk_set_error("invalid descriptor (need L > 0, dK > 0, Nmu >= 0, Np_max >= 1)");
        return -2;
    }
    TRY(fgk_init());
    fftgal_power_t *p = (fftgal_power_t *)calloc(1, sizeof(fftgal_power_t));
    if (!p) { fgk_set_error("out of host memory"); return -1; }
    p->d = *desc;
    p->d.wisdom = NULL;
    p->use_graph = 1;
    p->los[0] = desc->los[0];
    p->los[1] = desc->los[1];
    p->los[2] = desc->los[2];
    {   /* hat() of src/gal.h:19-26 */
        double a = sqrt(p->los[0] * p->los[0] + p->los[1] * p->los[1] + p->los[2] * p->los[2]);
        if (a > 1e-7) { p->los[0] /= a; p->los[1] /= a; p->los[2] /= a; }
    }
    p->KF = 2 * M_PI / desc->L;
    p->NK = fftgal_nk(Ng, desc->L, desc->dK);
    p->W = desc->Nmu > 0 ? desc->Nmu + 1 : 0;
    int rc = 0;
#define STEP(call) do { if (!rc) rc = (call); } while (0)
    fftgal_tune_lookup(desc->wisdom, Ng, desc->interlace, &p->tune);
    if (desc->paint_tile > 0) p->tune.paint_tile = desc->paint_tile;
    STEP(fgk_stream_create(&p->s));
    STEP(fgk_fft_plan_create(&p->plan, Ng, 4));
    if (!rc) STEP(fgk_fft_set_tuning(p->plan, p->tune.fft_cols, p->tune.fft_rows));
    if (!rc && p->tune.fft_generic) STEP(fgk_fft_force_generic(p->plan, 1));
    if (!rc && fgk_paint_nstages() != NPAINT) {
        fgk_set_error("internal: paint stage count mismatch");
        rc = -3;
    }
    STEP(fgk_paint_plan_create(&p->paint, Ng, p->tune.paint_tile, desc->Np_max, desc->interlace));
    const size_t n64 = (size_t)Ng * Ng * Ng * sizeof(int64_t);
    const size_t nf = (size_t)Ng * Ng * Ng * sizeof(float);   /* packed half spectrum */
    const int ngrids = desc->interlace ? 2 : 1;
    STEP(fgk_malloc((void **)&p->gA, ngrids * n64));
    STEP(fgk_malloc((void **)&p->fA, ngrids * nf));
    if (!rc && desc->interlace) {
        p->gB = p->gA + (size_t)Ng * Ng * Ng;
        p->fB = p->fA + (size_t)Ng * Ng * Ng;
    }
    STEP(fgk_malloc((void **)&p->d_phase, (size_t)3 * Ng * 2 * sizeof(double)));
    STEP(fgk_malloc((void **)&p->d_winv, (size_t)Ng * sizeof(double)));
    if (!rc) {
        p->spec.Ng = Ng;
        p->spec.NK = p->NK;
        p->spec.Nmu = desc->Nmu;
        p->spec.dKinv = p->KF / desc->dK;
        p->spec.los[0] = p->los[0];
        p->spec.los[1] = p->los[1];
        p->spec.los[2] = p->los[2];
        p->spec.do_pl = 1;
        p->spec.geometry = 0;
        p->spec.bk = p->tune.bin_bk;
        p->nbins = fgk_bins_doubles(&p->spec);
    }
    STEP(fgk_malloc((void **)&p->d_bins, 2 * p->nbins * sizeof(int64_t)));
    STEP(fgk_host_alloc((void **)&p->h_bins, 2 * p->nbins * sizeof(int64_t)));
    if (!rc && !(p->h_sums = (double *)malloc(p->nbins * sizeof(double)))) {
        fgk_set_error("out of host memory");
        rc = -1;
    }
    if (!rc) {
        p->gK = (double *)calloc((size_t)p->NK, sizeof(double));
        p->gN = (long *)calloc((size_t)p->NK, sizeof(long));
        p->gKmu = (double *)calloc((size_t)p->NK * (p->W ? p->W : 1), sizeof(double));
        p->gNmu = (long *)calloc((size_t)p->NK * (p->W ? p->W : 1), sizeof(long));
        if (!p->gK || !p->gN || !p->gKmu || !p->gNmu) {
            fgk_set_error("out of host memory");
            rc = -1;
        }
    }
    if (!rc) {
        double *tab = (double *)malloc((size_t)3 * Ng * 2 * sizeof(double));
        double *winv = (double *)malloc((size_t)Ng * sizeof(double));
        if (!tab || !winv) {
            fgk_set_error("out of host memory");
            rc = -1;
        } else {
            /* dual grid offset 1/2 per axis (src/fft.c:260), phase on */
            double off[3] = {remainder(0.5, Ng), remainder(0.5, Ng), remainder(0.5, Ng)};
            fftgal_phase_table(Ng, off, 1, tab);
            fftgal_winv_table(Ng, winv);
            rc = fgk_h2d(p->d_phase, tab, (size_t)3 * Ng * 2 * sizeof(double), p->s);
            if (!rc) rc = fgk_h2d(p->d_winv, winv, (size_t)Ng * sizeof(double), p->s);
            if (!rc) rc = fgk_stream_sync(p->s);
        }
        free(tab);
        free(winv);
    }
    STEP(fgk_fill_zero(p->gA, n64, p->s));
    if (desc->interlace) STEP(fgk_fill_zero(p->gB, n64, p->s));
    if (!rc && want_second_set((size_t)ngrids * n64)) {
        if (fgk_malloc((void **)&p->gA2, ngrids * n64) == 0)
            STEP(fgk_fill_zero(p->gA2, ngrids * n64, p->s));
        else
            p->gA2 = NULL;
    }
    STEP(compute_geometry(p));
    STEP(fgk_stream_sync(p->s));
#undef STEP
    if (rc) {
        fftgal_power_destroy(p);
        return rc;
    }
    *out = p;
    return 0;
}

int fftgal_power_NK(const fftgal_power_t *p) { return p->NK; }

/* pipeline stages, in order (fftgal_power_profile reports one time each):
 * zero_bins, the paint stages of fgk_paint_stage_name(), the FFT passes of
 * both grids, the fused bin kernel and the copy of the sums to the host */
static const char *const kTailStages[] = {"fft_rows", "fft_cols_j", "fft_cols_i", "fused_bin",
                                          "copy_bins"};
#define NSTAGES (1 + NPAINT + (int)(sizeof kTailStages / sizeof kTailStages[0]))

/* normalisations of one catalog's fused bin: norm = 1/(nbar H^3) of the
 * paint (src/fft.c:125-127) times H^3 of x2k; Vinv and Pshot of P
 * (src/power.c:15-17, no randoms) */
static void catalog_norms(int Ng, double L, long Np, double V, double *norm, double *Vinv,
                          double *Pshot) {
    const double H3 = (L / Ng) * (L / Ng) * (L / Ng);
    const double nbar = Np / V;
    const double Npcinv = 1 / nbar / H3;
    *norm = Npcinv * H3;
    *Vinv = 1 / V;
    *Pshot = V / Np;
}

/* enqueue the whole pipeline; ev (NSTAGES+1 events or NULL) brackets the
 * stages for profiling */
static int enqueue_work(fftgal_power_t *p, const float *x, const float *y, const float *z,
                        long Np, double V, fgk_event_t *ev) {
    const int Ng = p->d.Ng;
    const double Hinv = Ng / p->d.L;
    double norm, Vinv, Pshot;
    catalog_norms(Ng, p->d.L, Np, V, &norm, &Vinv, &Pshot);
    /* the FFT reads the int64 grids unscaled; the particle mass of one unit
     * goes into norm */
    norm *= fgk_paint_fixed_scale();
    p->spec.fs_n1 = fgk_fft_fourstep(p->plan);   /* the tuner may switch FFT kernels */
    const size_t Ng3 = (size_t)Ng * Ng * Ng;
    const int ngrids = p->d.interlace ? 2 : 1;
    /* grid set of this catalog; with two sets the paint clears the other */
    int64_t *gA = p->gA;
    if (p->gA2) {
        int64_t *other = p->parity ? p->gA : p->gA2;
        if (p->parity) gA = p->gA2;
        fgk_paint_set_zero(p->paint, other, (size_t)ngrids * Ng3 * sizeof(int64_t));
    } else {
        fgk_paint_set_zero(p->paint, NULL, 0);
    }
    int64_t *gB = p->d.interlace ? gA + Ng3 : NULL;
    int st = 0;
#define MARK() do { if (ev) TRY(fgk_event_record(ev[st], p->s)); ++st; } while (0)
    MARK();
    TRY(fgk_fill_zero(p->d_bins, 2 * p->nbins * sizeof(int64_t), p->s));
    for (int stage = 0; stage < NPAINT; ++stage) {
        MARK();
        TRY(fgk_paint_stage(p->paint, stage, x, y, z, Np, Hinv, gA, gB, p->s));
    }
    MARK();
    TRY(fgk_fft_rows_f32_from_fixed(p->plan, p->fA, gA, 1.0, p->gA2 ? 0 : 1, ngrids, p->s));
    MARK();
    TRY(fgk_fft_cols_f32(p->plan, p->fA, 1, -1, ngrids, p->s));
    MARK();
    TRY(fgk_fft_cols_f32(p->plan, p->fA, 0, -1, ngrids, p->s));
    MARK();
    TRY(fgk_fused_bin_f32(p->fA, p->d.interlace ? p->fB : NULL, &p->spec, p->d_phase, p->d_winv,
                          norm, Vinv, Pshot, (double)Np, p->d_bins, p->s));
    MARK();
    TRY(fgk_d2h(p->h_bins, p->d_bins, 2 * p->nbins * sizeof(int64_t), p->s));
    MARK();
#undef MARK
    return 0;
}

int fftgal_power_stages(void) { return NSTAGES; }

const char *fftgal_power_stage_name(int i) {
    if (i == 0) return "zero_bins";
    if (i >= 1 && i <= NPAINT) return fgk_paint_stage_name(i - 1);
    if (i > NPAINT && i < NSTAGES) return kTailStages[i - 1 - NPAINT];
    return NULL;
}

/* zero every int64 grid of the plan (after a paint_only) */
static int zero_grids(fftgal_power_t *p) {
    const size_t n = (size_t)p->d.Ng * p->d.Ng * p->d.Ng * sizeof(int64_t) * (p->d.interlace ? 2 : 1);
    TRY(fgk_fill_zero(p->gA, n, p->s));
    if (p->gA2) TRY(fgk_fill_zero(p->gA2, n, p->s));
    p->grids_dirty = 0;
    return 0;
}

int fftgal_power_profile(fftgal_power_t *p, const float *x, const float *y, const float *z,
                         long Np, double V, float *stage_ms) {
    if (Np < 1 || Np > p->d.Np_max || !(V > 0)) {
        fgk_set_error("profile: invalid Np or V");
        return -2;
    }
    fgk_event_t ev[NSTAGES + 1] = {0};
    int rc = 0;
    for (int i = 0; i <= NSTAGES && !rc; ++i) rc = fgk_event_create(&ev[i]);
    if (!rc && p->grids_dirty) rc = zero_grids(p);
    if (!rc) rc = enqueue_work(p, x, y, z, Np, V, ev);
    if (!rc && p->gA2) p->parity ^= 1;
    if (!rc) rc = fgk_stream_sync(p->s);
    for (int i = 0; i < NSTAGES && !rc; ++i) rc = fgk_event_elapsed_ms(&stage_ms[i], ev[i], ev[i + 1]);
    for (int i = 0; i <= NSTAGES; ++i) fgk_event_destroy(ev[i]);
    if (!rc) {
        p->Np = Np;
        p->V = V;
    }
    return rc;
}

int fftgal_power_enqueue(fftgal_power_t *p, const float *x, const float *y, const float *z,
                         long Np, double V) {
    if (Np < 1 || Np > p->d.Np_max) {
        fgk_set_error("enqueue: Np=%ld outside [1, Np_max=%ld]", Np, p->d.Np_max);
        return -2;
    }
    if (!(V > 0)) {
        fgk_set_error("enqueue: V must be > 0");
        return -2;
    }
    if (p->grids_dirty) TRY(zero_grids(p));
    p->Np = Np;
    p->V = V;
    /* grid set q for this catalog (enqueue_work reads p->parity); the next
     * catalog takes the other set */
    const int q = p->parity;
    if (!p->use_graph) {
        TRY(enqueue_work(p, x, y, z, Np, V, NULL));
        if (p->gA2) p->parity = q ^ 1;
        return 0;
    }
    if (!p->graph[q] || p->kx[q] != x || p->ky[q] != y || p->kz[q] != z || p->kNp[q] != Np ||
        p->kV[q] != V) {
        if (p->graph[q]) {
            TRY(fgk_stream_sync(p->s));
            fgk_graph_destroy(p->graph[q]);
            p->graph[q] = NULL;
        }
        TRY(fgk_graph_begin(p->s));
        int rc = enqueue_work(p, x, y, z, Np, V, NULL);
        fgk_graph_t g = NULL;
        int rc2 = fgk_graph_end(p->s, &g);
        if (rc) {
            if (g) fgk_graph_destroy(g);
            return rc;
        }
        TRY(rc2);
        p->graph[q] = g;
        p->kx[q] = x;
        p->ky[q] = y;
        p->kz[q] = z;
        p->kNp[q] = Np;
        p->kV[q] = V;
    }
    TRY(fgk_graph_launch(p->graph[q], p->s));
    if (p->gA2) p->parity = q ^ 1;
    return 0;
}

int fftgal_power_sync(fftgal_power_t *p) { return fgk_stream_sync(p->s); }

int fftgal_power_results(fftgal_power_t *p, double *K, long *N, double *Pl, double *Kmu,
                         long *Nmu_count, double *Pmu) {
    double norm, Vinv, Pshot;
    catalog_norms(p->d.Ng, p->d.L, p->Np, p->V, &norm, &Vinv, &Pshot);
    fixed_sums(p->h_bins, p->nbins, Pshot, p->h_sums);
    normalise_bins(p->NK, p->W, p->d.Nmu, p->KF, p->h_sums, p->gK, p->gN, p->gKmu, p->gNmu, K, N,
                   Pl, Kmu, Nmu_count, Pmu);
    return 0;
}

int fftgal_power_paint_only(fftgal_power_t *p, const float *x, const float *y, const float *z,
                            long Np, void **gridA, void **gridB) {
    const size_t n64 = (size_t)p->d.Ng * p->d.Ng * p->d.Ng * sizeof(int64_t);
    TRY(fgk_fill_zero(p->gA, n64, p->s));
    if (p->d.interlace) TRY(fgk_fill_zero(p->gB, n64, p->s));
    p->grids_dirty = 1;          /* every grid is zeroed before the next enqueue */
    fgk_paint_set_zero(p->paint, NULL, 0);
    TRY(fgk_paint_fixed(p->paint, x, y, z, Np, p->d.Ng / p->d.L, p->gA, p->gB, p->s));
    TRY(fgk_stream_sync(p->s));
    if (gridA) *gridA = p->gA;
    if (gridB) *gridB = p->gB;
    return 0;
}

int fftgal_power_grid_sets(const fftgal_power_t *p) { return p->gA2 ? 2 : 1; }

int fftgal_power_fixed_bits(long n) { return fgk_paint_tile_bits(n); }

double fftgal_power_fixed_scale(void) { return fgk_paint_fixed_scale(); }

int fftgal_power_spectra(fftgal_power_t *p, void **A, void **B) {
    if (A) *A = p->fA;
    if (B) *B = p->fB;
    return 0;
}

int fftgal_power_tuning(const fftgal_power_t *p, int *paint_tile, int *fft_cols, int *fft_rows,
                        int *fft_generic, int *bin_bk) {
    if (paint_tile) *paint_tile = fgk_paint_plan_tile(p->paint);
    if (fft_cols) *fft_cols = p->tune.fft_cols;
    if (fft_rows) *fft_rows = p->tune.fft_rows;
    if (fft_generic) *fft_generic = p->tune.fft_generic || !fgk_fft_f32_specialised(p->d.Ng);
    if (bin_bk) *bin_bk = p->spec.bk ? p->spec.bk : 16;
    return 0;
}

/* ---- autotuner ----
 * Times the candidates of each knob on the caller's catalog with eager
 * profiled runs (HIP events per stage), one knob at a time in pipeline
 * order: paint tile side (paint stages), FFT kernels (FFT stages), bin
 * block (fused bin stage); the fastest of `reps` runs of each candidate
 * counts.  The winner is stored in the wisdom file (tune.h) for
 * (Ng, interlace), where fftgal_power_create and fftgal_batch_create find
 * it. */

static float stage_sum(const float *ms, int a, int b) {
    float t = 0.0f;
    for (int i = a; i < b; ++i) t += ms[i];
    return t;
}

/* fastest of reps profiled runs: paint, fft and bin stage times */
static int tune_measure(fftgal_power_t *p, const float *x, const float *y, const float *z,
                        long Np, double V, int reps, float best[3]) {
    float ms[NSTAGES];
    best[0] = best[1] = best[2] = 1e30f;
    for (int r = 0; r <= reps; ++r) {   /* the first run warms up */
        TRY(fftgal_power_profile(p, x, y, z, Np, V, ms));
        if (r == 0) continue;
        const float t[3] = {stage_sum(ms, 1, 1 + NPAINT), stage_sum(ms, 1 + NPAINT, 4 + NPAINT),
                            ms[4 + NPAINT]};
        for (int k = 0; k < 3; ++k)
            if (t[k] < best[k]) best[k] = t[k];
    }
    return 0;
}

/* destroy the captured graphs (both grid sets): they hold the kernels and
 * launch shapes of the tuning in use */
static void drop_graphs(fftgal_power_t *p) {
    for (int q = 0; q < 2; ++q) {
        fgk_graph_destroy(p->graph[q]);
        p->graph[q] = NULL;
    }
}

/* switch a plan's FFT kernels: family (generic or specialised) and the
 * generic kernels' columns / rows per workgroup */
static int tune_fft(fftgal_power_t *p, int generic, int cols, int rows) {
    TRY(fgk_stream_sync(p->s));
    drop_graphs(p);
    TRY(fgk_fft_set_tuning(p->plan, cols, rows));
    TRY(fgk_fft_force_generic(p->plan, generic));
    p->tune.fft_generic = generic;
    p->tune.fft_cols = cols;
    p->tune.fft_rows = rows;
    return 0;
}

int fftgal_power_autotune(const fftgal_power_desc *desc, const float *x, const float *y,
                          const float *z, long Np, double V, int reps, int *paint_tile,
                          int *fft_generic, int *bin_bk) {
    if (!desc) { fgk_set_error("autotune: NULL descriptor"); return -2; }
    if (reps < 1) reps = 1;
    const int Ng = desc->Ng;
    fftgal_tuning best;
    fftgal_tune_lookup(NULL, Ng, desc->interlace, &best);
    /* 1. paint tile side */
    static const int kTiles[] = {16, 12, 8, 4};
    float tbest = 1e30f;
    for (int i = 0; i < 4; ++i) {
        const int T = kTiles[i];
        if (T > Ng) continue;
        fftgal_power_desc d = *desc;
        d.wisdom = NULL;
        d.paint_tile = T;
        fftgal_power_t *p = NULL;
        if (fftgal_power_create(&p, &d)) continue;   /* side not plannable here (tiles, memory) */
        float m[3];
        int rc = tune_measure(p, x, y, z, Np, V, reps, m);
        fftgal_power_destroy(p);
        if (rc) return rc;
        if (m[0] < tbest) {
            tbest = m[0];
            best.paint_tile = T;
        }
    }
    if (!(tbest < 1e30f)) {
        char why[256];
        snprintf(why, sizeof why, "%s", fgk_last_error());
        fgk_set_error("autotune: no paint tile side could be planned for Ng=%d (%s)", Ng, why);
        return -2;
    }
    /* 2. FFT kernels and 3. bin block, on a plan with the chosen tile */
    fftgal_power_desc d = *desc;
    d.wisdom = NULL;
    d.paint_tile = best.paint_tile;
    fftgal_power_t *p = NULL;
    TRY(fftgal_power_create(&p, &d));
    int rc = 0;
    float fbest = 1e30f;
    /* candidates: specialised kernels (where they exist), generic kernels with
     * the default launch shape, with half the columns, with half the rows */
    const int spec = fgk_fft_f32_specialised(Ng);
    int cand[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, -1, 0}, {1, 0, -1}};
    for (int c = spec ? 0 : 1; c < 4 && !rc; ++c) {
        int cols = cand[c][1], rows = cand[c][2];
        if (cols < 0) {   /* half the default columns per workgroup */
            int C = 16;
            while (C > 1 && (size_t)2 * Ng * C * 8 > 64 * 1024) C >>= 1;
            if (C < 2) continue;
            cols = C / 2;
        }
        if (rows < 0) {   /* half the default rows per workgroup */
            int P = 1;
            while (P * (Ng / 2) < 1024 && (size_t)2 * P * 2 * (Ng / 2) * 8 <= 64 * 1024) P *= 2;
            if (P < 2) continue;
            rows = P / 2;
        }
        rc = tune_fft(p, cand[c][0], cols, rows);
        float m[3];
        if (!rc) rc = tune_measure(p, x, y, z, Np, V, reps, m);
        if (!rc && m[1] < fbest) {
            fbest = m[1];
            best.fft_generic = cand[c][0];
            best.fft_cols = cols;
            best.fft_rows = rows;
        }
    }
    if (!rc) rc = tune_fft(p, best.fft_generic, best.fft_cols, best.fft_rows);
    float bbest = 1e30f;
    static const int kBk[] = {16, 32, 64, 128};
    for (int i = 0; i < 4 && !rc; ++i) {
        if (kBk[i] != 64 && kBk[i] > Ng / 2) continue;
        p->spec.bk = kBk[i];
        float m[3];
        rc = tune_measure(p, x, y, z, Np, V, reps, m);
        if (!rc && m[2] < bbest) {
            bbest = m[2];
            best.bin_bk = kBk[i];
        }
    }
    fftgal_power_destroy(p);
    if (rc) return rc;
    if (fftgal_tune_store(desc->wisdom, Ng, desc->interlace, &best)) {
        fgk_set_error("autotune: cannot write the wisdom file %s", desc->wisdom);
        return -1;
    }
    if (paint_tile) *paint_tile = best.paint_tile;
    if (fft_generic) *fft_generic = best.fft_generic;
    if (bin_bk) *bin_bk = best.bin_bk;
    return 0;
}

int fftgal_power_use_graph(fftgal_power_t *p, int on) {
    p->use_graph = on ? 1 : 0;
    return 0;
}

int fftgal_power_fft_generic(fftgal_power_t *p, int on) {
    TRY(fgk_stream_sync(p->s));
    drop_graphs(p);   /* the captured graphs hold the other kernels */
    TRY(fgk_fft_force_generic(p->plan, on));
    return fgk_fft_f32_specialised(p->d.Ng) && !on ? 0 : 1;
}

void *fftgal_power_stream(fftgal_power_t *p) { return p->s; }

int fftgal_memcpy_d2h(void *dst, const void *src, unsigned long bytes) {
    TRY(fgk_init());
    TRY(fgk_d2h(dst, src, bytes, NULL));
    return fgk_stream_sync(NULL);
}

/* ------------------------------------------------------------------------
 * Batches of catalogs (fftgal_batch_*): many catalogs, each on its own grid
 * (side, box = catalog box / fold, interlacing, k binning).
 *
 * Device memory is a bounded pool.  The catalogs are cut into rounds of
 * about equal size, in input order, as few as the pool allows; the rounds
 * run one after the other through the same memory:
 *   pool = [grid set 0 | grid set 1 | spectra of a round | sort words of a round]
 * With more than one round, a second stream sorts and paints the rounds
 * after the first while the batch's stream transforms and bins the one
 * before (batch_work_overlap); the sort words have a region of their own
 * so that a round's sort never writes the spectra the previous round's FFT
 * passes and bin still use.
 * A grid set holds the int64 grids of one round (the largest round's size);
 * the rounds alternate between the two sets, and each round's paint clears
 * the part of the other set that the previous round's row FFT has read.  With one set
 * (FFTGAL_GRID_SETS=1) the row FFT clears the grids it reads instead.  Both
 * keep the invariant that a round finds its grid set zero.
 * Within a round, one launch per stage:
 *   sort and paint  -- all its catalogs together (fgk_paintbatch; the paint
 *                      kernel dispatches on each catalog's tile side),
 *   FFT passes      -- the grids of every catalog of one side and
 *                      interlacing lie one after the other (a group); the
 *                      multi-size row and column kernels cover every group
 *                      (fgk_fftmulti), each workgroup finding its group,
 *   fused bin       -- one launch pair with a table of every catalog (grid
 *                      side, tables and first workgroup per catalog).
 * (FFTGAL_BATCH_MULTI=0, or a side without a multi-size body: one launch
 * per tile side, per group and per group respectively.)
 * Then one copy of every catalog's bin sums to pinned host memory.  The
 * whole sequence is one hipGraph; catalog pointers, Np and V live in device
 * tables, so a new set of catalogs updates the tables and replays the same
 * graph.  The pool's bound is FFTGAL_BATCH_BYTES (bytes, default 34 GiB).
 * ---------------------------------------------------------------------- */

#define BATCH_POOL_DEFAULT ((size_t)34 << 30)

typedef struct batch_side {      /* per grid side: FFT plan and k-space tables */
    int Ng;
    fgk_fftplan *plan;
    double *d_phase, *d_winv;
    fftgal_tuning tune;
} batch_side;

typedef struct batch_group {     /* the grids of one side and interlacing in a round */
    int Ng, interlace, n;        /* n catalogs */
    int side;                    /* its batch_side */
    int *jobs;                   /* batch catalog indices, in grid order */
    size_t goff;                 /* offset of its grids in a grid set, int64 words */
    int64_t *g, *g2;             /* n * G int64 grids in set 0 and set 1 (NULL: one set) */
    float *f;                    /* n * G packed spectra */
    fgk_binspec spec;            /* Ng, Nmu, los, do_pl (NK per catalog) */
    int NK_max;
    fgk_binjob *hjob, *djob;
} batch_group;

typedef struct batch_round {
    int njobs, ngroups;
    int *jobs;                   /* batch catalog indices, in paint order */
    batch_group *grp;
    fgk_paintbatch *paint;
    size_t gwords;               /* int64 words of its grids */
    /* one launch per stage over all its groups (NULL: a launch per group) */
    fgk_fftmulti *fm;            /* FFT row and column passes */
    fgk_binjob *hjall, *djall;   /* fused bin: every catalog, blocks laid out in order */
    fgk_binspec bspec;           /* its shared Nmu, los and bin block shape */
    int NK_max;
    unsigned bblocks, bpblocks;
} batch_round;

typedef struct batch_job {
    fftgal_job_desc d;
    int NK, round, group, slot, local;   /* local: index within its round */
    double KF;
    size_t boff;                 /* offset of its bin sums in d_bins / h_bins */
    double *gK, *gKmu;
    long *gN, *gNmu;
    long Np;
    double V;
} batch_job;

struct fftgal_batch {
    int njobs, nrounds, nsides, Nmu, W;
    double los[3];
    batch_job *job;
    batch_round *rnd;
    batch_side *side;
    fgk_stream_t s;
    fgk_stream_t s2;             /* a round's sort beside the previous round's FFT and bin */
    fgk_event_t *evp, *evr;      /* per round: its paint done, its row pass done */
    char *ws_region;             /* the sort words of every round */
    int64_t *d_bins, *h_bins;    /* fixed-point pairs (fgk.h), device and pinned */
    double *h_sums;              /* the values in FP64 */
    size_t nbins;                /* values */
    char *pool;
    size_t pool_bytes;
    int64_t *gset[2];            /* grid sets in the pool (gset[1] NULL: one set) */
    size_t gwords;               /* int64 words of a grid set */
    int tables_ok;
    int use_graph;
    fgk_graph_t graph[2];        /* one per grid set of the first round */
    int parity;                  /* set of the next enqueue's first round */
    const float **px, **py, **pz;   /* per-round argument arrays of fgk_paint_batch_set */
    long *pNp;
};

int fftgal_batch_destroy(fftgal_batch_t *b) {
    if (!b) return 0;
    if (b->s) fgk_stream_sync(b->s);
    if (b->s2) fgk_stream_sync(b->s2);   /* joined into s by every enqueue that completes */
    fgk_graph_destroy(b->graph[0]);
    fgk_graph_destroy(b->graph[1]);
    for (int r = 0; r < b->nrounds && b->rnd; ++r) {
        batch_round *R = &b->rnd[r];
        fgk_paint_batch_destroy(R->paint);
        for (int q = 0; q < R->ngroups && R->grp; ++q) {
            fgk_free(R->grp[q].djob);
            free(R->grp[q].hjob);
            free(R->grp[q].jobs);
        }
        fgk_fft_multi_destroy(R->fm);
        fgk_free(R->djall);
        free(R->hjall);
        free(R->grp);
        free(R->jobs);
    }
    for (int i = 0; i < b->nsides && b->side; ++i) {
        fgk_fft_plan_destroy(b->side[i].plan);
        fgk_free(b->side[i].d_phase);
        fgk_free(b->side[i].d_winv);
    }
    for (int j = 0; j < b->njobs && b->job; ++j) {
        free(b->job[j].gK);
        free(b->job[j].gKmu);
        free(b->job[j].gN);
        free(b->job[j].gNmu);
    }
    free(b->rnd);
    free(b->side);
    free(b->job);
    free(b->px);
    free(b->py);
    free(b->pz);
    free(b->pNp);
    fgk_free(b->pool);
    fgk_free(b->d_bins);
    fgk_host_free(b->h_bins);
    free(b->h_sums);
    for (int r = 0; r < b->nrounds; ++r) {
        if (b->evp && b->evp[r]) fgk_event_destroy(b->evp[r]);
        if (b->evr && b->evr[r]) fgk_event_destroy(b->evr[r]);
    }
    free(b->evp);
    free(b->evr);
    if (b->s2) fgk_stream_destroy(b->s2);
    if (b->s) fgk_stream_destroy(b->s);
    free(b);
    return 0;
}

static int side_of(fftgal_batch_t *b, int Ng) {
    for (int i = 0; i < b->nsides; ++i)
        if (b->side[i].Ng == Ng) return i;
    b->side[b->nsides].Ng = Ng;
    return b->nsides++;
}

static int side_setup(fftgal_batch_t *b, batch_side *S, int interlace, const char *wisdom) {
    const int Ng = S->Ng;
    fftgal_tune_lookup(wisdom, Ng, interlace, &S->tune);
    TRY(fgk_fft_plan_create(&S->plan, Ng, 4));
    TRY(fgk_fft_set_tuning(S->plan, S->tune.fft_cols, S->tune.fft_rows));
    if (S->tune.fft_generic) TRY(fgk_fft_force_generic(S->plan, 1));
    TRY(fgk_malloc((void **)&S->d_phase, (size_t)3 * Ng * 2 * sizeof(double)));
    TRY(fgk_malloc((void **)&S->d_winv, (size_t)Ng * sizeof(double)));
    double *tab = (double *)malloc((size_t)3 * Ng * 2 * sizeof(double));
    double *winv = (double *)malloc((size_t)Ng * sizeof(double));
    int rc = (!tab || !winv) ? -1 : 0;
    if (rc) fgk_set_error("out of host memory");
    if (!rc) {
        /* dual grid offset 1/2 per axis (src/fft.c:260), phase on */
        double off[3] = {remainder(0.5, Ng), remainder(0.5, Ng), remainder(0.5, Ng)};
        fftgal_phase_table(Ng, off, 1, tab);
        fftgal_winv_table(Ng, winv);
        rc = fgk_h2d(S->d_phase, tab, (size_t)3 * Ng * 2 * sizeof(double), b->s);
        if (!rc) rc = fgk_h2d(S->d_winv, winv, (size_t)Ng * sizeof(double), b->s);
        if (!rc) rc = fgk_stream_sync(b->s);
    }
    free(tab);
    free(winv);
    return rc;
}

/* device bytes of catalog j's grids (one set), spectra and sort words */
/* a catalog's int64 grids, FP32 spectra and sort words (the sort words of a
 * round of catalogs take at most the sum of theirs) */
static void job_bytes(const fftgal_job_desc *d, size_t *gbytes, size_t *fbytes, size_t *wbytes) {
    const size_t cells = (size_t)d->Ng * d->Ng * d->Ng, NG = d->interlace ? 2 : 1;
    *gbytes = (NG * cells * sizeof(int64_t) + 255) & ~(size_t)255;
    *fbytes = (NG * cells * sizeof(float) + 255) & ~(size_t)255;
    *wbytes = (fgk_paint_batch_ws_bytes(d->Np_max) + 255) & ~(size_t)255;
}

static size_t pool_size(size_t gmax, size_t fmax, size_t wmax, int sets) {
    return sets * gmax + fmax + wmax;
}

/* one launch per stage over the groups of a round (FFT passes and the
 * fused bin), unless FFTGAL_BATCH_MULTI=0 */
static int want_multi(void) {
    const char *e = getenv("FFTGAL_BATCH_MULTI");
    return !(e && atoi(e) == 0);
}

/* the one-launch tables of round R (groups of several grid sides): the
 * FFT passes (when every side has a multi-size body) and the fused bin over
 * every catalog, which shares Nmu, los and one bin block shape (the largest
 * grid's; the exact fixed-point sums do not depend on it) */
static int round_multi(fftgal_batch_t *b, batch_round *R) {
    fgk_fftplan **plans = (fgk_fftplan **)calloc((size_t)R->ngroups, sizeof(fgk_fftplan *));
    float **f = (float **)calloc((size_t)R->ngroups, sizeof(float *));
    int64_t **s0 = (int64_t **)calloc((size_t)R->ngroups, sizeof(int64_t *));
    int64_t **s1 = (int64_t **)calloc((size_t)R->ngroups, sizeof(int64_t *));
    int *nb = (int *)calloc((size_t)R->ngroups, sizeof(int));
    int rc = (!plans || !f || !s0 || !s1 || !nb) ? -1 : 0;
    if (rc) fgk_set_error("out of host memory");
    int big = 0;
    for (int q = 0; q < R->ngroups && !rc; ++q) {
        const batch_group *G = &R->grp[q];
        plans[q] = b->side[G->side].plan;
        f[q] = G->f;
        s0[q] = G->g;
        s1[q] = G->g2;
        nb[q] = G->n * (G->interlace ? 2 : 1);
        if (G->Ng > R->grp[big].Ng) big = q;
    }
    if (!rc) rc = fgk_fft_multi_create(&R->fm, R->ngroups, plans, f, s0, b->gset[1] ? s1 : NULL, nb);
    free(plans);
    free(f);
    free(s0);
    free(s1);
    free(nb);
    if (rc) return rc;
    for (int q = 0; q < R->ngroups; ++q)   /* the bin kernel reads natural plane order */
        if (fgk_fft_fourstep(b->side[R->grp[q].side].plan)) return 0;
    R->bspec = R->grp[big].spec;
    R->hjall = (fgk_binjob *)calloc((size_t)R->njobs, sizeof(fgk_binjob));
    if (!R->hjall) { fgk_set_error("out of host memory"); return -1; }
    unsigned long blk = 0, pblk = 0;
    int k = 0;
    for (int q = 0; q < R->ngroups; ++q) {
        const batch_group *G = &R->grp[q];
        const batch_side *S = &b->side[G->side];
        unsigned nbk = 0, npk = 0;
        fgk_bin_multi_blocks(&R->bspec, G->Ng, &nbk, &npk);
        for (int i = 0; i < G->n; ++i, ++k) {
            fgk_binjob *h = &R->hjall[k];
            *h = G->hjob[i];
            h->Ng = G->Ng;
            h->fs_n1 = 0;
            h->phase = S->d_phase;
            h->winv = S->d_winv;
            h->blk0 = (unsigned)blk;
            h->pblk0 = (unsigned)pblk;
            blk += nbk;
            pblk += npk;
        }
        if (G->NK_max > R->NK_max) R->NK_max = G->NK_max;
    }
    if (blk >= (1ul << 31) || pblk >= (1ul << 31)) {
        free(R->hjall);
        R->hjall = NULL;
        return 0;
    }
    R->bblocks = (unsigned)blk;
    R->bpblocks = (unsigned)pblk;
    TRY(fgk_malloc((void **)&R->djall, (size_t)R->njobs * sizeof(fgk_binjob)));
    return 0;
}

/* the grids, spectra, bin tables and paint batch of round r, in the pool */
static int round_setup(fftgal_batch_t *b, int r) {
    batch_round *R = &b->rnd[r];
    /* groups by grid side and interlacing, in the round's catalog order */
    R->grp = (batch_group *)calloc((size_t)R->njobs, sizeof(batch_group));
    if (!R->grp) { fgk_set_error("out of host memory"); return -1; }
    for (int k = 0; k < R->njobs; ++k) {
        batch_job *J = &b->job[R->jobs[k]];
        int q = 0;
        while (q < R->ngroups && !(R->grp[q].Ng == J->d.Ng && R->grp[q].interlace == J->d.interlace))
            ++q;
        if (q == R->ngroups) {
            R->grp[q].Ng = J->d.Ng;
            R->grp[q].interlace = J->d.interlace;
            R->grp[q].side = side_of(b, J->d.Ng);
            ++R->ngroups;
        }
        J->round = r;
        J->group = q;
        J->slot = R->grp[q].n++;
        J->local = k;
    }
    char *fw = b->pool + (size_t)(b->gset[1] ? 2 : 1) * b->gwords * sizeof(int64_t);
    for (int q = 0; q < R->ngroups; ++q) {
        batch_group *G = &R->grp[q];
        const int NG = G->interlace ? 2 : 1;
        const size_t cells = (size_t)G->Ng * G->Ng * G->Ng;
        G->jobs = (int *)calloc((size_t)G->n, sizeof(int));
        G->hjob = (fgk_binjob *)calloc((size_t)G->n, sizeof(fgk_binjob));
        if (!G->jobs || !G->hjob) { fgk_set_error("out of host memory"); return -1; }
        G->goff = R->gwords;
        R->gwords += (((size_t)G->n * NG * cells * sizeof(int64_t) + 255) & ~(size_t)255) / 8;
        G->g = b->gset[0] + G->goff;
        G->g2 = b->gset[1] ? b->gset[1] + G->goff : NULL;
        G->f = (float *)fw;
        fw += ((size_t)G->n * NG * cells * sizeof(float) + 255) & ~(size_t)255;
        TRY(fgk_malloc((void **)&G->djob, (size_t)G->n * sizeof(fgk_binjob)));
    }
    for (int k = 0; k < R->njobs; ++k) {
        const batch_job *J = &b->job[R->jobs[k]];
        R->grp[J->group].jobs[J->slot] = R->jobs[k];
    }
    for (int q = 0; q < R->ngroups; ++q) {
        batch_group *G = &R->grp[q];
        const batch_side *S = &b->side[G->side];
        const int NG = G->interlace ? 2 : 1;
        const size_t cells = (size_t)G->Ng * G->Ng * G->Ng;
        G->spec.Ng = G->Ng;
        G->spec.Nmu = b->Nmu;
        G->spec.los[0] = b->los[0];
        G->spec.los[1] = b->los[1];
        G->spec.los[2] = b->los[2];
        G->spec.do_pl = 1;
        G->spec.geometry = 0;
        G->spec.bk = S->tune.bin_bk;
        for (int k = 0; k < G->n; ++k) {
            batch_job *J = &b->job[G->jobs[k]];
            fgk_binjob *h = &G->hjob[k];
            h->A = G->f + (size_t)k * NG * cells;
            h->B = G->interlace ? h->A + cells : NULL;
            h->sums = b->d_bins + 2 * J->boff;
            h->NK = J->NK;
            h->dKinv = J->KF / J->d.dK;
            if (J->NK > G->NK_max) G->NK_max = J->NK;
        }
    }
    /* the paint batch writes each catalog's grids in its group; its sort
     * words follow the round's spectra */
    fgk_paintjob_desc *pd = (fgk_paintjob_desc *)calloc((size_t)R->njobs, sizeof *pd);
    if (!pd) { fgk_set_error("out of host memory"); return -1; }
    long Np_round = 0;
    for (int k = 0; k < R->njobs; ++k) {
        const batch_job *J = &b->job[R->jobs[k]];
        const batch_group *G = &R->grp[J->group];
        const batch_side *S = &b->side[G->side];
        const size_t cells = (size_t)G->Ng * G->Ng * G->Ng, NG = G->interlace ? 2 : 1;
        pd[k].Ng = J->d.Ng;
        pd[k].tile = J->d.paint_tile > 0 ? J->d.paint_tile : S->tune.paint_tile;
        pd[k].Hinv = J->d.Ng / J->d.L;
        pd[k].Np_max = J->d.Np_max;
        pd[k].interlace = J->d.interlace;
        pd[k].gridA = G->g + (size_t)J->slot * NG * cells;
        pd[k].gridB = G->interlace ? pd[k].gridA + cells : NULL;
        if (G->g2) {
            pd[k].gridA2 = G->g2 + (size_t)J->slot * NG * cells;
            pd[k].gridB2 = G->interlace ? pd[k].gridA2 + cells : NULL;
        }
        Np_round += J->d.Np_max;
    }
    const size_t ws = fgk_paint_batch_ws_bytes(Np_round);
    int rc = 0;
    if (fw > b->ws_region || b->ws_region + ws > b->pool + b->pool_bytes) {
        fgk_set_error("batch: round %d needs more than the pool", r);
        rc = -1;
    }
    if (!rc) rc = fgk_paint_batch_create_ws(&R->paint, R->njobs, pd, b->ws_region, ws);
    free(pd);
    if (!rc) fgk_paint_batch_per_side(R->paint, !want_multi());
    if (!rc && R->ngroups > 1 && want_multi()) rc = round_multi(b, R);
    return rc;
}

/* rounds of the catalogs in `order`: a catalog starts a new round when
 * adding it would take the pool past `budget` or the round past `target`
 * bytes (sets grid sets, spectra and sort words).  Fills b->rnd when b is
 * given; returns the number of rounds (-1: out of host memory) and the
 * largest round's grid, spectra and sort-word bytes. */
static int form_rounds(fftgal_batch_t *b, const fftgal_job_desc *jobs, const int *order,
                       int njobs, size_t budget, int sets, size_t target, size_t *gmax_out,
                       size_t *fmax_out, size_t *wmax_out) {
    size_t gmax = 0, fmax = 0, wmax = 0, gcur = 0, fcur = 0, wcur = 0;
    int nr = 0;
    for (int i = 0; i < njobs; ++i) {
        size_t gj, fj, wj;
        job_bytes(&jobs[order[i]], &gj, &fj, &wj);
        const size_t g2 = gcur + gj, f2 = fcur + fj, w2 = wcur + wj;
        const int fits = pool_size(g2 > gmax ? g2 : gmax, f2 > fmax ? f2 : fmax,
                                   w2 > wmax ? w2 : wmax, sets) <= budget &&
                         (size_t)sets * g2 + f2 + w2 <= target;
        if (nr == 0 || !fits) {
            if (b) {
                b->rnd[nr].jobs = (int *)calloc((size_t)njobs, sizeof(int));
                if (!b->rnd[nr].jobs) return -1;
            }
            ++nr;
            gcur = fcur = wcur = 0;
        }
        if (b) b->rnd[nr - 1].jobs[b->rnd[nr - 1].njobs++] = order[i];
        gcur += gj;
        fcur += fj;
        wcur += wj;
        if (gcur > gmax) gmax = gcur;
        if (fcur > fmax) fmax = fcur;
        if (wcur > wmax) wmax = wcur;
    }
    if (b) b->nrounds = nr;
    *gmax_out = gmax;
    *fmax_out = fmax;
    *wmax_out = wmax;
    return nr;
}

int fftgal_batch_create(fftgal_batch_t **out, int njobs, const fftgal_job_desc *jobs,
                        const double los[3], int Nmu, const char *wisdom) {
    *out = NULL;
    if (njobs < 1 || njobs > 65535 || !jobs || Nmu < 0) {
        fgk_set_error("batch: need 1..65535 catalogs and Nmu >= 0");
        return -2;
    }
    for (int j = 0; j < njobs; ++j) {
        const fftgal_job_desc *d = &jobs[j];
        if (!fgk_fft_supported(d->Ng)) {
            fgk_set_error("batch: catalog %d, Ng=%d unsupported (even, prime factors <= 13)", j,
                          d->Ng);
            return -2;
        }
        if (!(d->L > 0) || !(d->dK > 0) || d->Np_max < 1) {
            fgk_set_error("batch: catalog %d needs L > 0, dK > 0, Np_max >= 1", j);
            return -2;
        }
    }
    TRY(fgk_init());
    fftgal_batch_t *b = (fftgal_batch_t *)calloc(1, sizeof(fftgal_batch_t));
    if (!b) { fgk_set_error("out of host memory"); return -1; }
    b->njobs = njobs;
    b->Nmu = Nmu;
    b->W = Nmu > 0 ? Nmu + 1 : 0;
    b->use_graph = 1;
    b->los[0] = los ? los[0] : 0.0;
    b->los[1] = los ? los[1] : 0.0;
    b->los[2] = los ? los[2] : 1.0;
    {   /* hat() of src/gal.h:19-26 */
        double a = sqrt(b->los[0] * b->los[0] + b->los[1] * b->los[1] + b->los[2] * b->los[2]);
        if (a > 1e-7) { b->los[0] /= a; b->los[1] /= a; b->los[2] /= a; }
    }
    b->job = (batch_job *)calloc((size_t)njobs, sizeof(batch_job));
    b->rnd = (batch_round *)calloc((size_t)njobs, sizeof(batch_round));
    b->side = (batch_side *)calloc((size_t)njobs, sizeof(batch_side));
    b->px = (const float **)calloc((size_t)njobs, sizeof(float *));
    b->py = (const float **)calloc((size_t)njobs, sizeof(float *));
    b->pz = (const float **)calloc((size_t)njobs, sizeof(float *));
    b->pNp = (long *)calloc((size_t)njobs, sizeof(long));
    int *order = (int *)calloc((size_t)njobs, sizeof(int));
    int rc = (!b->job || !b->rnd || !b->side || !b->px || !b->py || !b->pz || !b->pNp || !order)
                 ? -1 : 0;
    if (rc) fgk_set_error("out of host memory");
#define STEP(call) do { if (!rc) rc = (call); } while (0)
    STEP(fgk_stream_create(&b->s));
    /* catalogs and bin layout */
    for (int j = 0; j < njobs && !rc; ++j) {
        batch_job *J = &b->job[j];
        J->d = jobs[j];
        J->d.interlace = jobs[j].interlace ? 1 : 0;
        J->KF = 2 * M_PI / J->d.L;
        J->NK = fftgal_nk(J->d.Ng, J->d.L, J->d.dK);
        fgk_binspec bs = {0};
        bs.NK = J->NK;
        bs.Nmu = Nmu;
        bs.do_pl = 1;
        J->boff = b->nbins;
        b->nbins += fgk_bins_doubles(&bs);
        const size_t W1 = b->W ? (size_t)b->W : 1;
        J->gK = (double *)calloc((size_t)J->NK, sizeof(double));
        J->gN = (long *)calloc((size_t)J->NK, sizeof(long));
        J->gKmu = (double *)calloc((size_t)J->NK * W1, sizeof(double));
        J->gNmu = (long *)calloc((size_t)J->NK * W1, sizeof(long));
        if (!J->gK || !J->gN || !J->gKmu || !J->gNmu) {
            fgk_set_error("out of host memory");
            rc = -1;
        }
        order[j] = j;
        side_of(b, J->d.Ng);
    }
    /* FFT plans and k-space tables per grid side (tuning lines are kept per
     * interlacing: the side's first catalog's) */
    for (int i = 0; i < b->nsides && !rc; ++i) {
        int inter = 1;
        for (int j = 0; j < njobs; ++j)
            if (b->job[j].d.Ng == b->side[i].Ng) { inter = b->job[j].d.interlace; break; }
        rc = side_setup(b, &b->side[i], inter, wisdom);
    }
    /* rounds: catalogs in input order (each round then holds a mix of grid
     * sides, as the caller's list does), each round as large as the pool
     * bound allows with the grid sets of the largest */
    const char *e = getenv("FFTGAL_BATCH_BYTES");
    const size_t budget = e && atof(e) > 0 ? (size_t)atof(e) : BATCH_POOL_DEFAULT;
    const char *es = getenv("FFTGAL_GRID_SETS");
    const int sets = (es && atoi(es) == 1) ? 1 : 2;
    /* the fewest rounds the bound allows (greedy), then rounds of about
     * equal size: a round's tail (its last, longest workgroups) repeats per
     * round, so one small last round costs as much as a full one */
    size_t total = 0;
    for (int i = 0; i < njobs; ++i) {
        size_t gj, fj, wj;
        job_bytes(&jobs[i], &gj, &fj, &wj);
        total += (size_t)sets * gj + fj + wj;
    }
    int nr = 0;
    size_t gmax = 0, fmax = 0, wmax = 0;
    if (!rc)
        nr = form_rounds(NULL, jobs, order, njobs, budget, sets, (size_t)-1, &gmax, &fmax, &wmax);
    if (!rc && nr > 1) {
        size_t g2 = 0, f2 = 0, w2 = 0;
        const size_t target = (total + nr - 1) / nr;
        const int n2 = form_rounds(NULL, jobs, order, njobs, budget, sets, target, &g2, &f2, &w2);
        nr = form_rounds(b, jobs, order, njobs, budget, sets, n2 == nr ? target : (size_t)-1,
                         &gmax, &fmax, &wmax);
    } else if (!rc) {
        nr = form_rounds(b, jobs, order, njobs, budget, sets, (size_t)-1, &gmax, &fmax, &wmax);
    }
    if (!rc && nr < 0) { fgk_set_error("out of host memory"); rc = -1; }
    free(order);
    STEP(fgk_malloc((void **)&b->d_bins, 2 * b->nbins * sizeof(int64_t)));
    STEP(fgk_host_alloc((void **)&b->h_bins, 2 * b->nbins * sizeof(int64_t)));
    if (!rc && !(b->h_sums = (double *)malloc(b->nbins * sizeof(double)))) {
        fgk_set_error("out of host memory");
        rc = -1;
    }
    /* the pool: grid sets (zeroed once; every round leaves its set zero),
     * the spectra of one round, the sort words of one round */
    b->gwords = gmax / 8;
    b->pool_bytes = pool_size(gmax, fmax, wmax, sets);
    STEP(fgk_malloc((void **)&b->pool, b->pool_bytes));
    if (!rc) {
        b->gset[0] = (int64_t *)b->pool;
        b->gset[1] = sets == 2 ? (int64_t *)(b->pool + gmax) : NULL;
        b->ws_region = b->pool + (size_t)sets * gmax + fmax;
    }
    /* more than one round: each round's sort runs on s2 beside the previous
     * round's FFT passes and bin */
    if (!rc && b->nrounds > 1) {
        STEP(fgk_stream_create(&b->s2));
        b->evp = (fgk_event_t *)calloc((size_t)b->nrounds, sizeof(fgk_event_t));
        b->evr = (fgk_event_t *)calloc((size_t)b->nrounds, sizeof(fgk_event_t));
        if (!rc && (!b->evp || !b->evr)) {
            fgk_set_error("out of host memory");
            rc = -1;
        }
        for (int r = 0; r < b->nrounds && !rc; ++r) {
            STEP(fgk_event_create_dep(&b->evp[r]));
            STEP(fgk_event_create_dep(&b->evr[r]));
        }
    }
    STEP(fgk_fill_zero(b->pool, (size_t)sets * gmax, b->s));
    for (int r = 0; r < b->nrounds && !rc; ++r) rc = round_setup(b, r);
    /* K sums and mode counts per catalog */
    for (int j = 0; j < njobs && !rc; ++j) {
        batch_job *J = &b->job[j];
        fgk_binspec g = b->rnd[J->round].grp[J->group].spec;
        g.NK = J->NK;
        g.dKinv = J->KF / J->d.dK;
        rc = bin_geometry(&g, J->NK, b->W, b->s, J->gK, J->gN, J->gKmu, J->gNmu);
    }
    STEP(fgk_stream_sync(b->s));
#undef STEP
    if (rc) {
        fftgal_batch_destroy(b);
        return rc;
    }
    *out = b;
    return 0;
}

int fftgal_batch_NK(const fftgal_batch_t *b, int job) {
    if (job < 0 || job >= b->njobs) return -2;
    return b->job[job].NK;
}

/* the device tables for this set of catalogs (waits for queued work when
 * they change) */
static int batch_tables(fftgal_batch_t *b, const float *const *x, const float *const *y,
                        const float *const *z, const long *Np, const double *V) {
    for (int j = 0; j < b->njobs; ++j) {
        if (Np[j] < 1 || Np[j] > b->job[j].d.Np_max || !(V[j] > 0)) {
            fgk_set_error("batch: catalog %d has Np=%ld (Np_max=%ld) or V=%g invalid", j, Np[j],
                          b->job[j].d.Np_max, V[j]);
            return -2;
        }
    }
    for (int r = 0; r < b->nrounds; ++r) {
        const batch_round *R = &b->rnd[r];
        for (int k = 0; k < R->njobs; ++k) {
            const int j = R->jobs[k];
            b->px[k] = x[j];
            b->py[k] = y[j];
            b->pz[k] = z[j];
            b->pNp[k] = Np[j];
        }
        TRY(fgk_paint_batch_set(R->paint, b->px, b->py, b->pz, b->pNp, b->s));
    }
    int same = b->tables_ok;
    for (int j = 0; j < b->njobs; ++j)
        if (b->job[j].Np != Np[j] || b->job[j].V != V[j]) same = 0;
    if (same) return 0;
    TRY(fgk_stream_sync(b->s));
    for (int r = 0; r < b->nrounds; ++r) {
        for (int q = 0; q < b->rnd[r].ngroups; ++q) {
            batch_group *G = &b->rnd[r].grp[q];
            for (int k = 0; k < G->n; ++k) {
                batch_job *J = &b->job[G->jobs[k]];
                fgk_binjob *h = &G->hjob[k];
                const long n = Np[G->jobs[k]];
                const double v = V[G->jobs[k]];
                catalog_norms(G->Ng, J->d.L, n, v, &h->norm, &h->Vinv, &h->Pshot);
                /* the FFT reads the int64 grids unscaled */
                h->norm *= fgk_paint_fixed_scale();
                h->npart = (double)n;
                J->Np = n;
                J->V = v;
            }
            TRY(fgk_h2d(G->djob, G->hjob, (size_t)G->n * sizeof(fgk_binjob), b->s));
        }
        batch_round *R = &b->rnd[r];
        if (R->djall) {   /* the one-launch bin table: the same catalogs in group order */
            int k = 0;
            for (int q = 0; q < R->ngroups; ++q)
                for (int i = 0; i < R->grp[q].n; ++i, ++k) {
                    const fgk_binjob *h = &R->grp[q].hjob[i];
                    R->hjall[k].norm = h->norm;
                    R->hjall[k].Vinv = h->Vinv;
                    R->hjall[k].Pshot = h->Pshot;
                    R->hjall[k].npart = h->npart;
                }
            TRY(fgk_h2d(R->djall, R->hjall, (size_t)R->njobs * sizeof(fgk_binjob), b->s));
        }
    }
    TRY(fgk_stream_sync(b->s));
    b->tables_ok = 1;
    return 0;
}

/* the batch's rounds, each through the stages in the order and under the
 * names of fftgal_power_stage_name; ev (NSTAGES+1 events per round, or
 * NULL) brackets them.  set0: grid set of the first round. */
/* host-side set-up of round r before its paint: spectra order of its FFT
 * plans, and with two grid sets the set its paint writes and the part of
 * the other set it clears (as far as the previous round, or for the first
 * round the last round of the previous enqueue, wrote it; on the first
 * enqueue that part is zero already) */
static int round_select(fftgal_batch_t *b, int r, int set) {
    batch_round *R = &b->rnd[r];
    for (int q = 0; q < R->ngroups; ++q)
        R->grp[q].spec.fs_n1 = fgk_fft_fourstep(b->side[R->grp[q].side].plan);
    if (b->gset[1]) {
        const batch_round *P = &b->rnd[(r + b->nrounds - 1) % b->nrounds];
        TRY(fgk_paint_batch_select(R->paint, set, b->gset[set ^ 1], P->gwords * sizeof(int64_t)));
    }
    return 0;
}

/* the row pass, both column passes and the fused bin of round r on s (and
 * after the last round the copy of the bin sums); with ev, each stage
 * boundary records the next of the profile's events (*st counts them);
 * after_rows, when given, is recorded once the row pass is issued */
static int round_fft_bin(fftgal_batch_t *b, int r, int set, fgk_stream_t s, fgk_event_t *ev,
                         int *st, fgk_event_t after_rows) {
    batch_round *R = &b->rnd[r];
#define MARK() do { if (ev) TRY(fgk_event_record(ev[*st], s)); ++*st; } while (0)
    if (R->fm) {
        TRY(fgk_fft_multi_rows(R->fm, set, b->gset[1] ? 0 : 1, s));
    } else {
        for (int q = 0; q < R->ngroups; ++q) {
            batch_group *G = &R->grp[q];
            TRY(fgk_fft_rows_f32_from_fixed(b->side[G->side].plan, G->f, set ? G->g2 : G->g, 1.0,
                                            b->gset[1] ? 0 : 1, G->n * (G->interlace ? 2 : 1),
                                            s));
        }
    }
    if (after_rows) TRY(fgk_event_record(after_rows, s));
    for (int axis = 1; axis >= 0; --axis) {
        MARK();
        if (R->fm) {
            TRY(fgk_fft_multi_cols(R->fm, axis, -1, s));
            continue;
        }
        for (int q = 0; q < R->ngroups; ++q) {
            batch_group *G = &R->grp[q];
            TRY(fgk_fft_cols_f32(b->side[G->side].plan, G->f, axis, -1,
                                 G->n * (G->interlace ? 2 : 1), s));
        }
    }
    MARK();
    if (R->djall) {
        TRY(fgk_fused_bin_f32_multi(&R->bspec, R->djall, R->njobs, R->NK_max, R->bblocks,
                                    R->bpblocks, s));
    } else {
        for (int q = 0; q < R->ngroups; ++q) {
            batch_group *G = &R->grp[q];
            const batch_side *S = &b->side[G->side];
            TRY(fgk_fused_bin_f32_batch(&G->spec, G->djob, G->n, G->NK_max, S->d_phase,
                                        S->d_winv, s));
        }
    }
    MARK();
    if (r == b->nrounds - 1)
        TRY(fgk_d2h(b->h_bins, b->d_bins, 2 * b->nbins * sizeof(int64_t), s));
    MARK();
#undef MARK
    return 0;
}

/* every round on the batch's stream, with stage events when ev is given */
static int batch_work_serial(fftgal_batch_t *b, int set0, fgk_event_t *ev) {
    int st = 0;
#define MARK() do { if (ev) TRY(fgk_event_record(ev[st], b->s)); ++st; } while (0)
    for (int r = 0; r < b->nrounds; ++r) {
        const int set = b->gset[1] ? (set0 + r) & 1 : 0;
        TRY(round_select(b, r, set));
        MARK();
        if (r == 0) TRY(fgk_fill_zero(b->d_bins, 2 * b->nbins * sizeof(int64_t), b->s));
        for (int stage = 0; stage < NPAINT; ++stage) {
            MARK();
            TRY(fgk_paint_batch_stage(b->rnd[r].paint, stage, b->s));
        }
        MARK();
        TRY(round_fft_bin(b, r, set, b->s, ev, &st, NULL));
    }
#undef MARK
    return 0;
}

/* Rounds on two streams: s2 sorts and paints rounds 1.., s transforms and
 * bins every round (and sorts and paints round 0).  Round r+1's sort starts
 * once round r's paint has read the sort words (the same stream), and its
 * paint once round r's row pass has read (and with one grid set cleared)
 * the grids that paint clears or writes, so both run beside round r's
 * column passes and bin; round r+1's row pass waits for its paint, and
 * after round r's bin has read the spectra (the same stream). */
static int batch_work_overlap(fftgal_batch_t *b, int set0) {
    int st = 0;
    const int n = b->nrounds;
    for (int r = 0; r < n; ++r) {
        const int set = b->gset[1] ? (set0 + r) & 1 : 0;
        if (r == 0) {
            TRY(round_select(b, 0, set));
            TRY(fgk_fill_zero(b->d_bins, 2 * b->nbins * sizeof(int64_t), b->s));
            for (int stage = 0; stage < NPAINT; ++stage)
                TRY(fgk_paint_batch_stage(b->rnd[0].paint, stage, b->s));
            TRY(fgk_event_record(b->evp[0], b->s));
            TRY(fgk_stream_wait_event(b->s2, b->evp[0]));
            for (int stage = 0; stage < NPAINT - 1; ++stage)
                TRY(fgk_paint_batch_stage(b->rnd[1].paint, stage, b->s2));
        } else {
            TRY(fgk_stream_wait_event(b->s, b->evp[r]));
        }
        const int next = r + 1 < n;
        TRY(round_fft_bin(b, r, set, b->s, NULL, &st, next ? b->evr[r] : NULL));
        if (!next) break;
        /* issued after round r's work on s, so that s2's wait finds the
         * record of round r's row pass */
        const int set1 = b->gset[1] ? (set0 + r + 1) & 1 : 0;
        TRY(round_select(b, r + 1, set1));
        TRY(fgk_stream_wait_event(b->s2, b->evr[r]));
        TRY(fgk_paint_batch_stage(b->rnd[r + 1].paint, NPAINT - 1, b->s2));
        TRY(fgk_event_record(b->evp[r + 1], b->s2));
        if (r + 2 < n)
            for (int stage = 0; stage < NPAINT - 1; ++stage)
                TRY(fgk_paint_batch_stage(b->rnd[r + 2].paint, stage, b->s2));
    }
    return 0;
}

static int batch_work(fftgal_batch_t *b, int set0, fgk_event_t *ev) {
    if (!ev && b->s2 && !getenv("FFTGAL_BATCH_SERIAL")) return batch_work_overlap(b, set0);
    return batch_work_serial(b, set0, ev);
}

/* grid set of the round after this enqueue's last */
static int next_parity(const fftgal_batch_t *b, int set0) {
    return b->gset[1] ? (set0 + b->nrounds) & 1 : 0;
}

int fftgal_batch_enqueue(fftgal_batch_t *b, const float *const *x, const float *const *y,
                         const float *const *z, const long *Np, const double *V) {
    TRY(batch_tables(b, x, y, z, Np, V));
    const int q = b->parity;
    if (!b->use_graph) {
        TRY(batch_work(b, q, NULL));
        b->parity = next_parity(b, q);
        return 0;
    }
    if (!b->graph[q]) {
        TRY(fgk_graph_begin(b->s));
        int rc = batch_work(b, q, NULL);
        fgk_graph_t g = NULL;
        int rc2 = fgk_graph_end(b->s, &g);
        if (rc) {
            if (g) fgk_graph_destroy(g);
            return rc;
        }
        TRY(rc2);
        b->graph[q] = g;
    }
    TRY(fgk_graph_launch(b->graph[q], b->s));
    b->parity = next_parity(b, q);
    return 0;
}

int fftgal_batch_grid_sets(const fftgal_batch_t *b) { return b->gset[1] ? 2 : 1; }

int fftgal_batch_rounds(const fftgal_batch_t *b) { return b->nrounds; }

size_t fftgal_batch_pool_bytes(const fftgal_batch_t *b) { return b->pool_bytes; }

int fftgal_batch_sync(fftgal_batch_t *b) { return fgk_stream_sync(b->s); }

int fftgal_batch_profile(fftgal_batch_t *b, const float *const *x, const float *const *y,
                         const float *const *z, const long *Np, const double *V,
                         float *stage_ms) {
    TRY(batch_tables(b, x, y, z, Np, V));
    const int n = b->nrounds * (NSTAGES + 1);
    fgk_event_t *ev = (fgk_event_t *)calloc((size_t)n, sizeof(fgk_event_t));
    if (!ev) { fgk_set_error("out of host memory"); return -1; }
    int rc = 0;
    for (int i = 0; i < n && !rc; ++i) rc = fgk_event_create(&ev[i]);
    const int q = b->parity;
    if (!rc) rc = batch_work(b, q, ev);
    if (!rc) b->parity = next_parity(b, q);
    if (!rc) rc = fgk_stream_sync(b->s);
    for (int i = 0; i < NSTAGES; ++i) stage_ms[i] = 0.0f;
    for (int r = 0; r < b->nrounds && !rc; ++r)
        for (int i = 0; i < NSTAGES && !rc; ++i) {
            float ms = 0.0f;
            rc = fgk_event_elapsed_ms(&ms, ev[r * (NSTAGES + 1) + i], ev[r * (NSTAGES + 1) + i + 1]);
            stage_ms[i] += ms;
        }
    for (int i = 0; i < n; ++i) fgk_event_destroy(ev[i]);
    free(ev);
    return rc;
}

int fftgal_batch_results(fftgal_batch_t *b, int job, double *K, long *N, double *Pl, double *Kmu,
                         long *Nmu_count, double *Pmu) {
    if (job < 0 || job >= b->njobs) {
        fgk_set_error("batch: no catalog %d", job);
        return -2;
    }
    const batch_job *J = &b->job[job];
    double norm, Vinv, Pshot;
    catalog_norms(J->d.Ng, J->d.L, J->Np, J->V, &norm, &Vinv, &Pshot);
    fgk_binspec bs = {0};
    bs.NK = J->NK;
    bs.Nmu = b->Nmu;
    bs.do_pl = 1;
    fixed_sums(b->h_bins + 2 * J->boff, fgk_bins_doubles(&bs), Pshot, b->h_sums + J->boff);
    normalise_bins(J->NK, b->W, b->Nmu, J->KF, b->h_sums + J->boff, J->gK, J->gN, J->gKmu,
                   J->gNmu, K, N, Pl, Kmu, Nmu_count, Pmu);
    return 0;
}

int fftgal_batch_use_graph(fftgal_batch_t *b, int on) {
    b->use_graph = on ? 1 : 0;
    return 0;
}

int fftgal_batch_tile(const fftgal_batch_t *b, int job) {
    if (job < 0 || job >= b->njobs) return -2;
    const batch_job *J = &b->job[job];
    return fgk_paint_batch_tile(b->rnd[J->round].paint, J->local);
}

void *fftgal_batch_stream(fftgal_batch_t *b) { return b->s; }
