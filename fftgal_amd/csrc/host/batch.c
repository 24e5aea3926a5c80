/* Mixed-precision power-spectrum pipeline (include/fftgal_batch.h).
 *
 * One fftgal_power_t holds everything a catalog needs on the device:
 *   int64 fixed-point grids  gA (plain) and gB (dual)     Ng^3 each
 *   FP32 spectra             fA, fB  (packed half spectrum) Ng^3 floats each
 *   paint plan (sort buffers), FFT plan (FP32 twiddles from FP64),
 *   offset-phase and deconvolution tables (FP64, built as src/fft.c:204-245),
 *   bin sums (FP64) and the geometry of the bins (K sums, mode counts),
 *   computed once at creation because they depend only on Ng, L, dK, los
 *   and Nmu.
 * An enqueue runs: zero bins -> fixed-point paint -> per grid (rows r2c
 * from int64, cols j, cols i) -> fused bin kernel -> copy sums to pinned
 * host memory.  The sequence is captured in a hipGraph keyed by the
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
    float *fA, *fB;
    double *d_phase, *d_winv;
    fgk_binspec spec;          /* per-run sums */
    double *d_bins;
    size_t nbins;
    double *h_bins;            /* pinned */
    /* geometry */
    double *gK, *gKmu;
    long *gN, *gNmu;
    /* last enqueue */
    long Np;
    double V;
    int grids_dirty;
    /* graph */
    int use_graph;
    fgk_graph_t graph;
    const float *kx, *ky, *kz;
    long kNp;
    double kV;
    /* tuning */
    fftgal_tuning tune;
};

const char *fftgal_last_error(void) { return fgk_last_error(); }

#define TRY(call)                 \
    do {                          \
        int rc_ = (call);         \
        if (rc_) return rc_;      \
    } while (0)

static int compute_geometry(fftgal_power_t *p) {
    fgk_binspec g = p->spec;
    g.geometry = 1;
    g.do_pl = 1;
    size_t nb = fgk_bins_doubles(&g), nc = fgk_bins_counts(&g);
    double *db = NULL;
    int64_t *dc = NULL;
    TRY(fgk_malloc((void **)&db, nb * sizeof(double)));
    if (fgk_malloc((void **)&dc, nc * sizeof(int64_t))) { fgk_free(db); return -1; }
    double *hb = (double *)malloc(nb * sizeof(double));
    long *hc = (long *)malloc(nc * sizeof(long));
    int rc = (!hb || !hc) ? -1 : 0;
    if (!rc) rc = fgk_fill_zero(db, nb * sizeof(double), p->s);
    if (!rc) rc = fgk_fill_zero(dc, nc * sizeof(int64_t), p->s);
    if (!rc) rc = fgk_bin_geometry(&g, db, dc, p->s);
    if (!rc) rc = fgk_d2h(hb, db, nb * sizeof(double), p->s);
    if (!rc) rc = fgk_d2h(hc, dc, nc * sizeof(int64_t), p->s);
    if (!rc) rc = fgk_stream_sync(p->s);
    if (!rc) {
        const int NK = p->NK, W = p->W;
        const double *Ks = hb + (size_t)NK * (4 + W);
        const double *Kmu = Ks + NK;
        for (int b = 0; b < NK; ++b) {
            p->gK[b] = Ks[b];
            p->gN[b] = hc[b];
            for (int m = 0; m < W; ++m) {
                p->gKmu[(size_t)b * W + m] = Kmu[(size_t)b * W + m];
                p->gNmu[(size_t)b * W + m] = hc[NK + (size_t)b * W + m];
            }
        }
        long tot = 0;
        for (int b = 0; b < NK; ++b) tot += p->gN[b];
        if (tot != (long)p->d.Ng * p->d.Ng * p->d.Ng - 1) {
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

int fftgal_power_destroy(fftgal_power_t *p) {
    if (!p) return 0;
    if (p->s) fgk_stream_sync(p->s);
    fgk_graph_destroy(p->graph);
    fgk_fft_plan_destroy(p->plan);
    fgk_paint_plan_destroy(p->paint);
    fgk_free(p->gA);
    fgk_free(p->gB);
    fgk_free(p->fA);
    fgk_free(p->fB);
    fgk_free(p->d_phase);
    fgk_free(p->d_winv);
    fgk_free(p->d_bins);
    fgk_host_free(p->h_bins);
    free(p->gK);
    free(p->gKmu);
    free(p->gN);
    free(p->gNmu);
    if (p->s) fgk_stream_destroy(p->s);
    free(p);
    return 0;
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
    if (!rc && fgk_paint_nstages() != NPAINT) {
        fgk_set_error("internal: paint stage count mismatch");
        rc = -3;
    }
    STEP(fgk_paint_plan_create(&p->paint, Ng, p->tune.paint_tile, desc->Np_max, desc->interlace));
    const size_t n64 = (size_t)Ng * Ng * Ng * sizeof(int64_t);
    const size_t nf = (size_t)Ng * Ng * Ng * sizeof(float);   /* packed half spectrum */
    STEP(fgk_malloc((void **)&p->gA, n64));
    STEP(fgk_malloc((void **)&p->fA, nf));
    if (desc->interlace) {
        STEP(fgk_malloc((void **)&p->gB, n64));
        STEP(fgk_malloc((void **)&p->fB, nf));
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
        p->nbins = fgk_bins_doubles(&p->spec);
    }
    STEP(fgk_malloc((void **)&p->d_bins, p->nbins * sizeof(double)));
    STEP(fgk_host_alloc((void **)&p->h_bins, p->nbins * sizeof(double)));
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
static const char *const kTailStages[] = {"fft_rows_A", "fft_cols_j_A", "fft_cols_i_A",
                                          "fft_rows_B", "fft_cols_j_B", "fft_cols_i_B",
                                          "fused_bin",  "copy_bins"};
#define NSTAGES (1 + NPAINT + (int)(sizeof kTailStages / sizeof kTailStages[0]))

/* enqueue the whole pipeline; ev (NSTAGES+1 events or NULL) brackets the
 * stages for profiling */
static int enqueue_work(fftgal_power_t *p, const float *x, const float *y, const float *z,
                        long Np, double V, fgk_event_t *ev) {
    const int Ng = p->d.Ng;
    const double Hinv = Ng / p->d.L;
    const double H3 = (p->d.L / Ng) * (p->d.L / Ng) * (p->d.L / Ng);
    const double nbar = Np / V;
    const double Npcinv = 1 / nbar / H3;       /* src/fft.c:125-127 */
    const double norm = Npcinv * H3;           /* paint normalisation x H^3 of x2k */
    const double Vinv = 1 / V, Pshot = V / Np; /* src/power.c:15-17, no randoms */
    const double in_scale = fgk_paint_fixed_scale(p->paint);
    int st = 0;
#define MARK() do { if (ev) TRY(fgk_event_record(ev[st], p->s)); ++st; } while (0)
    MARK();
    TRY(fgk_fill_zero(p->d_bins, p->nbins * sizeof(double), p->s));
    for (int stage = 0; stage < NPAINT; ++stage) {
        MARK();
        TRY(fgk_paint_stage(p->paint, stage, x, y, z, Np, Hinv, p->gA, p->gB, p->s));
    }
    for (int g = 0; g < 2; ++g) {
        int64_t *src = g ? p->gB : p->gA;
        float *dst = g ? p->fB : p->fA;
        const int on = g == 0 || p->d.interlace;
        MARK();
        if (on) TRY(fgk_fft_rows_f32_from_fixed(p->plan, dst, src, in_scale, 1, p->s));
        MARK();
        if (on) TRY(fgk_fft_cols_f32(p->plan, dst, 1, -1, p->s));
        MARK();
        if (on) TRY(fgk_fft_cols_f32(p->plan, dst, 0, -1, p->s));
    }
    MARK();
    TRY(fgk_fused_bin_f32(p->fA, p->d.interlace ? p->fB : NULL, &p->spec, p->d_phase, p->d_winv,
                          norm, Vinv, Pshot, (double)Np, p->d_bins, p->s));
    MARK();
    TRY(fgk_d2h(p->h_bins, p->d_bins, p->nbins * sizeof(double), p->s));
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

int fftgal_power_profile(fftgal_power_t *p, const float *x, const float *y, const float *z,
                         long Np, double V, float *stage_ms) {
    if (Np < 1 || Np > p->d.Np_max || !(V > 0)) {
        fgk_set_error("profile: invalid Np or V");
        return -2;
    }
    fgk_event_t ev[NSTAGES + 1] = {0};
    int rc = 0;
    for (int i = 0; i <= NSTAGES && !rc; ++i) rc = fgk_event_create(&ev[i]);
    if (!rc && p->grids_dirty) {
        const size_t n64 = (size_t)p->d.Ng * p->d.Ng * p->d.Ng * sizeof(int64_t);
        rc = fgk_fill_zero(p->gA, n64, p->s);
        if (!rc && p->d.interlace) rc = fgk_fill_zero(p->gB, n64, p->s);
        p->grids_dirty = 0;
    }
    if (!rc) rc = enqueue_work(p, x, y, z, Np, V, ev);
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
    if (p->grids_dirty) {
        const size_t n64 = (size_t)p->d.Ng * p->d.Ng * p->d.Ng * sizeof(int64_t);
        TRY(fgk_fill_zero(p->gA, n64, p->s));
        if (p->d.interlace) TRY(fgk_fill_zero(p->gB, n64, p->s));
        p->grids_dirty = 0;
    }
    p->Np = Np;
    p->V = V;
    if (!p->use_graph) return enqueue_work(p, x, y, z, Np, V, NULL);
    if (!p->graph || p->kx != x || p->ky != y || p->kz != z || p->kNp != Np || p->kV != V) {
        if (p->graph) {
            TRY(fgk_stream_sync(p->s));
            fgk_graph_destroy(p->graph);
            p->graph = NULL;
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
        p->graph = g;
        p->kx = x;
        p->ky = y;
        p->kz = z;
        p->kNp = Np;
        p->kV = V;
    }
    return fgk_graph_launch(p->graph, p->s);
}

int fftgal_power_sync(fftgal_power_t *p) { return fgk_stream_sync(p->s); }

int fftgal_power_results(fftgal_power_t *p, double *K, long *N, double *Pl, double *Kmu,
                         long *Nmu_count, double *Pmu) {
    const int NK = p->NK, W = p->W, Nmu = p->d.Nmu;
    const double KF = p->KF;
    const double *Psum = p->h_bins;                  /* [NK][4] */
    const double *Pw = p->h_bins + (size_t)NK * 4;   /* [NK][W] */
    for (int b = 0; b < NK; ++b) {
        /* src/power.c:70-74 (no guard: empty bins give NaN like the reference) */
        long n = p->gN[b];
        if (K) K[b] = p->gK[b] * (KF / n);
        if (N) N[b] = n;
        if (Pl)
            for (int l = 0; l < 4; ++l) Pl[(size_t)b * 4 + l] = Psum[(size_t)b * 4 + l] * ((4 * l + 1.0) / n);
        /* src/power.c:139-151 */
        for (int m = 0; m < Nmu; ++m) {
            double ks = p->gKmu[(size_t)b * W + m], ps = Pw[(size_t)b * W + m];
            long c = p->gNmu[(size_t)b * W + m];
            if (m == Nmu - 1) {
                ks += p->gKmu[(size_t)b * W + Nmu];
                ps += Pw[(size_t)b * W + Nmu];
                c += p->gNmu[(size_t)b * W + Nmu];
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
    return 0;
}

int fftgal_power_paint_only(fftgal_power_t *p, const float *x, const float *y, const float *z,
                            long Np, void **gridA, void **gridB) {
    const size_t n64 = (size_t)p->d.Ng * p->d.Ng * p->d.Ng * sizeof(int64_t);
    TRY(fgk_fill_zero(p->gA, n64, p->s));
    if (p->d.interlace) TRY(fgk_fill_zero(p->gB, n64, p->s));
    p->grids_dirty = 1;
    TRY(fgk_paint_fixed(p->paint, x, y, z, Np, p->d.Ng / p->d.L, p->gA, p->gB, p->s));
    TRY(fgk_stream_sync(p->s));
    if (gridA) *gridA = p->gA;
    if (gridB) *gridB = p->gB;
    return 0;
}

int fftgal_power_fixed_bits(const fftgal_power_t *p) {
    return (int)lround(-log2(fgk_paint_fixed_scale(p->paint)));
}

int fftgal_power_spectra(fftgal_power_t *p, void **A, void **B) {
    if (A) *A = p->fA;
    if (B) *B = p->fB;
    return 0;
}

int fftgal_power_tuning(const fftgal_power_t *p, int *paint_tile, int *fft_cols, int *fft_rows) {
    if (paint_tile) *paint_tile = fgk_paint_plan_tile(p->paint);
    if (fft_cols) *fft_cols = p->tune.fft_cols;
    if (fft_rows) *fft_rows = p->tune.fft_rows;
    return 0;
}

int fftgal_power_use_graph(fftgal_power_t *p, int on) {
    p->use_graph = on ? 1 : 0;
    return 0;
}

int fftgal_power_fft_generic(fftgal_power_t *p, int on) {
    TRY(fgk_stream_sync(p->s));
    if (p->graph) {   /* the captured graph holds the other kernels */
        fgk_graph_destroy(p->graph);
        p->graph = NULL;
    }
    TRY(fgk_fft_force_generic(p->plan, on));
    return fgk_fft_f32_specialised(p->d.Ng) && !on ? 0 : 1;
}

void *fftgal_power_stream(fftgal_power_t *p) { return p->s; }

int fftgal_memcpy_d2h(void *dst, const void *src, unsigned long bytes) {
    TRY(fgk_init());
    TRY(fgk_d2h(dst, src, bytes, NULL));
    return fgk_stream_sync(NULL);
}
