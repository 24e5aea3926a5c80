/* Drop-in grid API (include/fft.h) on one MI355X.
 *
 * Each function keeps the signature, side effects and error behaviour of
 * its counterpart in eelregit/fftgal src/fft.c (cited per function) and
 * runs the arithmetic as FP64 HIP kernels through the fgk C-ABI.  The
 * grid lives on the device; the host array f mirrors it according to the
 * residency mode (fft.h).
 */
#define _GNU_SOURCE
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <time.h>
#include "internal.h"
#include "tune.h"
#include "../../../include/fft.h"
#include "../../../include/fftgal_ext.h"

int fftgal_verbose(void) {
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("FFTGAL_VERBOSE");
        v = (e && *e && strcmp(e, "0")) ? 1 : 0;
    }
    return v;
}

double fftgal_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static inline double cube(double x) { return x * x * x; }
static inline double quart(double x) { double s = x * x; return s * s; }

/* src/fft.c:239-245 */
void fftgal_winv_table(int Ng, double *winv) {
    winv[0] = 1;
    for (int i = 1; i <= Ng / 2; ++i) {
        double arg = M_PI * i / Ng;
        winv[i] = quart(arg / sin(arg));
        winv[Ng - i] = winv[i];
    }
}

/* src/fft.c:204-216: phase[d][j] = exp(-i onoff 2 pi j offset[d] / Ng) for
 * j <= Ng/2, conjugate mirrored; the mirror overwrites j = Ng/2, so the
 * Nyquist entry carries the -Ng/2 convention */
void fftgal_phase_table(int Ng, const double offset[3], int onoff, double *ph) {
    for (int d = 0; d < 3; ++d) {
        ph[2 * (d * Ng)] = 1;
        ph[2 * (d * Ng) + 1] = 0;
    }
    for (int d = 0; d < 3; ++d)
        for (int j = 1; j <= Ng / 2; ++j) {
            double arg = -onoff * 2 * M_PI * j / Ng * offset[d];
            double c = cos(arg), s = sin(arg);
            ph[2 * (d * Ng + j)] = c;
            ph[2 * (d * Ng + j) + 1] = s;
            ph[2 * (d * Ng + Ng - j)] = c;
            ph[2 * (d * Ng + Ng - j) + 1] = -s;
        }
}

static int env_resident(void) {
    const char *e = getenv("FFTGAL_DEVICE_RESIDENT");
    return e && *e && strcmp(e, "0") != 0;
}

void fftgal_dev_pull(fft_t *g) {
    fft_dev *d = fft_dev_of(g);
    if (!d->resident)
        DI_CHECK(fgk_h2d(d->d_f, g->f, (size_t)g->Ng3_pad * sizeof(double), d->s));
}

void fftgal_dev_push(fft_t *g) {
    fft_dev *d = fft_dev_of(g);
    if (!d->resident) {
        DI_CHECK(fgk_d2h(g->f, d->d_f, (size_t)g->Ng3_pad * sizeof(double), d->s));
        DI_CHECK(fgk_stream_sync(d->s));
    }
}

/* The role of FFTW_MEASURE planning (src/fft.c:36-43): time the FP64
 * transform pair on the grid's own device buffer with the generic kernels'
 * default launch shape and with half its columns / rows per workgroup, keep
 * the fastest (the buffer is zeroed afterwards, as the reference's paint
 * zeroes after planning, src/fft.c:64). */
static void measure_fft_shape(fft_dev *d, long n, fftgal_tuning *best) {
    fftgal_tune_lookup(NULL, 0, FFTGAL_TUNE_FP64_GRID, best);
    int C = 0, R = 0;
    DI_CHECK(fgk_fft_launch_shape(d->plan, &C, &R));
    const int cand[4][2] = {{0, 0}, {C / 2, 0}, {0, R / 2}, {C / 2, R / 2}};
    fgk_event_t e0 = NULL, e1 = NULL;
    DI_CHECK(fgk_event_create(&e0));
    DI_CHECK(fgk_event_create(&e1));
    float tbest = 1e30f;
    for (int c = 0; c < 4; ++c) {
        if ((c & 1) && C < 2) continue;
        if ((c & 2) && R < 2) continue;
        DI_CHECK(fgk_fft_set_tuning(d->plan, cand[c][0], cand[c][1]));
        float ms = 0.0f;
        for (int rep = 0; rep < 3; ++rep) {   /* the first pair warms up */
            DI_CHECK(fgk_event_record(e0, d->s));
            DI_CHECK(fgk_fft_r2c_f64(d->plan, d->d_f, 1.0, d->s));
            DI_CHECK(fgk_fft_c2r_f64(d->plan, d->d_f, 1.0, d->s));
            DI_CHECK(fgk_event_record(e1, d->s));
            DI_CHECK(fgk_event_sync(e1));
            float m = 0.0f;
            DI_CHECK(fgk_event_elapsed_ms(&m, e0, e1));
            if (rep) ms += m;
        }
        if (ms < tbest) {
            tbest = ms;
            best->fft_cols = cand[c][0];
            best->fft_rows = cand[c][1];
        }
    }
    DI_CHECK(fgk_fft_set_tuning(d->plan, best->fft_cols, best->fft_rows));
    DI_CHECK(fgk_fill_zero(d->d_f, (size_t)n * sizeof(double), d->s));
    DI_CHECK(fgk_stream_sync(d->s));
    fgk_event_destroy(e0);
    fgk_event_destroy(e1);
}

/* src/fft.c:13-47 */
fft_t *fft_init(int Ng, double L, char *wisdom) {
    if (fftgal_verbose())
        fprintf(stderr, "fft_init() making a %d^3 grid in a %g^3 box\n", Ng, L);
    DI_ASSERT(fgk_fft_supported_f64(Ng), "fft_init: Ng must be in [1, 1024] (or an even "
              "Ng with prime factors in {2,3,5,7,11,13})");
    /* a wisdom file that has no line for this grid will be written after
     * measuring: fail before any device work if it cannot be (the reference
     * asserts on the export, src/fft.c:40) */
    if (wisdom && strcmp(wisdom, "FFTW_ESTIMATE") != 0) {
        fftgal_tuning t;
        if (!fftgal_tune_lookup(wisdom, Ng, FFTGAL_TUNE_FP64_GRID, &t)) {
            FILE *fp = fopen(wisdom, "a");
            DI_ASSERT(fp != NULL, "fft_init: cannot write the wisdom file");
            fclose(fp);
        }
    }
    DI_CHECK(fgk_init());
    fft_t *self = (fft_t *)malloc(sizeof(fft_t));
    DI_ASSERT(self != NULL, "fft_init: out of memory");
    self->Ng = Ng;
    self->L = L;
    for (int i = 0; i < 3; ++i) self->offset[i] = 0;
    self->Ng_pad = 2 * (Ng / 2 + 1);
    self->Ng3_pad = (long)Ng * Ng * self->Ng_pad;
    /* calloc: a large grid comes from fresh zero pages, touched only when a
     * host transfer or the caller writes them (68.8 GB at Ng = 2048) */
    self->f = (double *)calloc((size_t)self->Ng3_pad, sizeof(double));
    DI_ASSERT(self->f != NULL, "fft_init: out of memory");

    fft_dev *d = (fft_dev *)calloc(1, sizeof(fft_dev));
    DI_ASSERT(d != NULL, "fft_init: out of memory");
    DI_CHECK(fgk_stream_create(&d->s));
    DI_CHECK(fgk_fft_plan_create(&d->plan, Ng, 8));
    DI_CHECK(fgk_malloc((void **)&d->d_f, (size_t)self->Ng3_pad * sizeof(double)));
    DI_CHECK(fgk_malloc((void **)&d->d_phase, (size_t)3 * Ng * 2 * sizeof(double)));
    DI_CHECK(fgk_malloc((void **)&d->d_winv, (size_t)Ng * sizeof(double)));
    DI_CHECK(fgk_fill_zero(d->d_f, (size_t)self->Ng3_pad * sizeof(double), d->s));
    double *winv = (double *)malloc((size_t)Ng * sizeof(double));
    DI_ASSERT(winv != NULL, "fft_init: out of memory");
    fftgal_winv_table(Ng, winv);
    DI_CHECK(fgk_h2d(d->d_winv, winv, (size_t)Ng * sizeof(double), d->s));
    DI_CHECK(fgk_stream_sync(d->s));
    free(winv);
    d->resident = env_resident();
    self->x2k = self->k2x = d;

    if (wisdom && strcmp(wisdom, "FFTW_ESTIMATE") != 0) {
        fftgal_tuning t;
        if (fftgal_tune_lookup(wisdom, Ng, FFTGAL_TUNE_FP64_GRID, &t)) {
            DI_CHECK(fgk_fft_set_tuning(d->plan, t.fft_cols, t.fft_rows));
            if (fftgal_verbose()) fprintf(stderr, "fft_init() imported wisdom from %s\n", wisdom);
        } else {
            const double t0 = fftgal_now();
            measure_fft_shape(d, self->Ng3_pad, &t);
            DI_ASSERT(fftgal_tune_store(wisdom, Ng, FFTGAL_TUNE_FP64_GRID, &t) == 0,
                      "fft_init: cannot write the wisdom file");
            if (fftgal_verbose()) {
                fprintf(stderr, "fft_init() exported wisdom to %s\n", wisdom);
                fprintf(stderr, "fft_init() %.2fs on measuring\n", fftgal_now() - t0);
            }
        }
    }
    return self;
}

/* src/fft.c:50-54 */
void fft_free(fft_t *self) {
    if (self != NULL) {
        fft_dev *d = fft_dev_of(self);
        if (d) {
            fgk_stream_sync(d->s);
            fgk_fft_plan_destroy(d->plan);
            fgk_free(d->d_f);
            fgk_free(d->d_dual);
            fgk_free(d->d_phase);
            fgk_free(d->d_winv);
            fgk_free(d->d_pos);
            fgk_stream_destroy(d->s);
            free(d);
        }
        free(self->f);
    }
    free(self);
}

static double *stage_positions(fft_dev *d, const gal_t *c) {
    if (c->Np > d->pos_cap) {
        fgk_stream_sync(d->s);
        fgk_free(d->d_pos);
        d->d_pos = NULL;
        DI_CHECK(fgk_malloc((void **)&d->d_pos, (size_t)3 * c->Np * sizeof(double)));
        d->pos_cap = c->Np;
    }
    size_t b = (size_t)c->Np * sizeof(double);
    DI_CHECK(fgk_h2d(d->d_pos, c->x, b, d->s));
    DI_CHECK(fgk_h2d(d->d_pos + c->Np, c->y, b, d->s));
    DI_CHECK(fgk_h2d(d->d_pos + 2 * c->Np, c->z, b, d->s));
    return d->d_pos;
}

/* the device part of fft_p2g: zero, paint data (+1) and randoms (-alpha),
 * normalise by 1/(nbar H^3) (src/fft.c:64-129) */
static void p2g_device(fft_t *self, const double *x, const double *y, const double *z, long Np,
                       double V, const gal_t *rand, const double offset[3]) {
    fft_dev *d = fft_dev_of(self);
    const int Ng = self->Ng;
    double Hinv = Ng / self->L;
    DI_CHECK(fgk_fill_zero(d->d_f, (size_t)self->Ng3_pad * sizeof(double), d->s));
    DI_CHECK(fgk_paint_f64(d->d_f, Ng, x, y, z, Np, Hinv, offset, 1.0, +1, d->s));
    if (rand != NULL) {
        double alpha = (double)Np / rand->Np;
        /* randoms need their own staging; data positions are no longer needed */
        DI_CHECK(fgk_stream_sync(d->s));
        const double *r = stage_positions(d, rand);
        DI_CHECK(fgk_paint_f64(d->d_f, Ng, r, r + rand->Np, r + 2 * rand->Np, rand->Np, Hinv,
                               offset, alpha, -1, d->s));
    }
    double H3 = cube(self->L / Ng);
    double nbar = Np / V;
    double Npcinv = 1 / nbar / H3;
    DI_CHECK(fgk_scale_f64(d->d_f, self->Ng3_pad, Npcinv, d->s));
}

static void reduce_offset(fft_t *self, double offset[3]) {
    for (int i = 0; i < 3; ++i) {
        offset[i] = remainder(offset[i], self->Ng);
        self->offset[i] = offset[i];
    }
}

/* src/fft.c:57-131 */
void fft_p2g(fft_t *self, gal_t *part, double offset[3]) {
    double t = fftgal_now();
    fft_dev *d = fft_dev_of(self);
    reduce_offset(self, offset);
    const double *p = stage_positions(d, part);
    p2g_device(self, p, p + part->Np, p + 2 * part->Np, part->Np, part->V, part->rand, offset);
    fftgal_dev_push(self);
    if (fftgal_verbose()) {
        fgk_stream_sync(d->s);
        fprintf(stderr, "fft_p2g() %.2fs\n", fftgal_now() - t);
    }
}

void fft_p2g_device(fft_t *self, const double *x, const double *y, const double *z, long Np,
                    double V, double offset[3]) {
    reduce_offset(self, offset);
    p2g_device(self, x, y, z, Np, V, NULL, offset);
    fftgal_dev_push(self);
}

/* src/fft.c:134-169, on the host with the reference's behaviour */
double *fft_g2p(fft_t *self, gal_t *pos) {
    fft_dev *d = fft_dev_of(self);
    if (d->resident) fft_sync_host(self);
    double *out = (double *)malloc((pos->Np > 0 ? (size_t)pos->Np : 1) * sizeof(double));
    DI_ASSERT(out != NULL, "fft_g2p: out of memory");
    const int Ng = self->Ng;
    const double Hinv = Ng / self->L;
    for (long p = 0; p < pos->Np; ++p) {
        double u[3] = {pos->x[p] * Hinv - self->offset[0], pos->y[p] * Hinv - self->offset[1],
                       pos->z[p] * Hinv - self->offset[2]};
        int g[3][4];
        double w[3][4];
        for (int a = 0; a < 3; ++a) {
            int g1 = (int)floor(u[a]);
            double dd = u[a] - g1;
            g1 = (g1 % Ng + Ng) % Ng;
            g[a][0] = (g1 - 1 + Ng) % Ng;
            g[a][1] = g1;
            g[a][2] = (g1 + 1) % Ng;
            g[a][3] = (g1 + 2) % Ng;
            double e = 1 - dd;
            w[a][0] = e * e * e / 6;
            w[a][1] = ((3 * dd - 6) * (dd * dd) + 4) / 6;
            w[a][2] = (((-3 * dd + 3) * dd + 3) * dd + 1) / 6;
            w[a][3] = dd * dd * dd / 6;
        }
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k)
                    out[p] = F(self, g[0][i], g[1][j], g[2][k]) * w[0][i] * w[1][j] * w[2][k];
    }
    return out;
}

static void phase_device(fft_t *self, int onoff) {
    fft_dev *d = fft_dev_of(self);
    const int Ng = self->Ng;
    double *ph = (double *)malloc((size_t)3 * Ng * 2 * sizeof(double));
    DI_ASSERT(ph != NULL, "fft_offset_phase: out of memory");
    fftgal_phase_table(Ng, self->offset, onoff > 0 ? 1 : -1, ph);
    DI_CHECK(fgk_h2d(d->d_phase, ph, (size_t)3 * Ng * 2 * sizeof(double), d->s));
    DI_CHECK(fgk_offset_phase_f64(d->d_f, Ng, d->d_phase, d->s));
    DI_CHECK(fgk_stream_sync(d->s));  /* ph is host memory in flight */
    free(ph);
}

static void x2k_device(fft_t *self, int phase_on) {
    fft_dev *d = fft_dev_of(self);
    double H3 = cube(self->L / self->Ng);
    DI_CHECK(fgk_fft_r2c_f64(d->plan, d->d_f, H3, d->s));
    if (phase_on) phase_device(self, 1);
}

/* src/fft.c:172-181 */
void fft_x2k(fft_t *self, int offset_phase_on) {
    double t = fftgal_now();
    fftgal_dev_pull(self);
    x2k_device(self, offset_phase_on);
    fftgal_dev_push(self);
    if (fftgal_verbose()) {
        fgk_stream_sync(fft_dev_of(self)->s);
        fprintf(stderr, "fft_x2k() %.2fs\n", fftgal_now() - t);
    }
}

/* src/fft.c:184-193 */
void fft_k2x(fft_t *self, int offset_phase_off) {
    double t = fftgal_now();
    fft_dev *d = fft_dev_of(self);
    fftgal_dev_pull(self);
    DI_CHECK(fgk_fft_c2r_f64(d->plan, d->d_f, 1 / cube(self->L), d->s));
    if (offset_phase_off) phase_device(self, -1);
    fftgal_dev_push(self);
    if (fftgal_verbose()) {
        fgk_stream_sync(d->s);
        fprintf(stderr, "fft_k2x() %.2fs\n", fftgal_now() - t);
    }
}

/* super-sample modes of sub-volumes: the k-space loop, inverse transform
 * and corner reads of src/app/DLsb.c:66-96 (kind 0, with the window tables
 * of DLsb.c:42-49) and src/app/DLss.c:62-89 (kind 1) on the device */
void fft_dl_filter(fft_t *self, int kind, int Nsub, const double los[3], int ell, double bias,
                   double *delta) {
    const int Ng = self->Ng;
    DI_ASSERT(kind == 0 || kind == 1, "fft_dl_filter: kind must be 0 (sub box) or 1 (sub sphere)");
    DI_ASSERT(Nsub >= 1 && Ng % Nsub == 0, "fft_dl_filter: Nsub must divide Ng");
    DI_ASSERT(ell == 0 || ell == 2 || ell == 4, "fft_dl_filter: ell must be 0, 2 or 4");
    fft_dev *d = fft_dev_of(self);
    const int Ngsub = Ng / Nsub, n = Nsub * Nsub * Nsub;
    const size_t ntab = (size_t)Ng + (size_t)3 * Ng * 2;
    double *tab = (double *)malloc(ntab * sizeof(double));
    DI_ASSERT(tab != NULL, "fft_dl_filter: out of memory");
    double *wamp = tab, *wph = tab + Ng;
    wamp[0] = 1;
    for (int i = 1; i < Ng; ++i) wamp[i] = sin(M_PI * i / Nsub) / sin(M_PI * i / Ng) / Ngsub;
    for (int i = 0; i < 3 * Ng; ++i) {
        wph[2 * i] = cos(M_PI * i * (Ngsub - 1) / Ng);
        wph[2 * i + 1] = sin(M_PI * i * (Ngsub - 1) / Ng);
    }
    double *dtab = NULL, *dout = NULL;
    DI_CHECK(fgk_malloc((void **)&dtab, ntab * sizeof(double)));
    DI_CHECK(fgk_malloc((void **)&dout, (size_t)n * sizeof(double)));
    DI_CHECK(fgk_h2d(dtab, tab, ntab * sizeof(double), d->s));
    fftgal_dev_pull(self);
    DI_CHECK(fgk_dl_filter_f64(d->d_f, Ng, kind, Nsub, los, ell, dtab, dtab + Ng, d->s));
    DI_CHECK(fgk_fft_c2r_f64(d->plan, d->d_f, 1 / cube(self->L), d->s));   /* fft_k2x(grid, 0) */
    DI_CHECK(fgk_dl_gather_f64(d->d_f, Ng, Nsub, bias, dout, d->s));
    DI_CHECK(fgk_d2h(delta, dout, (size_t)n * sizeof(double), d->s));
    DI_CHECK(fgk_stream_sync(d->s));
    fftgal_dev_push(self);
    fgk_free(dtab);
    fgk_free(dout);
    free(tab);
}

/* the window deconvolution of src/app/Psb.c:73-85 (kind 0) and
 * src/app/Pss.c:70-87 (kind 1): fft_k2x(grid, 0), the correlation function
 * divided by the sub-volume's window autocorrelation, fft_x2k(grid, 0) */
void fft_subvolume_deconv(fft_t *self, int kind, int Nsub) {
    const int Ng = self->Ng;
    DI_ASSERT(kind == 0 || kind == 1, "fft_subvolume_deconv: kind must be 0 (box) or 1 (sphere)");
    DI_ASSERT(Nsub >= 1 && Ng % Nsub == 0, "fft_subvolume_deconv: Nsub must divide Ng");
    double t = fftgal_now();
    fft_dev *d = fft_dev_of(self);
    fftgal_dev_pull(self);
    DI_CHECK(fgk_fft_c2r_f64(d->plan, d->d_f, 1 / cube(self->L), d->s));
    DI_CHECK(fgk_subvolume_winv_f64(d->d_f, Ng, kind, Nsub, d->s));
    DI_CHECK(fgk_fft_r2c_f64(d->plan, d->d_f, cube(self->L / Ng), d->s));
    fftgal_dev_push(self);
    if (fftgal_verbose()) {
        fgk_stream_sync(d->s);
        fprintf(stderr, "fft_subvolume_deconv() %.2fs\n", fftgal_now() - t);
    }
}

/* src/fft.c:196-233 */
void fft_offset_phase(fft_t *self, int onoff) {
    DI_ASSERT(onoff != 0, "fft_offset_phase: onoff must be nonzero");
    fftgal_dev_pull(self);
    phase_device(self, onoff);
    fftgal_dev_push(self);
}

/* src/fft.c:236-254 */
void fft_deconv(fft_t *self) {
    double t = fftgal_now();
    fft_dev *d = fft_dev_of(self);
    fftgal_dev_pull(self);
    DI_CHECK(fgk_deconv_f64(d->d_f, self->Ng, d->d_winv, d->s));
    fftgal_dev_push(self);
    if (fftgal_verbose()) {
        fgk_stream_sync(d->s);
        fprintf(stderr, "fft_deconv() %.2fs\n", fftgal_now() - t);
    }
}

/* src/fft.c:257-276 for a catalog without randoms, with both grids kept on
 * the device (the dual grid is copied device to device, not exported) */
static void p2k_device(fft_t *self, const double *x, const double *y, const double *z, long Np,
                       double V) {
    fft_dev *d = fft_dev_of(self);
    if (!d->d_dual)
        DI_CHECK(fgk_malloc((void **)&d->d_dual, (size_t)self->Ng3_pad * sizeof(double)));
    double off_dual[3] = {0.5, 0.5, 0.5};
    reduce_offset(self, off_dual);
    p2g_device(self, x, y, z, Np, V, NULL, off_dual);
    x2k_device(self, 1);
    DI_CHECK(fgk_d2d(d->d_dual, d->d_f, (size_t)self->Ng3_pad * sizeof(double), d->s));
    double off_none[3] = {0, 0, 0};
    reduce_offset(self, off_none);
    p2g_device(self, x, y, z, Np, V, NULL, off_none);
    x2k_device(self, 0);
    DI_CHECK(fgk_avg_f64(d->d_f, d->d_dual, self->Ng3_pad, d->s));
    DI_CHECK(fgk_deconv_f64(d->d_f, self->Ng, d->d_winv, d->s));
}

void fft_p2k(fft_t *self, gal_t *part) {
    double t = fftgal_now();
    fft_dev *d = fft_dev_of(self);
    if (fftgal_verbose()) fprintf(stderr, "fft_p2k() interlacing\n");
    if (part->rand == NULL) {
        const double *p = stage_positions(d, part);
        p2k_device(self, p, p + part->Np, p + 2 * part->Np, part->Np, part->V);
    } else {
        /* with randoms the staging buffer alternates between catalogs: use
         * the per-function path, which re-stages before each paint */
        if (!d->d_dual)
            DI_CHECK(fgk_malloc((void **)&d->d_dual, (size_t)self->Ng3_pad * sizeof(double)));
        double off_dual[3] = {0.5, 0.5, 0.5};
        reduce_offset(self, off_dual);
        const double *p = stage_positions(d, part);
        p2g_device(self, p, p + part->Np, p + 2 * part->Np, part->Np, part->V, part->rand,
                   off_dual);
        x2k_device(self, 1);
        DI_CHECK(fgk_d2d(d->d_dual, d->d_f, (size_t)self->Ng3_pad * sizeof(double), d->s));
        double off_none[3] = {0, 0, 0};
        reduce_offset(self, off_none);
        DI_CHECK(fgk_stream_sync(d->s));
        p = stage_positions(d, part);
        p2g_device(self, p, p + part->Np, p + 2 * part->Np, part->Np, part->V, part->rand,
                   off_none);
        x2k_device(self, 0);
        DI_CHECK(fgk_avg_f64(d->d_f, d->d_dual, self->Ng3_pad, d->s));
        DI_CHECK(fgk_deconv_f64(d->d_f, self->Ng, d->d_winv, d->s));
    }
    fftgal_dev_push(self);
    if (fftgal_verbose()) {
        fgk_stream_sync(d->s);
        fprintf(stderr, "fft_p2k() %.2fs\n", fftgal_now() - t);
    }
}

void fft_p2k_device(fft_t *self, const double *x, const double *y, const double *z, long Np,
                    double V) {
    p2k_device(self, x, y, z, Np, V);
    fftgal_dev_push(self);
}

/* src/fft.c:279-283: malloc'd, released by the caller with free() */
double *fft_exportf(fft_t *self) {
    fft_dev *d = fft_dev_of(self);
    double *dest = (double *)malloc((size_t)self->Ng3_pad * sizeof(double));
    DI_ASSERT(dest != NULL, "fft_exportf: out of memory");
    if (d->resident) {
        DI_CHECK(fgk_d2h(dest, d->d_f, (size_t)self->Ng3_pad * sizeof(double), d->s));
        DI_CHECK(fgk_stream_sync(d->s));
    } else {
        memcpy(dest, self->f, (size_t)self->Ng3_pad * sizeof(double));
    }
    return dest;
}

/* src/fft.c:286-289 */
void fft_importf(fft_t *self, double *src) {
    DI_ASSERT(src != NULL, "fft_importf: NULL source");
    fft_dev *d = fft_dev_of(self);
    memcpy(self->f, src, (size_t)self->Ng3_pad * sizeof(double));
    if (d->resident) {
        DI_CHECK(fgk_h2d(d->d_f, src, (size_t)self->Ng3_pad * sizeof(double), d->s));
        DI_CHECK(fgk_stream_sync(d->s));
    }
}

/* ---- fftgal_ext.h ---- */

void fft_set_device_resident(fft_t *self, int on) {
    fft_dev *d = fft_dev_of(self);
    if (d->resident && !on) fft_sync_host(self);
    if (!d->resident && on) {
        DI_CHECK(fgk_h2d(d->d_f, self->f, (size_t)self->Ng3_pad * sizeof(double), d->s));
        DI_CHECK(fgk_stream_sync(d->s));
    }
    d->resident = on ? 1 : 0;
}

void fft_sync_host(fft_t *self) {
    fft_dev *d = fft_dev_of(self);
    DI_CHECK(fgk_d2h(self->f, d->d_f, (size_t)self->Ng3_pad * sizeof(double), d->s));
    DI_CHECK(fgk_stream_sync(d->s));
}

double *fft_device_grid(fft_t *self) { return fft_dev_of(self)->d_f; }
