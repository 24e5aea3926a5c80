/*
 * FP64 CPU restatement of the reference power-spectrum path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker, never the product.
 *
 * Follows eelregit/fftgal src/fft.c, src/power.c, src/gal.c, src/gal.h in
 * evaluation order; the FFT is the oracle's own (fft64.c) instead of FFTW.
 * Deliberate differences from the reference, none of which changes a result:
 *   - no stderr timing lines, no FFTW planning / wisdom (fft.c:26-45);
 *   - the sum-of-counts check uses long arithmetic (power.c:76,153 overflow
 *     int for Ng >= 1291) and is returned, not asserted;
 *   - Pmu takes Nmu as a parameter (power.c:111 fixes it to 3).
 */
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <assert.h>
#include "oracle.h"

/* sqrt(3) to double precision (the value gsl_math.h names M_SQRT3, used at power.c:36,110) */
#define ORC_SQRT3 1.73205080756887729352744634151

#define FIDX(g, i, j, k) (((long)(i) * (g)->Ng + (j)) * (g)->Ng_pad + (k))

static inline double sq(double x) { return x * x; }          /* gsl_pow_2 */
static inline double cube(double x) { return x * x * x; }    /* gsl_pow_3 */
static inline double quart(double x) { double s = x * x; return s * s; } /* gsl_pow_4 */

/* ---------------------------------------------------------------- catalog */

/* src/io/qpm.c:9-34: count the '\n' characters, then read six numbers per
 * line with fscanf, skipping the halo mass, satellite flag and halo id.
 * Returns NULL (where the reference asserts) if the file cannot be opened or
 * a line does not match. */
ogal_t *oracle_loadqpm_cubic(const char *file, double L) {
    FILE *fp = fopen(file, "r");
    if (!fp) return NULL;
    long Np = 0;
    int c;
    while ((c = fgetc(fp)) != EOF)
        if (c == '\n') ++Np;
    rewind(fp);
    ogal_t *part = oracle_gal_init(Np, L * L * L);
    size_t n = Np > 0 ? (size_t)Np : 1;
    part->vx = (double *)malloc(n * sizeof(double));
    part->vy = (double *)malloc(n * sizeof(double));
    part->vz = (double *)malloc(n * sizeof(double));
    for (long i = 0; i < Np; ++i) {
        int r = fscanf(fp, "%lf %lf %lf %lf %lf %lf %*f %*d %*d\n", part->x + i, part->y + i,
                       part->z + i, part->vx + i, part->vy + i, part->vz + i);
        if (r != 6) {
            fclose(fp);
            oracle_gal_free(part);
            return NULL;
        }
    }
    fclose(fp);
    return part;
}

/* MT19937 (Matsumoto and Nishimura 1998, 2002 seeding), the generator behind
 * gsl_rng_default in src/gal.c:101-106,142-147: seed 0 means 4357, and
 * gsl_rng_uniform is the 32-bit output over 2^32.  Written in the original
 * paper's two-loop form. */
typedef struct orc_mt {
    unsigned long mt[624];
    int mti;
} orc_mt;

static void orc_mt_seed(orc_mt *r, unsigned long s) {
    if (s == 0) s = 4357;
    r->mt[0] = s & 0xffffffffUL;
    for (int i = 1; i < 624; ++i)
        r->mt[i] = (1812433253UL * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + i) & 0xffffffffUL;
    r->mti = 624;
}

static double orc_mt_uniform(orc_mt *r) {
    static const unsigned long mag01[2] = {0x0UL, 0x9908b0dfUL};
    unsigned long y;
    if (r->mti >= 624) {
        int kk;
        for (kk = 0; kk < 624 - 397; ++kk) {
            y = (r->mt[kk] & 0x80000000UL) | (r->mt[kk + 1] & 0x7fffffffUL);
            r->mt[kk] = r->mt[kk + 397] ^ (y >> 1) ^ mag01[y & 0x1UL];
        }
        for (; kk < 623; ++kk) {
            y = (r->mt[kk] & 0x80000000UL) | (r->mt[kk + 1] & 0x7fffffffUL);
            r->mt[kk] = r->mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 0x1UL];
        }
        y = (r->mt[623] & 0x80000000UL) | (r->mt[0] & 0x7fffffffUL);
        r->mt[623] = r->mt[396] ^ (y >> 1) ^ mag01[y & 0x1UL];
        r->mti = 0;
    }
    y = r->mt[r->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680UL;
    y ^= (y << 15) & 0xefc60000UL;
    y ^= (y >> 18);
    return (double)(y & 0xffffffffUL) / 4294967296.0;
}

static unsigned long orc_env_seed(void) {
    const char *e = getenv("GSL_RNG_SEED");
    return e ? strtoul(e, NULL, 0) : 0;
}

static double orc_pdist2(double x1, double y1, double z1, double x0, double y0, double z0,
                         double L) {
    double dx = remainder(x1 - x0, L), dy = remainder(y1 - y0, L), dz = remainder(z1 - z0, L);
    return dx * dx + dy * dy + dz * dz;
}

/* src/gal.c:78-118 */
ogal_t *oracle_gal_subbox(const ogal_t *g, double box[7], double alpha) {
    static orc_mt rng;
    static int seeded = 0;
    double x0 = box[0], x1 = box[1], y0 = box[2], y1 = box[3], z0 = box[4], z1 = box[5];
    double L = box[6];
    x1 = x0 + remainder(x1 - x0, L);
    y1 = y0 + remainder(y1 - y0, L);
    z1 = z0 + remainder(z1 - z0, L);
    double Vsub = (x1 - x0) * (y1 - y0) * (z1 - z0);
    ogal_t *sub = oracle_gal_init(g->Np, Vsub);
    long n = 0;
    for (long p = 0; p < g->Np; ++p)
        if (remainder(g->x[p] - x0, L) >= 0 && remainder(g->x[p] - x1, L) < 0 &&
            remainder(g->y[p] - y0, L) >= 0 && remainder(g->y[p] - y1, L) < 0 &&
            remainder(g->z[p] - z0, L) >= 0 && remainder(g->z[p] - z1, L) < 0) {
            sub->x[n] = g->x[p];
            sub->y[n] = g->y[p];
            sub->z[n] = g->z[p];
            ++n;
        }
    sub->Np = n;
    long nr = lround(n / alpha);
    ogal_t *rand = oracle_gal_init(nr, Vsub);
    if (!seeded) {
        orc_mt_seed(&rng, orc_env_seed());
        seeded = 1;
    }
    for (long p = 0; p < nr; ++p) {
        rand->x[p] = x0 + (x1 - x0) * orc_mt_uniform(&rng);
        rand->y[p] = y0 + (y1 - y0) * orc_mt_uniform(&rng);
        rand->z[p] = z0 + (z1 - z0) * orc_mt_uniform(&rng);
    }
    sub->rand = rand;
    return sub;
}

/* src/gal.c:121-158 */
ogal_t *oracle_gal_subsphere(const ogal_t *g, double sphere[5], double alpha) {
    static orc_mt rng;
    static int seeded = 0;
    double x0 = sphere[0], y0 = sphere[1], z0 = sphere[2], R = sphere[3], L = sphere[4];
    double R2 = R * R, Vsub = 4 * M_PI / 3 * R2 * R;
    ogal_t *sub = oracle_gal_init(g->Np, Vsub);
    long n = 0;
    for (long p = 0; p < g->Np; ++p)
        if (orc_pdist2(g->x[p], g->y[p], g->z[p], x0, y0, z0, L) < R2) {
            sub->x[n] = g->x[p];
            sub->y[n] = g->y[p];
            sub->z[n] = g->z[p];
            ++n;
        }
    sub->Np = n;
    long nr = lround(n / alpha);
    ogal_t *rand = oracle_gal_init(nr, Vsub);
    if (!seeded) {
        orc_mt_seed(&rng, orc_env_seed());
        seeded = 1;
    }
    for (long p = 0; p < nr; ++p) {
        do {
            rand->x[p] = x0 + R * (2 * orc_mt_uniform(&rng) - 1);
            rand->y[p] = y0 + R * (2 * orc_mt_uniform(&rng) - 1);
            rand->z[p] = z0 + R * (2 * orc_mt_uniform(&rng) - 1);
        } while (orc_pdist2(rand->x[p], rand->y[p], rand->z[p], x0, y0, z0, L) >= R2);
    }
    sub->rand = rand;
    return sub;
}

/* src/gal.h:19-26 */
void oracle_hat(double vec[3]) {
    double amp = sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
    if (amp > 1e-7) {
        vec[0] /= amp;
        vec[1] /= amp;
        vec[2] /= amp;
    }
}

/* src/gal.c:11-21 */
ogal_t *oracle_gal_init(long Np, double V) {
    ogal_t *g = (ogal_t *)calloc(1, sizeof(ogal_t));
    assert(g != NULL);
    g->Np = Np;
    g->V = V;
    size_t n = Np > 0 ? (size_t)Np : 1;
    g->x = (double *)malloc(n * sizeof(double));
    g->y = (double *)malloc(n * sizeof(double));
    g->z = (double *)malloc(n * sizeof(double));
    assert(g->x && g->y && g->z);
    return g;
}

/* src/gal.c:41-48 */
void oracle_gal_free(ogal_t *g) {
    if (g != NULL) {
        oracle_gal_free(g->rand);
        free(g->x); free(g->y); free(g->z);
        free(g->vx); free(g->vy); free(g->vz);
    }
    free(g);
}

/* src/gal.c:51-60 */
void oracle_gal_wrap(ogal_t *g, double L) {
    double Linv = 1 / L;
    for (long p = 0; p < g->Np; ++p) {
        g->x[p] -= L * floor(g->x[p] * Linv);
        g->y[p] -= L * floor(g->y[p] * Linv);
        g->z[p] -= L * floor(g->z[p] * Linv);
    }
}

/* src/gal.c:63-77: plane-parallel shift x + (v.los)/(aH) los, new catalog */
ogal_t *oracle_gal_rsd(const ogal_t *g, double los[3], double aH) {
    oracle_hat(los);
    double aHinv = 1 / aH;
    ogal_t *d = oracle_gal_init(g->Np, g->V);
    for (long p = 0; p < d->Np; ++p) {
        double vdl = los[0] * g->vx[p] + los[1] * g->vy[p] + los[2] * g->vz[p];
        d->x[p] = g->x[p] + aHinv * vdl * los[0];
        d->y[p] = g->y[p] + aHinv * vdl * los[1];
        d->z[p] = g->z[p] + aHinv * vdl * los[2];
    }
    return d;
}

/* ------------------------------------------------------------------- grid */

/* src/fft.c:13-24 (allocation; the plans are the oracle's own) */
offt_t *oracle_fft_init(int Ng, double L) {
    assert(Ng >= 1);
    offt_t *g = (offt_t *)calloc(1, sizeof(offt_t));
    assert(g != NULL);
    g->Ng = Ng;
    g->L = L;
    g->Ng_pad = 2 * (Ng / 2 + 1);
    g->Ng3_pad = (long)Ng * Ng * g->Ng_pad;
    g->f = (double *)calloc((size_t)g->Ng3_pad, sizeof(double));
    assert(g->f != NULL);
    g->pN = (offt64_plan *)calloc(1, sizeof(offt64_plan));
    g->pH = (offt64_plan *)calloc(1, sizeof(offt64_plan));
    assert(g->pN && g->pH);
    assert(offt64_plan_init(g->pN, Ng) == 0);
    assert(offt64_plan_init(g->pH, Ng / 2 > 0 ? Ng / 2 : 1) == 0);
    return g;
}

/* src/fft.c:50-54 */
void oracle_fft_free(offt_t *g) {
    if (g) {
        free(g->f);
        offt64_plan_free(g->pN);
        offt64_plan_free(g->pH);
        free(g->pN);
        free(g->pH);
    }
    free(g);
}

/* PCS kernel weights and wrapped node indices for one axis (fft.c:68-88) */
static inline void pcs_axis(double u, int Ng, int idx[4], double w[4]) {
    int g1 = (int)floor(u);
    double d = u - g1;
    g1 = (g1 % Ng + Ng) % Ng;
    idx[0] = (g1 - 1 + Ng) % Ng;
    idx[1] = g1;
    idx[2] = (g1 + 1) % Ng;
    idx[3] = (g1 + 2) % Ng;
    w[0] = cube(1 - d) / 6;
    w[1] = ((3 * d - 6) * sq(d) + 4) / 6;
    w[2] = (((-3 * d + 3) * d + 3) * d + 1) / 6;
    w[3] = cube(d) / 6;
}

typedef struct paint_args {
    offt_t *g;
    const ogal_t *c;
    const double *off;
    double Hinv, scale;
    int negate;
} paint_args;

/* The particle loop of src/fft.c:67-93 (and its randoms twin 94-124),
 * restricted to the cells of i-planes [i0, i1).  Every cell is owned by one
 * range and receives its terms in the serial particle order, so any split
 * over threads gives the serial result bit for bit. */
static void paint_slab(void *arg, long i0, long i1) {
    const paint_args *a = (const paint_args *)arg;
    offt_t *g = a->g;
    const ogal_t *c = a->c;
    const int Ng = g->Ng;
    const int all = i0 == 0 && i1 == Ng;
    for (long p = 0; p < c->Np; ++p) {
        int gi[4], gj[4], gk[4];
        double wx[4], wy[4], wz[4];
        pcs_axis(c->x[p] * a->Hinv - a->off[0], Ng, gi, wx);
        if (!all) {
            int hit = 0;
            for (int i = 0; i < 4; ++i) hit |= gi[i] >= i0 && gi[i] < i1;
            if (!hit) continue;
        }
        pcs_axis(c->y[p] * a->Hinv - a->off[1], Ng, gj, wy);
        pcs_axis(c->z[p] * a->Hinv - a->off[2], Ng, gk, wz);
        for (int i = 0; i < 4; ++i) {
            if (gi[i] < i0 || gi[i] >= i1) continue;
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    if (a->negate)
                        g->f[FIDX(g, gi[i], gj[j], gk[k])] -= a->scale * wx[i] * wy[j] * wz[k];
                    else
                        g->f[FIDX(g, gi[i], gj[j], gk[k])] += wx[i] * wy[j] * wz[k];
                }
        }
    }
}

static void paint_catalog(offt_t *g, const ogal_t *c, const double off[3],
                          double Hinv, double scale, int negate) {
    paint_args a = {g, c, off, Hinv, scale, negate};
    orc_parallel(g->Ng, paint_slab, &a);
}

typedef struct sweep_args {
    offt_t *g;
    double s;
} sweep_args;

static void scale_range(void *arg, long q0, long q1) {
    const sweep_args *a = (const sweep_args *)arg;
    for (long q = q0; q < q1; ++q)
        a->g->f[q] *= a->s;
}

static void zero_range(void *arg, long q0, long q1) {
    const sweep_args *a = (const sweep_args *)arg;
    memset(a->g->f + q0, 0, (size_t)(q1 - q0) * sizeof(double));
}

/* f[q] *= s over the whole padded grid (elementwise, any split) */
static void grid_scale(offt_t *g, double s) {
    sweep_args a = {g, s};
    orc_parallel(g->Ng3_pad, scale_range, &a);
}

/* src/fft.c:57-131 */
void oracle_fft_p2g(offt_t *g, const ogal_t *part, double offset[3]) {
    const int Ng = g->Ng;
    for (int i = 0; i < 3; ++i) {
        offset[i] = remainder(offset[i], Ng);
        g->offset[i] = offset[i];
    }
    sweep_args z = {g, 0};
    orc_parallel(g->Ng3_pad, zero_range, &z);
    double Hinv = Ng / g->L;
    paint_catalog(g, part, offset, Hinv, 1.0, 0);
    if (part->rand != NULL) {
        double alpha = (double)part->Np / part->rand->Np;
        paint_catalog(g, part->rand, offset, Hinv, alpha, 1);
    }
    double H3 = cube(g->L / Ng);
    double nbar = part->Np / part->V;
    double Npcinv = 1 / nbar / H3;
    grid_scale(g, Npcinv);
}

typedef struct phase_args {
    offt_t *g;
    const double *ph;   /* [3][Ng] complex factors */
} phase_args;

/* the mode loop of src/fft.c:217-232 over i-planes [i0, i1) */
static void phase_planes(void *arg, long i0, long i1) {
    const phase_args *a = (const phase_args *)arg;
    offt_t *g = a->g;
    const int Ng = g->Ng;
    const double *ph = a->ph;
#define PH(d, j, c) ph[((d) * Ng + (j)) * 2 + (c)]
    for (long i = i0; i < i1; ++i)
        for (int j = 0; j < Ng; ++j)
            for (int k = 0; k <= Ng / 2; ++k) {
                double pr = PH(0, i, 0) * PH(1, j, 0) * PH(2, k, 0)
                          - PH(0, i, 0) * PH(1, j, 1) * PH(2, k, 1)
                          - PH(0, i, 1) * PH(1, j, 0) * PH(2, k, 1)
                          - PH(0, i, 1) * PH(1, j, 1) * PH(2, k, 0);
                double pi = PH(0, i, 1) * PH(1, j, 0) * PH(2, k, 0)
                          + PH(0, i, 0) * PH(1, j, 1) * PH(2, k, 0)
                          + PH(0, i, 0) * PH(1, j, 0) * PH(2, k, 1)
                          - PH(0, i, 1) * PH(1, j, 1) * PH(2, k, 1);
                double *c = g->f + FIDX(g, i, j, 2 * k);
                double fr = c[0], fi = c[1];
                c[0] = fr * pr - fi * pi;
                c[1] = fr * pi + fi * pr;
            }
#undef PH
}

/* src/fft.c:196-233 */
void oracle_fft_offset_phase(offt_t *g, int onoff) {
    assert(onoff != 0);
    onoff = onoff > 0 ? 1 : -1;
    const int Ng = g->Ng;
    double *ph = (double *)malloc((size_t)3 * Ng * 2 * sizeof(double));
    assert(ph != NULL);
#define PH(d, j, c) ph[((d) * Ng + (j)) * 2 + (c)]
    for (int d = 0; d < 3; ++d) { PH(d, 0, 0) = 1; PH(d, 0, 1) = 0; }
    for (int d = 0; d < 3; ++d)
        for (int j = 1; j <= Ng / 2; ++j) {
            double arg = -onoff * 2 * M_PI * j / Ng * g->offset[d];
            PH(d, j, 0) = cos(arg);
            PH(d, j, 1) = sin(arg);
            PH(d, Ng - j, 0) = PH(d, j, 0);      /* overwrites j = Ng/2: Nyquist = -Ng/2 */
            PH(d, Ng - j, 1) = -PH(d, j, 1);
        }
    phase_args a = {g, ph};
    orc_parallel(Ng, phase_planes, &a);
#undef PH
    free(ph);
}

/* src/fft.c:172-181 */
void oracle_fft_x2k(offt_t *g, int offset_phase_on) {
    offt64_r2c_3d(g);
    double H3 = cube(g->L / g->Ng);
    grid_scale(g, H3);
    if (offset_phase_on)
        oracle_fft_offset_phase(g, 1);
}

/* src/fft.c:184-193 */
void oracle_fft_k2x(offt_t *g, int offset_phase_off) {
    offt64_c2r_3d(g);
    double L3inv = 1 / cube(g->L);
    grid_scale(g, L3inv);
    if (offset_phase_off)
        oracle_fft_offset_phase(g, -1);
}

typedef struct deconv_args {
    offt_t *g;
    const double *Winv;
} deconv_args;

/* the mode loop of src/fft.c:246-252 over i-planes [i0, i1) */
static void deconv_planes(void *arg, long i0, long i1) {
    const deconv_args *a = (const deconv_args *)arg;
    offt_t *g = a->g;
    const int Ng = g->Ng;
    const double *Winv = a->Winv;
    for (long i = i0; i < i1; ++i)
        for (int j = 0; j < Ng; ++j)
            for (int k = 0; k <= Ng / 2; ++k) {
                double W3 = Winv[i] * Winv[j] * Winv[k];
                g->f[FIDX(g, i, j, 2 * k)] *= W3;
                g->f[FIDX(g, i, j, 2 * k + 1)] *= W3;
            }
}

/* src/fft.c:236-254 */
void oracle_fft_deconv(offt_t *g) {
    const int Ng = g->Ng;
    double *Winv = (double *)malloc((size_t)Ng * sizeof(double));
    assert(Winv != NULL);
    Winv[0] = 1;
    for (int i = 1; i <= Ng / 2; ++i) {
        double arg = M_PI * i / Ng;
        Winv[i] = quart(arg / sin(arg));
        Winv[Ng - i] = Winv[i];
    }
    deconv_args a = {g, Winv};
    orc_parallel(Ng, deconv_planes, &a);
    free(Winv);
}

/* src/fft.c:279-283 */
double *oracle_fft_exportf(const offt_t *g) {
    double *d = (double *)malloc((size_t)g->Ng3_pad * sizeof(double));
    assert(d != NULL);
    memcpy(d, g->f, (size_t)g->Ng3_pad * sizeof(double));
    return d;
}

/* src/fft.c:286-289 */
void oracle_fft_importf(offt_t *g, const double *src) {
    assert(src != NULL);
    memcpy(g->f, src, (size_t)g->Ng3_pad * sizeof(double));
}

typedef struct avg_args {
    offt_t *g;
    const double *fd;
} avg_args;

static void avg_range(void *arg, long q0, long q1) {
    const avg_args *a = (const avg_args *)arg;
    for (long q = q0; q < q1; ++q)
        a->g->f[q] = 0.5 * (a->fd[q] + a->g->f[q]);
}

/* src/fft.c:257-276: interlacing = dual (offset 1/2, phase on) + plain, average, deconvolve */
void oracle_fft_p2k(offt_t *g, const ogal_t *part) {
    double od[3] = {0.5, 0.5, 0.5};
    oracle_fft_p2g(g, part, od);
    oracle_fft_x2k(g, 1);
    double *fd = oracle_fft_exportf(g);
    double on[3] = {0, 0, 0};
    oracle_fft_p2g(g, part, on);
    oracle_fft_x2k(g, 0);
    avg_args a = {g, fd};
    orc_parallel(g->Ng3_pad, avg_range, &a);
    oracle_fft_deconv(g);
    free(fd);
}

/* ------------------------------------------------------------- estimators */

typedef struct P_args {
    offt_t *g;
    double Vinv, Pshot;
} P_args;

/* the mode loop of src/power.c:19-25 over i-planes [i0, i1) */
static void P_planes(void *arg, long i0, long i1) {
    const P_args *a = (const P_args *)arg;
    offt_t *g = a->g;
    const int Ng = g->Ng;
    for (long i = i0; i < i1; ++i)
        for (int j = 0; j < Ng; ++j)
            for (int k = (i == 0 && j == 0); k <= Ng / 2; ++k) {
                double *c = g->f + FIDX(g, i, j, 2 * k);
                double d2 = sq(c[0]) + sq(c[1]);
                c[0] = d2 * a->Vinv - a->Pshot;
                c[1] = 0;
            }
}

/* src/power.c:12-27 */
void oracle_P(offt_t *g, const ogal_t *part) {
    const int Ng = g->Ng;
    double Vinv = 1 / part->V;
    double alpha = part->rand == NULL ? 0 : (double)part->Np / part->rand->Np;
    double Pshot = (1 + alpha) * part->V / part->Np;
    g->f[FIDX(g, 0, 0, 0)] = g->f[FIDX(g, 0, 0, 1)] = 0;
    P_args a = {g, Vinv, Pshot};
    orc_parallel(Ng, P_planes, &a);
}

/* power.c:36,110 */
int oracle_NK(int Ng, double L, double dK) {
    double KF = 2 * M_PI / L;
    double dKinv = KF / dK;
    return (int)floor(ORC_SQRT3 * (Ng / 2) * dKinv) + 1;
}

/* signed integer frequency, Nyquist index carries -Ng/2 (power.c:46-51) */
static double *kval_table(int Ng) {
    double *kv = (double *)malloc((size_t)(Ng > 0 ? Ng : 1) * sizeof(double));
    assert(kv != NULL);
    kv[0] = 0;
    for (int i = 1; i <= Ng / 2; ++i) {
        kv[i] = i;
        kv[Ng - i] = -i;
    }
    return kv;
}

/* src/power.c:30-76 (accumulation + normalisation; no file) */
int oracle_Pl_bins(const offt_t *g, double los[3], double dK,
                   double *K, long *N, double *P) {
    const int Ng = g->Ng;
    oracle_hat(los);
    double KF = 2 * M_PI / g->L;
    double dKinv = KF / dK;
    int NK = (int)floor(ORC_SQRT3 * (Ng / 2) * dKinv) + 1;
    const int Nl = 4;
    for (int b = 0; b < NK; ++b) {
        K[b] = 0;
        N[b] = 0;
        for (int l = 0; l < Nl; ++l) P[b * Nl + l] = 0;
    }
    double *kv = kval_table(Ng);
    for (int i = 0; i < Ng; ++i)
        for (int j = 0; j < Ng; ++j)
            for (int k = (i == 0 && j == 0); k <= Ng / 2; ++k) {
                int count = 1 + (2 * k % Ng > 0);
                double d2 = g->f[FIDX(g, i, j, 2 * k)];
                d2 *= count;
                double Kamp = sqrt(sq(kv[i]) + sq(kv[j]) + sq(kv[k]));
                int iK = (int)floor(Kamp * dKinv);
                double mu2 = sq((kv[i] * los[0] + kv[j] * los[1] + kv[k] * los[2]) / Kamp);
                K[iK] += count * Kamp;
                N[iK] += count;
                P[iK * Nl + 0] += d2;
                P[iK * Nl + 1] += d2 * (1.5 * mu2 - 0.5);
                P[iK * Nl + 2] += d2 * ((4.375 * mu2 - 3.75) * mu2 + 0.375);
                P[iK * Nl + 3] += d2 * (((14.4375 * mu2 - 19.6875) * mu2 + 6.5625) * mu2 - 0.3125);
            }
    free(kv);
    long Ntot = 0;
    for (int b = 0; b < NK; ++b) {
        K[b] *= KF / N[b];
        for (int l = 0; l < Nl; ++l)
            P[b * Nl + l] *= (4 * l + 1.0) / N[b];
        Ntot += N[b];
    }
    assert(Ntot == (long)Ng * Ng * Ng - 1);
    return NK;
}

/* src/power.c:104-153 with Nmu as a parameter */
int oracle_Pmu_bins(const offt_t *g, double los[3], double dK, int Nmu,
                    double *Kout, long *Nout, double *Pout) {
    const int Ng = g->Ng;
    oracle_hat(los);
    double KF = 2 * M_PI / g->L;
    double dKinv = KF / dK;
    int NK = (int)floor(ORC_SQRT3 * (Ng / 2) * dKinv) + 1;
    const int W = Nmu + 1;
    double *K = (double *)calloc((size_t)NK * W, sizeof(double));
    double *P = (double *)calloc((size_t)NK * W, sizeof(double));
    long *N = (long *)calloc((size_t)NK * W, sizeof(long));
    assert(K && P && N);
    double *kv = kval_table(Ng);
    for (int i = 0; i < Ng; ++i)
        for (int j = 0; j < Ng; ++j)
            for (int k = (i == 0 && j == 0); k <= Ng / 2; ++k) {
                int count = 1 + (2 * k % Ng > 0);
                double d2 = g->f[FIDX(g, i, j, 2 * k)];
                d2 *= count;
                double Kamp = sqrt(sq(kv[i]) + sq(kv[j]) + sq(kv[k]));
                int iK = (int)floor(Kamp * dKinv);
                double mu = (kv[i] * los[0] + kv[j] * los[1] + kv[k] * los[2]) / Kamp;
                int imu = (int)floor(fabs(mu * Nmu));
                K[iK * W + imu] += count * Kamp;
                N[iK * W + imu] += count;
                P[iK * W + imu] += d2;
            }
    free(kv);
    for (int b = 0; b < NK; ++b) {     /* corner case mu = 1 (power.c:139-143) */
        K[b * W + Nmu - 1] += K[b * W + Nmu];
        N[b * W + Nmu - 1] += N[b * W + Nmu];
        P[b * W + Nmu - 1] += P[b * W + Nmu];
    }
    long Ntot = 0;
    for (int b = 0; b < NK; ++b)
        for (int m = 0; m < Nmu; ++m) {
            long n = N[b * W + m];
            if (n) {
                K[b * W + m] *= KF / n;
                P[b * W + m] /= n;
                Ntot += n;
            }
            Kout[b * Nmu + m] = K[b * W + m];
            Nout[b * Nmu + m] = n;
            Pout[b * Nmu + m] = P[b * W + m];
        }
    assert(Ntot == (long)Ng * Ng * Ng - 1);
    free(K); free(P); free(N);
    return NK;
}

/* src/app/DLsb.c:42-49,66-96 (kind 0) and src/app/DLss.c:18-20,62-89
 * (kind 1): Re of the DC mode to 0, every other mode times the window and
 * L_ell(mu) (ell = 0, 2, 4 is the reference's iL = 0, 1, 2), fft_k2x(g, 0),
 * then the sub-volume corners over bias */
void oracle_dl_filter(offt_t *g, int kind, int Nsub, const double los[3], int ell, double bias,
                      double *delta) {
    const int Ng = g->Ng, Ngsub = Ng / Nsub;
    double *Wamp = (double *)malloc((size_t)Ng * sizeof(double));
    double *Wph = (double *)malloc((size_t)3 * Ng * 2 * sizeof(double));
    double *Kval = (double *)malloc((size_t)Ng * sizeof(double));
    assert(Wamp && Wph && Kval);
    Wamp[0] = 1;
    for (int i = 1; i < Ng; ++i) Wamp[i] = sin(M_PI * i / Nsub) / sin(M_PI * i / Ng) / Ngsub;
    for (int i = 0; i < 3 * Ng; ++i) {
        Wph[2 * i] = cos(M_PI * i * (Ngsub - 1) / Ng);
        Wph[2 * i + 1] = sin(M_PI * i * (Ngsub - 1) / Ng);
    }
    Kval[0] = 0;
    for (int i = 1; i <= Ng / 2; ++i) {
        Kval[i] = i;
        Kval[Ng - i] = -i;
    }
    const int iL = ell / 2;
    g->f[0] = 0;
    for (int i = 0; i < Ng; ++i)
        for (int j = 0; j < Ng; ++j)
            for (int k = (i == 0 && j == 0); k <= Ng / 2; ++k) {
                double Kamp = sqrt(sq(Kval[i]) + sq(Kval[j]) + sq(Kval[k]));
                double mu2 = sq((Kval[i] * los[0] + Kval[j] * los[1] + Kval[k] * los[2]) / Kamp);
                double Legendre[3] = {1, 1.5 * mu2 - 0.5, (4.375 * mu2 - 3.75) * mu2 + 0.375};
                double *c = g->f + FIDX(g, i, j, 2 * k);
                if (kind == 0) {
                    double Wamp3 = Wamp[i] * Wamp[j] * Wamp[k];
                    double W_re = Wph[2 * (i + j + k)] * Wamp3;
                    double W_im = Wph[2 * (i + j + k) + 1] * Wamp3;
                    double f_re = c[0], f_im = c[1];
                    c[0] = Legendre[iL] * (f_re * W_re - f_im * W_im);
                    c[1] = Legendre[iL] * (f_re * W_im + f_im * W_re);
                } else {
                    double KR = Kamp * M_PI * ORC_SQRT3 / Nsub;
                    double W = 3 * (sin(KR) - KR * cos(KR)) / cube(KR);
                    c[0] *= Legendre[iL] * W;
                    c[1] *= Legendre[iL] * W;
                }
            }
    oracle_fft_k2x(g, 0);
    for (int a = 0; a < Nsub; ++a)
        for (int b = 0; b < Nsub; ++b)
            for (int cc = 0; cc < Nsub; ++cc)
                delta[(a * Nsub + b) * Nsub + cc] =
                    g->f[FIDX(g, a * Ngsub, b * Ngsub, cc * Ngsub)] / bias;
    free(Wamp);
    free(Wph);
    free(Kval);
}

/* src/app/Psb.c:73-85 (kind 0) and src/app/Pss.c:70-87 (kind 1) */
void oracle_subvolume_deconv(offt_t *g, int kind, int Nsub) {
    const int Ng = g->Ng, Ngsub = Ng / Nsub;
    oracle_fft_k2x(g, 0);
    if (kind == 0) {
        double Ngsub3 = (double)Ngsub * Ngsub * Ngsub;
        for (int i = -Ngsub + 1; i < Ngsub; ++i)
            for (int j = -Ngsub + 1; j < Ngsub; ++j)
                for (int k = -Ngsub + 1; k < Ngsub; ++k) {
                    double Winv = Ngsub3 / ((double)(Ngsub - abs(i)) * (Ngsub - abs(j)) *
                                            (Ngsub - abs(k)));
                    g->f[FIDX(g, (i + Ng) % Ng, (j + Ng) % Ng, (k + Ng) % Ng)] *= Winv;
                }
    } else {
        double D2 = 3.0 * Ngsub * Ngsub;
        int Ngbnd = (int)ceil(sqrt(D2));
        for (int i = -Ngbnd + 1; i < Ngbnd; ++i)
            for (int j = -Ngbnd + 1; j < Ngbnd; ++j)
                for (int k = -Ngbnd + 1; k < Ngbnd; ++k) {
                    double d2 = (double)i * i + (double)j * j + (double)k * k;
                    if (d2 < 0.99 * D2) {
                        double dD = sqrt(d2 / D2);
                        double Winv = 1 / ((1 - dD) * (1 - dD) * (1 + 0.5 * dD));
                        g->f[FIDX(g, (i + Ng) % Ng, (j + Ng) % Ng, (k + Ng) % Ng)] *= Winv;
                    }
                }
    }
    oracle_fft_x2k(g, 0);
}

/* src/power.c:78-99 */
void oracle_Pl(offt_t *g, const ogal_t *part, double los[3], double dK, const char *file) {
    int NK = oracle_NK(g->Ng, g->L, dK);
    double *K = (double *)malloc((size_t)NK * sizeof(double));
    long *N = (long *)malloc((size_t)NK * sizeof(long));
    double *P = (double *)malloc((size_t)NK * 4 * sizeof(double));
    assert(K && N && P);
    oracle_Pl_bins(g, los, dK, K, N, P);
    double alpha = part->rand == NULL ? 0 : (double)part->Np / part->rand->Np;
    FILE *fp = fopen(file, "w");
    assert(fp != NULL);
    fprintf(fp, "# NK %d\n# dK %g\n# Nl %d\n# Ng %d\n# L %g\n# Np %ld\n# V %g\n# alpha %g\n",
            NK, dK, 4, g->Ng, g->L, part->Np, part->V, alpha);
    fprintf(fp, "# los[3] %g %g %g\n# K N", los[0], los[1], los[2]);
    for (int l = 0; l < 4; ++l) fprintf(fp, " P%d", 2 * l);
    fprintf(fp, "\n");
    for (int b = 0; b < NK; ++b) {
        fprintf(fp, "%e %9ld", K[b], N[b]);
        for (int l = 0; l < 4; ++l) fprintf(fp, " % e", P[b * 4 + l]);
        fprintf(fp, "\n");
    }
    fclose(fp);
    free(K); free(N); free(P);
}

/* src/power.c:155-176 (Nmu = 3 as in power.c:111) */
void oracle_Pmu(offt_t *g, const ogal_t *part, double los[3], double dK, const char *file) {
    const int Nmu = 3;
    int NK = oracle_NK(g->Ng, g->L, dK);
    double *K = (double *)malloc((size_t)NK * Nmu * sizeof(double));
    long *N = (long *)malloc((size_t)NK * Nmu * sizeof(long));
    double *P = (double *)malloc((size_t)NK * Nmu * sizeof(double));
    assert(K && N && P);
    oracle_Pmu_bins(g, los, dK, Nmu, K, N, P);
    double alpha = part->rand == NULL ? 0 : (double)part->Np / part->rand->Np;
    FILE *fp = fopen(file, "w");
    assert(fp != NULL);
    fprintf(fp, "# NK %d\n# dK %g\n# Nmu %d\n# Ng %d\n# L %g\n# Np %ld\n# V %g\n# alpha %g\n",
            NK, dK, Nmu, g->Ng, g->L, part->Np, part->V, alpha);
    fprintf(fp, "# los[3] %g %g %g\n#", los[0], los[1], los[2]);
    for (int m = 0; m < Nmu; ++m) fprintf(fp, " K%d N%d P%d", m, m, m);
    fprintf(fp, "\n");
    for (int b = 0; b < NK; ++b) {
        fprintf(fp, "%e %9ld % e", K[b * Nmu], N[b * Nmu], P[b * Nmu]);
        for (int m = 1; m < Nmu; ++m)
            fprintf(fp, " %e %9ld % e", K[b * Nmu + m], N[b * Nmu + m], P[b * Nmu + m]);
        fprintf(fp, "\n");
    }
    fclose(fp);
    free(K); free(N); free(P);
}
