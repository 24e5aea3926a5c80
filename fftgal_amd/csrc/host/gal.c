/* Catalog type on the host: src/gal.c:11-158 of eelregit/fftgal, written
 * for this library (plain C, no GSL: the sub-volume randoms use MT19937
 * written below, the generator GSL's default is). */
#include <stdlib.h>
#include <stdio.h>
#include <math.h>
#include "../../../include/gal.h"
#include "internal.h"

/* src/gal.c:11-21 */
gal_t *gal_init(long Np, double V) {
    gal_t *self = (gal_t *)malloc(sizeof(gal_t));
    DI_ASSERT(self != NULL, "gal_init: out of memory");
    self->Np = Np;
    self->V = V;
    size_t n = Np > 0 ? (size_t)Np : 1;
    self->x = (double *)malloc(n * sizeof(double));
    self->y = (double *)malloc(n * sizeof(double));
    self->z = (double *)malloc(n * sizeof(double));
    DI_ASSERT(self->x && self->y && self->z, "gal_init: out of memory");
    self->vx = self->vy = self->vz = NULL;
    self->rand = NULL;
    return self;
}

static double *grow(double *p, long Np) {
    double *q = (double *)realloc(p, (Np > 0 ? (size_t)Np : 1) * sizeof(double));
    DI_ASSERT(q != NULL, "gal_resize: out of memory");
    return q;
}

/* src/gal.c:24-38 */
void gal_resize(gal_t *self, long Np) {
    self->Np = Np;
    self->x = grow(self->x, Np);
    self->y = grow(self->y, Np);
    self->z = grow(self->z, Np);
    if (self->vx) self->vx = grow(self->vx, Np);
    if (self->vy) self->vy = grow(self->vy, Np);
    if (self->vz) self->vz = grow(self->vz, Np);
}

/* src/gal.c:41-48: frees the randoms too */
void gal_free(gal_t *self) {
    if (self != NULL) {
        gal_free(self->rand);
        free(self->x);
        free(self->y);
        free(self->z);
        free(self->vx);
        free(self->vy);
        free(self->vz);
    }
    free(self);
}

/* src/gal.c:51-60 */
void gal_wrap(gal_t *self, double L) {
    double t = fftgal_now();
    double Linv = 1 / L;
    for (long p = 0; p < self->Np; ++p) {
        self->x[p] -= L * floor(self->x[p] * Linv);
        self->y[p] -= L * floor(self->y[p] * Linv);
        self->z[p] -= L * floor(self->z[p] * Linv);
    }
    if (fftgal_verbose()) fprintf(stderr, "gal_wrap() %.2fs\n", fftgal_now() - t);
}

/* src/gal.c:63-77 */
gal_t *gal_rsd(gal_t *self, double los[3], double aH) {
    double t = fftgal_now();
    hat(los);
    DI_ASSERT(self->vx && self->vy && self->vz, "gal_rsd: catalog has no velocities");
    double aHinv = 1 / aH;
    gal_t *out = gal_init(self->Np, self->V);
    for (long p = 0; p < out->Np; ++p) {
        double vl = los[0] * self->vx[p] + los[1] * self->vy[p] + los[2] * self->vz[p];
        out->x[p] = self->x[p] + aHinv * vl * los[0];
        out->y[p] = self->y[p] + aHinv * vl * los[1];
        out->z[p] = self->z[p] + aHinv * vl * los[2];
    }
    if (fftgal_verbose())
        fprintf(stderr, "gal_rsd() %.2fs, los[]={%.3f,%.3f,%.3f}\n", fftgal_now() - t, los[0],
                los[1], los[2]);
    return out;
}

/* ---- sub volumes with randoms (src/gal.c:78-158) ----
 *
 * The randoms come from the reference's generator: GSL's default,
 * gsl_rng_mt19937, set up by gsl_rng_env_setup (seed from GSL_RNG_SEED,
 * default 0, which MT19937 replaces by 4357), one generator per function,
 * kept for the life of the process, and gsl_rng_uniform = 32-bit output /
 * 2^32.  The generator is Matsumoto and Nishimura's MT19937 (ACM TOMACS 8,
 * 3, 1998) with its 2002 seeding, written here from that description. */

typedef struct mt19937 {
    uint32_t mt[624];
    int mti;
} mt19937;

static void mt_seed(mt19937 *g, unsigned long s) {
    if (s == 0) s = 4357;
    g->mt[0] = (uint32_t)s;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = (uint32_t)(1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i);
    g->mti = 624;
}

static uint32_t mt_next(mt19937 *g) {
    if (g->mti >= 624) {   /* regenerate the 624 words */
        for (int i = 0; i < 624; ++i) {
            const uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->mti = 0;
    }
    uint32_t y = g->mt[g->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

static double mt_uniform(mt19937 *g) { return mt_next(g) / 4294967296.0; }

/* the seed gsl_rng_env_setup would take */
static unsigned long env_seed(void) {
    const char *e = getenv("GSL_RNG_SEED");
    return e ? strtoul(e, NULL, 0) : 0;
}

/* src/gal.c:78-118 */
gal_t *gal_subbox(gal_t *self, double box[7], double alpha) {
    static mt19937 rng;
    static int seeded = 0;
    double t = fftgal_now();
    double x0 = box[0], x1 = box[1];
    double y0 = box[2], y1 = box[3];
    double z0 = box[4], z1 = box[5], L = box[6];
    x1 = x0 + remainder(x1 - x0, L);
    y1 = y0 + remainder(y1 - y0, L);
    z1 = z0 + remainder(z1 - z0, L);
    DI_ASSERT(x1 > x0 && y1 > y0 && z1 > z0, "gal_subbox: empty box");
    double Vsub = (x1 - x0) * (y1 - y0) * (z1 - z0);
    gal_t *sub = gal_init(self->Np, Vsub);
    long Npsub = 0;
    for (long p = 0; p < self->Np; ++p)
        if (remainder(self->x[p] - x0, L) >= 0 && remainder(self->x[p] - x1, L) < 0 &&
            remainder(self->y[p] - y0, L) >= 0 && remainder(self->y[p] - y1, L) < 0 &&
            remainder(self->z[p] - z0, L) >= 0 && remainder(self->z[p] - z1, L) < 0) {
            sub->x[Npsub] = self->x[p];
            sub->y[Npsub] = self->y[p];
            sub->z[Npsub] = self->z[p];
            ++Npsub;
        }
    gal_resize(sub, Npsub);
    long Nprand = lround(Npsub / alpha);
    gal_t *rand = gal_init(Nprand, Vsub);
    if (!seeded) {
        mt_seed(&rng, env_seed());
        seeded = 1;
    }
    for (long p = 0; p < Nprand; ++p) {
        rand->x[p] = x0 + (x1 - x0) * mt_uniform(&rng);
        rand->y[p] = y0 + (y1 - y0) * mt_uniform(&rng);
        rand->z[p] = z0 + (z1 - z0) * mt_uniform(&rng);
    }
    sub->rand = rand;
    if (fftgal_verbose())
        fprintf(stderr, "gal_subbox() %.2fs, picked out %ld+%ld particles\n", fftgal_now() - t,
                Npsub, Nprand);
    return sub;
}

/* src/gal.c:121-158 */
gal_t *gal_subsphere(gal_t *self, double sphere[5], double alpha) {
    static mt19937 rng;
    static int seeded = 0;
    double t = fftgal_now();
    double x0 = sphere[0], y0 = sphere[1], z0 = sphere[2];
    double R = sphere[3], L = sphere[4];
    DI_ASSERT(2 * R <= L, "gal_subsphere: the sphere must fit in the box");
    double R2 = R * R, Vsub = 4 * M_PI / 3 * R2 * R;
    gal_t *sub = gal_init(self->Np, Vsub);
    long Npsub = 0;
    for (long p = 0; p < self->Np; ++p)
        if (pdist2(self->x[p], self->y[p], self->z[p], x0, y0, z0, L) < R2) {
            sub->x[Npsub] = self->x[p];
            sub->y[Npsub] = self->y[p];
            sub->z[Npsub] = self->z[p];
            ++Npsub;
        }
    gal_resize(sub, Npsub);
    long Nprand = lround(Npsub / alpha);
    gal_t *rand = gal_init(Nprand, Vsub);
    if (!seeded) {
        mt_seed(&rng, env_seed());
        seeded = 1;
    }
    for (long p = 0; p < Nprand; ++p) {
        do {
            rand->x[p] = x0 + R * (2 * mt_uniform(&rng) - 1);
            rand->y[p] = y0 + R * (2 * mt_uniform(&rng) - 1);
            rand->z[p] = z0 + R * (2 * mt_uniform(&rng) - 1);
        } while (pdist2(rand->x[p], rand->y[p], rand->z[p], x0, y0, z0, L) >= R2);
    }
    sub->rand = rand;
    if (fftgal_verbose())
        fprintf(stderr, "gal_subsphere() %.2fs, picked out %ld+%ld particles\n",
                fftgal_now() - t, Npsub, Nprand);
    return sub;
}
