/* Catalog type on the host: src/gal.c:11-77 of eelregit/fftgal, written
 * for this library (plain C, no GSL). */
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
