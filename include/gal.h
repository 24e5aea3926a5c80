/* galaxies/particles data type -- drop-in for eelregit/fftgal src/gal.h.
 *
 * gal_t keeps the reference's fields, types and meaning (src/gal.h:29-39)
 * so that callers compile unchanged.  Implemented on the host in
 * fftgal_amd/csrc/host/gal.c.
 *
 * Differences from the reference header:
 *  - pdist2() and hat() are `static inline` (src/gal.h:10,19 declare plain
 *    `inline`, which links only when the compiler inlines every call);
 *  - gal_subbox()/gal_subsphere() draw their randoms from the generator
 *    the reference gets from GSL (gsl_rng_mt19937 set up from the
 *    environment: GSL_RNG_SEED, default seed 0 -> 4357), written in this
 *    library, so that the same catalogs come out without linking GSL.
 */
#ifndef GAL_H
#define GAL_H

#include <math.h>

#ifdef __cplusplus
extern "C" {
#endif

/* squared distance with periodic boundary (src/gal.h:10-16) */
static inline double pdist2(double x1, double y1, double z1,
                            double x0, double y0, double z0, double L) {
    double dx = remainder(x1 - x0, L);
    double dy = remainder(y1 - y0, L);
    double dz = remainder(z1 - z0, L);
    return dx * dx + dy * dy + dz * dz;
}

/* normalise a vector in place if its length exceeds 1e-7 (src/gal.h:19-26) */
static inline void hat(double vec[3]) {
    double amp = sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
    if (amp > 1e-7) {
        vec[0] /= amp;
        vec[1] /= amp;
        vec[2] /= amp;
    }
}

typedef struct gal {
    long Np;
    double V;              /* [Mpc^3/h^3] */
    double *x, *y, *z;     /* [Mpc/h] */
    double *vx, *vy, *vz;  /* [km/s], optional */
    struct gal *rand;      /* random catalog, optional */
} gal_t;

/* src/gal.h:42-44, src/gal.c:11-48 */
gal_t *gal_init(long Np, double V);
void gal_resize(gal_t *self, long Np);
void gal_free(gal_t *self);

/* move particles periodically back into [0,L)^3 (src/gal.c:51-60) */
void gal_wrap(gal_t *self, double L);

/* plane-parallel redshift-space shift along los, new catalog without
 * velocities or randoms (src/gal.c:63-77); normalises los in place */
gal_t *gal_rsd(gal_t *self, double los[3], double aH);

/* particles inside the periodic sub-box box = {x0, x1, y0, y1, z0, z1, L},
 * V = its volume, with lround(Np_sub / alpha) uniform randoms in it as
 * ->rand (src/gal.h:46-49, src/gal.c:78-118) */
gal_t *gal_subbox(gal_t *self, double box[7], double alpha);
/* the same for the periodic sub-sphere sphere = {x0, y0, z0, R, L}
 * (src/gal.h:50-53, src/gal.c:121-158) */
gal_t *gal_subsphere(gal_t *self, double sphere[5], double alpha);

#ifdef __cplusplus
}
#endif
#endif
