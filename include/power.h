/* power spectrum estimators -- drop-in for eelregit/fftgal src/power.h.
 * Implemented by HIP binning kernels (FP64 sums, integer mode counts).
 */
#ifndef POWER_H
#define POWER_H

#include "fft.h"
#include "gal.h"

#ifdef __cplusplus
extern "C" {
#endif

/* P(k) = |delta(k)|^2 / V - Pshot, DC set to 0, imaginary parts to 0;
 * call this first (src/power.c:12-27) */
void P(fft_t *grid, gal_t *part);

/* power spectrum multipoles l = 0,2,4,6 into a text file
 * (src/power.c:30-101); normalises los in place */
void Pl(fft_t *grid, gal_t *part, double los[3], double dK, char *file);
/* power spectrum in 3 |mu| wedges into a text file (src/power.c:104-178) */
void Pmu(fft_t *grid, gal_t *part, double los[3], double dK, char *file);

#ifdef __cplusplus
}
#endif
#endif
