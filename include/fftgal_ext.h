/* Additive entry points next to the drop-in API (fft.h, power.h).
 *
 * None of these exists in eelregit/fftgal; they expose what the GPU backing
 * makes possible without changing any reference signature:
 *  - keeping the grid on the device between calls (no PCIe traffic for
 *    callers such as src/app/P.c that never touch grid->f);
 *  - painting from catalogs already in device memory;
 *  - Pl/Pmu results as arrays instead of text files, with any number of
 *    mu wedges (src/power.c:111 fixes Nmu = 3).
 */
#ifndef FFTGAL_EXT_H
#define FFTGAL_EXT_H

#include "fft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* on != 0: the device grid is authoritative between calls; host f is only
 * refreshed by fft_sync_host() / fft_exportf().  Switching off copies the
 * device grid to f first. */
void fft_set_device_resident(fft_t *self, int on);
/* copy the device grid to host f */
void fft_sync_host(fft_t *self);
/* device address of the FP64 grid (same padded layout as f) */
double *fft_device_grid(fft_t *self);

/* fft_p2g / fft_p2k for a catalog of Np FP64 positions in device memory
 * (no randoms); V is the catalog volume as in gal_t.V */
void fft_p2g_device(fft_t *self, const double *x, const double *y, const double *z,
                    long Np, double V, double offset[3]);
void fft_p2k_device(fft_t *self, const double *x, const double *y, const double *z,
                    long Np, double V);
/* P() without a gal_t: alpha = Np/Nrand (Nrand = 0: no randoms) */
void P_params(fft_t *grid, long Np, double V, long Nrand);

/* number of k bins of Pl/Pmu (src/power.c:36) */
int fftgal_NK(int Ng, double L, double dK);
/* Pl as arrays: K[NK], N[NK], P[NK*4] (l = 0,2,4,6); returns NK */
int Pl_bins(fft_t *grid, double los[3], double dK, double *K, long *N, double *P);
/* No counterpart in the reference.  Pl_bins for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): the same results, with the device bin sums held
 * in a fixed 1 MiB scratch array that each GPU keeps for the life of the
 * process, and no device or host allocation.  Limit: NK <= 21840 bins
 * (fftgal_NK; e.g. Ng = 2048 down to dK = kF/12).  Returns NK, or -1 with
 * fftgal_last_error() set (too many bins, a failed device call, a count
 * mismatch) instead of aborting.  Every call on a GPU shares that GPU's
 * scratch: calls that overlap on one GPU (two threads, two grids) race. */
int Pl_bins_noalloc(fft_t *grid, double los[3], double dK, double *K, long *N, double *P);
/* Pmu with Nmu wedges as arrays: K, N, P each [NK*Nmu]; returns NK */
int Pmu_bins(fft_t *grid, double los[3], double dK, int Nmu, double *K, long *N, double *P);

/* Super-sample modes of sub-volumes (the k-space loop of src/app/DLsb.c:66-96
 * and DLss.c:62-89 on the device).  On the grid's half spectrum from
 * fft_x2k (restored e.g. by fft_importf): Re of the DC mode set to 0, every
 * other mode times W(k) L_ell(mu), with W the sub-box window with its phase
 * (kind 0, DLsb) or the spherical top hat of K R = |n| pi sqrt(3) / Nsub
 * (kind 1, DLss); then fft_k2x(grid, 0), and the Nsub^3 corners
 * f(a Ng/Nsub, b Ng/Nsub, c Ng/Nsub) / bias into delta[(a*Nsub + b)*Nsub + c].
 * los as the reference passes it (already normalised by hat()); ell = 0, 2
 * or 4; Nsub divides Ng.  grid->f holds the filtered configuration-space
 * grid afterwards, as after fft_k2x. */
void fft_dl_filter(fft_t *grid, int kind, int Nsub, const double los[3], int ell, double bias,
                   double *delta);

/* Window deconvolution of a sub-volume power spectrum (the loops of
 * src/app/Psb.c:73-85 and Pss.c:70-87 on the device).  The grid holds P(k)
 * from P() of a sub-volume catalog (gal_subbox / gal_subsphere); it is
 * transformed back (fft_k2x(grid, 0)), the correlation function at lag
 * (i, j, k) cells is divided by the window's autocorrelation -- kind 0, a
 * sub box of gs = Ng/Nsub cells: prod(gs - |lag|) / gs^3 for |lags| < gs;
 * kind 1, a sub sphere of diameter D = sqrt(3) gs cells:
 * (1 - d)^2 (1 + d/2), d = |lag| / D, for |lag|^2 < 0.99 D^2 -- and
 * transformed forward (fft_x2k(grid, 0)), ready for Pl / Pmu. */
void fft_subvolume_deconv(fft_t *grid, int kind, int Nsub);

#ifdef __cplusplus
}
#endif
#endif
