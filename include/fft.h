/* grid data type -- drop-in for eelregit/fftgal src/fft.h, backed by HIP
 * kernels on one MI355X (gfx950).
 *
 * fft_t keeps the reference's public fields with the same types and meaning
 * (src/fft.h:16-24) and the F/F_Re/F_Im macros (src/fft.h:10-13), so caller
 * code that reads or writes grid->f between calls compiles and behaves as
 * before.  The two fftw_plan slots become opaque pointers to the device
 * state (same size on LP64).
 *
 * Host/device contract: by default every call leaves the result valid in
 * the host array f (and uploads f first when the call reads it), exactly
 * the reference's semantics.  fft_set_device_resident() (fftgal_ext.h) or
 * FFTGAL_DEVICE_RESIDENT=1 keeps the grid on the device between calls;
 * then f is refreshed only by fft_sync_host() and fft_exportf().
 *
 * Precision: FP64 throughout (FP64 atomics in the paint, FP64 FFT), like
 * the reference.  The mixed-precision fast path is fftgal_batch.h.
 * Limits: Ng even with prime factors in {2,3,5,7,11,13}.  Failures print a
 * message on stderr and abort(), as the reference's assert()s do.
 */
#ifndef FFT_H
#define FFT_H

#include "gal.h"

#ifdef __cplusplus
extern "C" {
#endif

/* field value in configuration space */
#define F(self,i,j,k) (self->f[((long)(i) * self->Ng + (j)) * self->Ng_pad + (k)])
/* field value in Fourier space */
#define F_Re(self,i,j,k) F(self,i,j,2*k)
#define F_Im(self,i,j,k) F(self,i,j,2*k+1)

typedef struct fft {
    void *x2k, *k2x;   /* opaque device state (both point to it) */
    double *f;         /* host grid; in place: fk = (complex *)fx */
    int Ng;
    int Ng_pad;        /* 2 * (Ng/2 + 1) */
    long Ng3_pad;      /* sizeof(f) / sizeof(double) */
    double L;          /* [Mpc/h] */
    double offset[3];  /* grid wrt boundary [0,L)^3 in grid unit */
} fft_t;

/* src/fft.c:13-47.  Ng: any side in [1, 1024], or a larger even side with
 * prime factors in {2,3,5,7,11,13}.  wisdom: "FFTW_ESTIMATE" (default launch
 * shapes) or a tuning cache path: a line for this Ng is read if the file
 * holds one; otherwise the FP64 transform's launch shapes are timed on the
 * new grid (as FFTW_MEASURE) and the fastest is written to the file, and an
 * unwritable path aborts, as the reference's failed wisdom export does */
fft_t *fft_init(int Ng, double L, char *wisdom);
/* src/fft.c:50-54 */
void fft_free(fft_t *self);

/* PCS paint, src/fft.c:57-131; reduces offset[] with remainder(., Ng) in
 * place and stores it in self->offset; randoms (part->rand) with -alpha */
void fft_p2g(fft_t *self, gal_t *part, double offset[3]);
/* src/fft.c:134-169, reproduced as is on the host (it assigns each stencil
 * term instead of accumulating, so each result is the last term) */
double *fft_g2p(fft_t *self, gal_t *pos);

/* forward r2c times (L/Ng)^3, optional offset phase (src/fft.c:172-181) */
void fft_x2k(fft_t *self, int offset_phase_on);
/* inverse c2r times 1/L^3, optional phase removal applied after the
 * transform, as in src/fft.c:184-193 */
void fft_k2x(fft_t *self, int offset_phase_off);
/* apply (onoff>0) or remove (onoff<0) the offset phase (src/fft.c:196-233) */
void fft_offset_phase(fft_t *self, int onoff);

/* PCS window deconvolution (src/fft.c:236-254) */
void fft_deconv(fft_t *self);

/* interlacing = (paint + FFT) x 2 + average + deconvolve (src/fft.c:257-276) */
void fft_p2k(fft_t *self, gal_t *part);

/* malloc'd copy of the grid, released by the caller with free()
 * (src/fft.c:279-283) */
double *fft_exportf(fft_t *self);
/* copy a grid in (src/fft.c:286-289) */
void fft_importf(fft_t *self, double *src);

#ifdef __cplusplus
}
#endif
#endif
