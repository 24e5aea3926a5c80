/* Tuning parameters of the fast path and their cache file.
 *
 * The cache plays the role of the reference's FFTW wisdom file
 * (src/fft.c:26-45, named ${Ng}.wsdm in src/app/job_P.sh:13): fft_init's and
 * fftgal_power_desc's `wisdom` string names it, "FFTW_ESTIMATE" or NULL means
 * built-in defaults.  Text format, one line per configuration:
 *     Ng interlace paint_tile fft_cols fft_rows fft_generic bin_bk
 * (0 in fft_cols / fft_rows = automatic choice inside the FFT plan; the
 * last two fields may be absent: 0 = specialised FFT kernels where the
 * grid side has them, 0 = default bin block).  fftgal_power_autotune
 * (batch.c) measures the candidates and writes the line.  interlace = 2
 * marks the FP64 grid of the drop-in fft_init, whose line holds only the
 * FFT launch shape (fft_cols, fft_rows), measured by fft_init itself. */
#ifndef FFTGAL_TUNE_H
#define FFTGAL_TUNE_H

#define FFTGAL_TUNE_FP64_GRID 2

typedef struct fftgal_tuning {
    int paint_tile;
    int fft_cols;
    int fft_rows;
    int fft_generic;     /* 1: generic LDS Stockham kernels even where specialised exist */
    int bin_bk;          /* fused bin block k extent (0: default) */
} fftgal_tuning;

/* defaults, overridden by a matching line of the cache file if present;
 * returns 1 when the cache supplied the values */
int fftgal_tune_lookup(const char *path, int Ng, int interlace, fftgal_tuning *t);
/* append or replace the line for (Ng, interlace) */
int fftgal_tune_store(const char *path, int Ng, int interlace, const fftgal_tuning *t);

#endif
