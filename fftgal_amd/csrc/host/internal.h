/* Host-side internals shared by fft.c, power.c and batch.c. */
#ifndef FFTGAL_INTERNAL_H
#define FFTGAL_INTERNAL_H

#include <stdio.h>
#include <stdlib.h>
#include "../fgk.h"
#include "../../../include/fft.h"

/* abort with the library's last error, the drop-in's failure mode (the
 * reference asserts) */
#define DI_CHECK(call)                                                        \
    do {                                                                      \
        if ((call) != 0) {                                                    \
            fprintf(stderr, "fftgal: %s failed: %s\n", #call, fgk_last_error()); \
            abort();                                                          \
        }                                                                     \
    } while (0)

#define DI_ASSERT(cond, msg)                                                  \
    do {                                                                      \
        if (!(cond)) {                                                        \
            fprintf(stderr, "fftgal: %s\n", msg);                             \
            abort();                                                          \
        }                                                                     \
    } while (0)

/* device state behind fft_t.x2k */
typedef struct fft_dev {
    fgk_stream_t s;
    fgk_fftplan *plan;     /* FP64 */
    double *d_f;           /* device grid, same layout as f */
    double *d_dual;        /* interlacing copy */
    double *d_phase;       /* 3 * Ng complex */
    double *d_winv;        /* Ng */
    double *d_pos;         /* positions staging, 3 * pos_cap doubles */
    long pos_cap;
    int resident;
} fft_dev;

static inline fft_dev *fft_dev_of(fft_t *g) { return (fft_dev *)g->x2k; }

/* host f -> device when host is authoritative */
void fftgal_dev_pull(fft_t *g);
/* device -> host f when host is authoritative */
void fftgal_dev_push(fft_t *g);

int fftgal_verbose(void);
double fftgal_now(void);

/* PCS window deconvolution factors W^-1[i] (src/fft.c:239-245) */
void fftgal_winv_table(int Ng, double *winv);
/* offset phase table phase[d][j] (src/fft.c:204-216) */
void fftgal_phase_table(int Ng, const double offset[3], int onoff, double *phase);
/* NK of src/power.c:36 */
int fftgal_nk(int Ng, double L, double dK);

#endif
