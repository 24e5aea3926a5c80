/*
 * FP64 CPU oracle for the fftgal power-spectrum path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU
 * library (libfftgal_amd.so) is compared against; only tests, the smoke
 * check and the CPU-baseline leg of the benchmark load it.  No product code
 * links or calls it.
 *
 * It restates the reference's algorithm (eelregit/fftgal, src/fft.c,
 * src/power.c, src/gal.c, src/gal.h) in plain C with the same evaluation
 * order, and replaces FFTW (absent here) with its own FP64 mixed-radix FFT
 * (fft64.c).  Every function cites the reference lines it follows.
 *
 * Pinning (what checks each function independently of this restatement):
 *   oracle_fft_p2g      octet known answer (src/tests/octet.c:18-31), closed-form
 *                       PCS weights, partition of unity / mass, cyclic-shift
 *                       invariance, brute force on tiny grids
 *   oracle_fft_x2k/k2x  octet known answer (octet.c:33-56), numpy.fft.rfftn /
 *                       irfftn golden vectors (tests/golden), brute-force DFT,
 *                       Parseval, round trip
 *   oracle_fft_offset_phase  octet (Nyquist convention at Ng=2), closed form,
 *                       +1/-1 round trip, equivalence with a sub-cell shift
 *   oracle_fft_deconv   closed form (sin(x)/x)^-4
 *   oracle_P / _Pl / _Pmu  closed forms, brute-force mode counts, sum N = Ng^3-1,
 *                       planted isotropic and planted-Legendre fields
 *   oracle_gal_rsd, oracle_hat  closed forms
 *   oracle_dl_filter    the sub-box window against its defining sum (the
 *                       mean of e^{ik.x} over a sub box, brute force), the
 *                       top hat against its series at small KR, and
 *                       uniform-field and Legendre-plane closed forms
 *   oracle_gal_subbox/_subsphere  selection against a numpy mask, randoms
 *                       against numpy's MT19937 (legacy seeding) raw outputs
 *   oracle_subvolume_deconv  window autocorrelations against brute-force
 *                       overlap counts of a digitised box / quadrature of the
 *                       sphere lens volume
 *   oracle_loadqpm_cubic  every value equals Python's float() of its text
 *                       field (correctly rounded), line count = '\n' count
 *                       (a last line without '\n' is not read), velocities
 */
#ifndef FFTGAL_ORACLE_H
#define FFTGAL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* own FP64 FFT plan for one length (fft64.c) */
typedef struct offt64_plan {
    int n;          /* transform length */
    int nstage;
    int radix[64];
    double *tw;     /* 2n doubles: exp(-2 pi i t / n), t = 0..n-1 */
} offt64_plan;

/* grid, same field meaning as fft_t (src/fft.h:16-24) */
typedef struct offt {
    double *f;
    int Ng;
    int Ng_pad;        /* 2 * (Ng/2 + 1) */
    long Ng3_pad;
    double L;
    double offset[3];
    offt64_plan *pN;   /* length Ng complex plan */
    offt64_plan *pH;   /* length Ng/2 complex plan (real<->half-complex rows) */
} offt_t;

/* catalog, same field meaning as gal_t (src/gal.h:29-39) */
typedef struct ogal {
    long Np;
    double V;
    double *x, *y, *z;
    double *vx, *vy, *vz;
    struct ogal *rand;
} ogal_t;

/* ---- threads (fft64.c) ----
 * The reference is single-threaded.  For large test workloads the oracle can
 * split work over threads without changing a single bit of its results:
 * threads own disjoint output elements (paint: slabs of i-planes, each
 * filled in the serial particle order; FFT: whole 1D transforms; k-space
 * sweeps: i-planes), and the bin sums stay serial.  Default 1. */
void oracle_set_threads(int n);
int  oracle_threads(void);
/* fn(arg, lo, hi) over contiguous ranges covering [0, n), one per thread */
void orc_parallel(long n, void (*fn)(void *arg, long lo, long hi), void *arg);

/* ---- FP64 FFT (fft64.c) ---- */
int  offt64_plan_init(offt64_plan *p, int n);
void offt64_plan_free(offt64_plan *p);
/* complex FFT of x (2n doubles, interleaved), sign -1 forward, +1 inverse,
 * unnormalised; work holds 2n doubles */
void offt64_exec(const offt64_plan *p, double *x, double *work, int sign);
/* in-place 3D real<->complex on the padded layout f[(i*Ng+j)*Ng_pad+k] */
void offt64_r2c_3d(offt_t *g);
void offt64_c2r_3d(offt_t *g);

/* ---- grid API (src/fft.c) ---- */
offt_t *oracle_fft_init(int Ng, double L);
void oracle_fft_free(offt_t *self);
void oracle_fft_p2g(offt_t *self, const ogal_t *part, double offset[3]);
void oracle_fft_x2k(offt_t *self, int offset_phase_on);
void oracle_fft_k2x(offt_t *self, int offset_phase_off);
void oracle_fft_offset_phase(offt_t *self, int onoff);
void oracle_fft_deconv(offt_t *self);
void oracle_fft_p2k(offt_t *self, const ogal_t *part);
double *oracle_fft_exportf(const offt_t *self);
void oracle_fft_importf(offt_t *self, const double *src);

/* ---- estimators (src/power.c) ---- */
void oracle_P(offt_t *grid, const ogal_t *part);
/* number of k bins used by Pl/Pmu (power.c:36,110) */
int oracle_NK(int Ng, double L, double dK);
/* Pl bins: K[NK], N[NK], P[NK*4] (l = 0,2,4,6); returns NK */
int oracle_Pl_bins(const offt_t *grid, double los[3], double dK,
                   double *K, long *N, double *P);
/* Pmu bins with Nmu wedges (the reference fixes Nmu = 3): K,N,P[NK*Nmu] */
int oracle_Pmu_bins(const offt_t *grid, double los[3], double dK, int Nmu,
                    double *K, long *N, double *P);
/* text writers with the reference's format (power.c:79-99, 155-176) */
void oracle_Pl(offt_t *grid, const ogal_t *part, double los[3], double dK, const char *file);
void oracle_Pmu(offt_t *grid, const ogal_t *part, double los[3], double dK, const char *file);

/* ---- catalog (src/gal.c, src/gal.h) ---- */
void oracle_hat(double vec[3]);
ogal_t *oracle_gal_init(long Np, double V);
void oracle_gal_free(ogal_t *self);
void oracle_gal_wrap(ogal_t *self, double L);
ogal_t *oracle_gal_rsd(const ogal_t *self, double los[3], double aH);
ogal_t *oracle_loadqpm_cubic(const char *file, double L);
void oracle_dl_filter(offt_t *g, int kind, int Nsub, const double los[3], int ell, double bias,
                      double *delta);
ogal_t *oracle_gal_subbox(const ogal_t *g, double box[7], double alpha);
ogal_t *oracle_gal_subsphere(const ogal_t *g, double sphere[5], double alpha);
void oracle_subvolume_deconv(offt_t *g, int kind, int Nsub);

#ifdef __cplusplus
}
#endif
#endif
