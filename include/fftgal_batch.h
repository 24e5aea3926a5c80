/* Mixed-precision power-spectrum pipeline on one MI355X (gfx950).
 *
 * Additive API (no counterpart in eelregit/fftgal).  It computes, for
 * catalogs of float32 positions already in device memory, what the
 * reference computes with
 *     fft_p2k (src/fft.c:257-276) or fft_p2g + fft_x2k + fft_deconv,
 *     P (src/power.c:12-27), Pl (src/power.c:30-101), Pmu (power.c:104-178)
 * but as three stages on the device:
 *   paint  -- tile-sorted PCS scatter of the plain and half-cell dual grids
 *             in one pass, int64 fixed-point accumulation (bit-identical
 *             run to run);
 *   FFT    -- FP32 Stockham r2c (rows) + 2 strided complex passes, twiddles
 *             generated in FP64;
 *   bin    -- one fused kernel: dual-grid offset phase, interlace average,
 *             PCS deconvolution, |delta|^2/V - Pshot in FP64, Legendre
 *             l = 0,2,4,6 and |mu|-wedge binning with FP64 sums.
 * The stages of one enqueue are captured once in a hipGraph and replayed.
 * Mode counts equal the reference's exactly; P_l agrees with the FP64 path
 * within the FP32 FFT error.
 *
 * All functions return 0 on success and a negative code on failure;
 * fftgal_last_error() describes the failure.
 */
#ifndef FFTGAL_BATCH_H
#define FFTGAL_BATCH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fftgal_power fftgal_power_t;

typedef struct fftgal_power_desc {
    int Ng;              /* grid cells per side: even, prime factors <= 13 */
    double L;            /* grid box side [Mpc/h] = catalog box / fold */
    int interlace;       /* 1: plain + dual grid (fft_p2k); 0: plain grid */
    double dK;           /* k-bin width [h/Mpc] */
    double los[3];       /* line of sight (normalised internally) */
    int Nmu;             /* |mu| wedges (0: none) */
    long Np_max;         /* largest catalog accepted by enqueue */
    int paint_tile;      /* paint tile side in cells (0: tuned / default) */
    const char *wisdom;  /* tuning cache file (NULL or "FFTW_ESTIMATE": defaults) */
} fftgal_power_desc;

const char *fftgal_last_error(void);

int fftgal_power_create(fftgal_power_t **out, const fftgal_power_desc *desc);
int fftgal_power_destroy(fftgal_power_t *p);
/* number of k bins (reference NK, src/power.c:36) */
int fftgal_power_NK(const fftgal_power_t *p);

/* Enqueue one catalog: Np float32 positions x, y, z in device memory
 * (any real coordinates; the paint wraps them periodically into the grid
 * box), V = catalog volume (nbar = Np/V, Pshot = V/Np).  Asynchronous on
 * the plan's stream. */
int fftgal_power_enqueue(fftgal_power_t *p, const float *x, const float *y, const float *z,
                         long Np, double V);
/* wait for the last enqueue */
int fftgal_power_sync(fftgal_power_t *p);
/* normalised results of the last enqueue (after sync):
 * K[NK], N[NK], Pl[NK*4] as Pl(); Kmu, Nmu_count, Pmu [NK*Nmu] as Pmu().
 * Any pointer may be NULL. */
int fftgal_power_results(fftgal_power_t *p, double *K, long *N, double *Pl, double *Kmu,
                         long *Nmu_count, double *Pmu);

/* ---- stage access for profiling and tests ---- */
/* device address of the int64 grids (plain, dual) after the paint stage,
 * layout [Ng][Ng][Ng], in units of fftgal_power_fixed_scale() particles */
int fftgal_power_paint_only(fftgal_power_t *p, const float *x, const float *y, const float *z,
                            long Np, void **gridA, void **gridB);
/* Fixed point of the paint.  Each paint tile rounds its particles'
 * weights to 1 / (216 C_S) particle, C_S = 2427 * 2^(S-19), with S (19 to
 * 25) the most fraction bits its particle count allows without int32
 * overflow (fftgal_power_fixed_bits(n) for work items of n particles; 19
 * for 8192 or more, 25 for at most 216), with a deterministic dither that
 * makes every rounding unbiased.  The int64 grids hold units of
 * fftgal_power_fixed_scale() = 1 / (216 C_25) particle (within 1.1e-4 of
 * 2^-25); coarser tiles are shifted to it exactly. */
int fftgal_power_fixed_bits(long n);
double fftgal_power_fixed_scale(void);
/* int64 grid sets of the plan: 2 when device memory allowed a second set at
 * creation (catalogs alternate between the sets and each paint clears the
 * other one, so the row FFT does not clear the grid it reads), else 1.
 * FFTGAL_GRID_SETS=1 in the environment at creation keeps one set. */
int fftgal_power_grid_sets(const fftgal_power_t *p);
/* device address of the FP32 spectra (plain, dual) after an enqueue, in
 * the packed half-spectrum layout c[(i*Ng+j)*(Ng/2) + k] (complex float,
 * k < Ng/2, unnormalised forward transform of the int64 fixed-point grid
 * in its own units, fftgal_power_fixed_scale() particles each): column
 * k = 0 holds Z = F(X0) + i F(XN), from which
 * delta(i,j,0) = (Z[i,j] + conj Z[-i,-j]) / 2 and
 * delta(i,j,Ng/2) = (Z[i,j] - conj Z[-i,-j]) / 2i.  Grids of Ng >= 2048
 * (Ng a multiple of 64) run their i-axis pass as a four-step 64 x Ng/64,
 * which leaves frequency i on i-plane (i % 64) * (Ng / 64) + i / 64
 * (FFTGAL_FFT_FOURSTEP=n1 at plan creation sets the split n1, 0 keeps the
 * natural order). */
int fftgal_power_spectra(fftgal_power_t *p, void **A, void **B);
/* tuning in use: paint tile side, generic FFT kernels' columns / rows per
 * workgroup (0: automatic), 1 if the generic FFT kernels run, k extent of
 * the fused bin kernel's block of modes */
int fftgal_power_tuning(const fftgal_power_t *p, int *paint_tile, int *fft_cols, int *fft_rows,
                        int *fft_generic, int *bin_bk);
/* Autotuner: times, on the caller's catalog (device float32 positions, as
 * enqueue takes them), the paint tile sides 16, 12, 8, 4, the FFT kernels
 * (compile-time specialised where the grid side has them, generic LDS
 * Stockham with its default, halved-column and halved-row launch shapes)
 * and the fused bin block (k extent 16, 32, 64, 128), one knob at a time in
 * pipeline order, the fastest of `reps` eager runs per candidate.  Stores
 * the winner for (desc->Ng, desc->interlace) in the wisdom file
 * desc->wisdom, the role FFTW's wisdom has for fft_init (src/fft.c:26-45),
 * and returns it.  Plans and batches created with that wisdom use it. */
int fftgal_power_autotune(const fftgal_power_desc *desc, const float *x, const float *y,
                          const float *z, long Np, double V, int reps, int *paint_tile,
                          int *fft_generic, int *bin_bk);
/* use the hipGraph (1, default) or plain stream launches (0) */
int fftgal_power_use_graph(fftgal_power_t *p, int on);
/* FFT kernels: on = 1 forces the generic LDS Stockham kernels, on = 0
 * restores the default (compile-time specialised kernels where the grid
 * side has them).  Returns 1 if the generic kernels are now in use, 0 if
 * the specialised ones, < 0 on error. */
int fftgal_power_fft_generic(fftgal_power_t *p, int on);
/* the plan's stream (hipStream_t) */
void *fftgal_power_stream(fftgal_power_t *p);
/* run one catalog eagerly with events between the pipeline stages and
 * return each stage's time in milliseconds (fftgal_power_stages() values;
 * names from fftgal_power_stage_name); results as after an enqueue */
int fftgal_power_stages(void);
const char *fftgal_power_stage_name(int i);
int fftgal_power_profile(fftgal_power_t *p, const float *x, const float *y, const float *z,
                         long Np, double V, float *stage_ms);
/* synchronous copy from device memory (e.g. the grids above) to host */
int fftgal_memcpy_d2h(void *dst, const void *src, unsigned long bytes);

/* ---- batches of catalogs ----
 * Many catalogs, each on its own grid -- mocks and subvolumes of different
 * boxes and sizes, and the same catalog folded (grid box L/fold with the
 * positions in box L, src/app/P.c:35-36) -- through the pipeline above
 * with one launch per stage for all of them: one sort launch per stage,
 * one paint launch (tile sides 8, 12 and 16 in one kernel, each workgroup
 * dispatching on its catalog's side), one launch per FFT pass (grid sides
 * 32 to 512: multi-size kernels, each workgroup running the specialised
 * transform of its grid's side), one fused bin launch pair, one copy of all
 * bin sums.  Sides outside those sets, and FFTGAL_BATCH_MULTI=0 in the
 * environment at creation, run one launch per tile or grid side instead.
 * The catalogs are cut into rounds through a bounded device pool
 * (FFTGAL_BATCH_BYTES, default 34 GiB; fftgal_batch_pool_bytes); each
 * round's sort and paint run on a second stream beside the previous
 * round's FFT passes and bin (FFTGAL_BATCH_SERIAL=1 in the environment at
 * enqueue keeps one stream).  Every round, every stage, is captured once in
 * a hipGraph; the catalogs'
 * pointers, Np and V are read from device tables, so later enqueues replay
 * the same graph.  A catalog's grid may need at most 256 x 128 paint tiles
 * (Ng <= 512 with tiles of 16); larger grids take a fftgal_power_t of their
 * own. */
typedef struct fftgal_batch fftgal_batch_t;

typedef struct fftgal_job_desc {
    int Ng;              /* grid cells per side: even, prime factors <= 13 */
    double L;            /* grid box side = catalog box / fold */
    int interlace;       /* 1: plain + dual grid */
    double dK;           /* k-bin width */
    long Np_max;         /* largest catalog this entry accepts */
    int paint_tile;      /* paint tile side (0: tuned / default) */
} fftgal_job_desc;

/* los and Nmu are shared by all catalogs; wisdom as in fftgal_power_desc */
int fftgal_batch_create(fftgal_batch_t **out, int njobs, const fftgal_job_desc *jobs,
                        const double los[3], int Nmu, const char *wisdom);
int fftgal_batch_destroy(fftgal_batch_t *b);
int fftgal_batch_NK(const fftgal_batch_t *b, int job);
/* catalog j: Np[j] float32 positions x[j], y[j], z[j] in device memory,
 * volume V[j] (several entries may share one catalog's arrays).
 * Asynchronous on the batch's stream; when the arrays, Np or V differ
 * from the previous enqueue it first waits for the stream. */
int fftgal_batch_enqueue(fftgal_batch_t *b, const float *const *x, const float *const *y,
                         const float *const *z, const long *Np, const double *V);
int fftgal_batch_sync(fftgal_batch_t *b);
/* results of catalog `job` after sync, as fftgal_power_results */
int fftgal_batch_results(fftgal_batch_t *b, int job, double *K, long *N, double *Pl,
                         double *Kmu, long *Nmu_count, double *Pmu);
/* eager run with events between the stages (names: fftgal_power_stage_name) */
int fftgal_batch_profile(fftgal_batch_t *b, const float *const *x, const float *const *y,
                         const float *const *z, const long *Np, const double *V,
                         float *stage_ms);
int fftgal_batch_use_graph(fftgal_batch_t *b, int on);
/* paint tile side of catalog `job` */
int fftgal_batch_tile(const fftgal_batch_t *b, int job);
/* int64 grid sets of the batch, as fftgal_power_grid_sets */
int fftgal_batch_grid_sets(const fftgal_batch_t *b);
/* Device memory: the grids, spectra and sort words of a batch live in one
 * pool bounded by FFTGAL_BATCH_BYTES (environment, bytes, at creation;
 * default 34 GiB).  The catalogs run in rounds of about equal size, in
 * input order, as few as the bound allows (a catalog larger than the bound
 * gets a round of its own); results do not depend on the rounds.  The
 * number of rounds and the pool's size: */
int fftgal_batch_rounds(const fftgal_batch_t *b);
size_t fftgal_batch_pool_bytes(const fftgal_batch_t *b);
void *fftgal_batch_stream(fftgal_batch_t *b);

/* ---- input ----
 * A QPM catalog file (the format of gal_loadqpm_cubic, include/io.h)
 * straight into device float32 arrays x, y, z of at least Np_max elements:
 * threads parse the lines a piece at a time into pinned host buffers, and
 * each piece's copy to the device overlaps the parsing of the next.  Values
 * are those of the drop-in loader rounded to float32.  *Np receives the
 * number of galaxies; more than Np_max lines is an error. */
int fftgal_qpm_load_device(const char *file, long Np_max, float *x, float *y, float *z,
                           long *Np);

#ifdef __cplusplus
}
#endif
#endif
