/*
 * fgk.h -- the thin C-ABI between the C host layer (the host C files) and the HIP
 * kernels for gfx950 (the .hip files).  Plain C types only: device pointers are
 * void* / typed pointers, streams are opaque.  Every function returns 0 on
 * success or a negative code; fgk_last_error() gives the message.
 *
 * Internal to libfftgal_amd.so: the public API is include/fft.h, gal.h,
 * power.h, fftgal_ext.h and fftgal_batch.h.
 */
#ifndef FGK_H
#define FGK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *fgk_stream_t;     /* hipStream_t */
typedef void *fgk_event_t;      /* hipEvent_t */

/* ---- runtime (fgk_runtime.hip) ---- */
int  fgk_init(void);                       /* idempotent; selects device 0 */
const char *fgk_last_error(void);
void fgk_set_error(const char *fmt, ...);
int  fgk_malloc(void **ptr, size_t bytes);
size_t fgk_mem_free(void);                   /* free device memory, bytes (0 if unknown) */
int  fgk_free(void *ptr);
int  fgk_host_alloc(void **ptr, size_t bytes);   /* pinned host memory */
int  fgk_host_free(void *ptr);
int  fgk_h2d(void *dst, const void *src, size_t bytes, fgk_stream_t s);
int  fgk_d2h(void *dst, const void *src, size_t bytes, fgk_stream_t s);
int  fgk_d2d(void *dst, const void *src, size_t bytes, fgk_stream_t s);
int  fgk_fill_zero(void *dst, size_t bytes, fgk_stream_t s);  /* kernel, graph safe */
int  fgk_stream_create(fgk_stream_t *s);
int  fgk_stream_destroy(fgk_stream_t s);
int  fgk_stream_sync(fgk_stream_t s);
int  fgk_event_create(fgk_event_t *e);
int  fgk_event_destroy(fgk_event_t e);
int  fgk_event_record(fgk_event_t e, fgk_stream_t s);
int  fgk_event_elapsed_ms(float *ms, fgk_event_t a, fgk_event_t b);
int  fgk_event_sync(fgk_event_t e);         /* wait for the work before its record */
int  fgk_event_create_dep(fgk_event_t *e);  /* ordering only, no timestamps */
int  fgk_stream_wait_event(fgk_stream_t s, fgk_event_t e);  /* s waits for e's record */
int  fgk_device_info(char *name, size_t len, int *cu_count, size_t *mem_bytes);
/* hipGraph capture of the work enqueued on a stream between begin/end */
typedef void *fgk_graph_t;
int  fgk_graph_begin(fgk_stream_t s);
int  fgk_graph_end(fgk_stream_t s, fgk_graph_t *g);
int  fgk_graph_launch(fgk_graph_t g, fgk_stream_t s);
int  fgk_graph_destroy(fgk_graph_t g);

/* ---- FFT engine (fgk_fft.hip) ---- */
/* plan for cubic grids of side Ng (even), precision 4 (float) or 8 (double);
 * twiddle tables are generated in FP64 on the host and stored in the plan's
 * precision on the device */
typedef struct fgk_fftplan fgk_fftplan;
int  fgk_fft_plan_create(fgk_fftplan **plan, int Ng, int precision);
void fgk_fft_plan_destroy(fgk_fftplan *plan);
int  fgk_fft_supported(int Ng);     /* 1 if Ng = 2^a 3^b 5^c 7^d 11^e 13^f, even */
/* 1 if an FP64 plan takes Ng: any Ng in [1, 1024] (odd sides, or prime
 * factors above 13, run complex transforms of whole rows in pairs; Bluestein
 * for lengths with a prime factor above 13), or any side fgk_fft_supported */
int  fgk_fft_supported_f64(int Ng);
/* FP64, in place on the padded layout f[(i*Ng+j)*Ng_pad+k], unnormalised,
 * forward sign -1; the result is multiplied by scale */
int  fgk_fft_r2c_f64(fgk_fftplan *p, double *f, double scale, fgk_stream_t s);
int  fgk_fft_c2r_f64(fgk_fftplan *p, double *f, double scale, fgk_stream_t s);
/* FP32 forward r2c whose real input is an int64 fixed-point grid
 * src[(i*Ng+j)*Ng+k] times in_scale; src rows are cleared to zero after they
 * are read when clear_src != 0.  Output in the packed half-spectrum layout
 * c[(i*Ng+j)*(Ng/2) + k], k < Ng/2 (complex float): the real k = 0 and
 * k = Ng/2 values of every row transform share column 0 as (X0, XN), so
 * after the column passes column 0 holds Z = F(X0) + i F(XN) and
 *   F(X0)[m] = (Z[m] + conj Z[-m]) / 2,  F(XN)[m] = (Z[m] - conj Z[-m]) / 2i */
int  fgk_fft_r2c_f32_from_fixed(fgk_fftplan *p, float *f, int64_t *src,
                                double in_scale, int clear_src, fgk_stream_t s);
/* the three passes separately, each over nbatch grids stored one after the
 * other (src: Ng^3 int64 apart; f: Ng^3/2 complex apart), in one launch */
int  fgk_fft_rows_f32_from_fixed(fgk_fftplan *p, float *f, int64_t *src,
                                 double in_scale, int clear_src, int nbatch, fgk_stream_t s);
int  fgk_fft_cols_f32(fgk_fftplan *p, float *f, int axis, int sign, int nbatch,
                      fgk_stream_t s);
/* One launch per FP32 pass over groups of grids of different sides
 * (batches): group g = nbatch[g] grids of side plans[g]->Ng one after the
 * other, spectra at f[g], int64 grids at src0[g] (grid set 0) and src1[g]
 * (set 1, or src1 NULL).  *m stays NULL (return 0) when a group's side has
 * no multi-size body (sides 32..512 of the specialised kernels, no
 * four-step): the caller then launches per group.  The row pass reads the
 * int64 grids of `set` and zeroes them when clear_src. */
typedef struct fgk_fftmulti fgk_fftmulti;
int  fgk_fft_multi_create(fgk_fftmulti **m, int ngroups, fgk_fftplan *const *plans,
                          float *const *f, int64_t *const *src0, int64_t *const *src1,
                          const int *nbatch);
void fgk_fft_multi_destroy(fgk_fftmulti *m);
int  fgk_fft_multi_rows(fgk_fftmulti *m, int set, int clear_src, fgk_stream_t s);
int  fgk_fft_multi_cols(fgk_fftmulti *m, int axis, int sign, fgk_stream_t s);
/* 1 if the FP32 passes of this Ng run compile-time specialised kernels
 * (otherwise the generic LDS Stockham kernels) */
int  fgk_fft_f32_specialised(int Ng);
/* force the generic FP32 kernels (1) or the default choice (0); for tests */
int  fgk_fft_force_generic(fgk_fftplan *p, int on);
/* n1 > 0 when the plan's FP32 i-axis pass runs as a four-step Ng = n1 x
 * (Ng / n1), which leaves frequency i on i-plane (i % n1) * (Ng / n1) +
 * i / n1; 0 when the planes are in natural order */
int  fgk_fft_fourstep(const fgk_fftplan *p);
/* tuning knobs: columns per workgroup for the strided passes, rows per
 * workgroup for the row pass (0 = default); generic kernels only */
int  fgk_fft_set_tuning(fgk_fftplan *p, int cols_per_wg, int rows_per_wg);
/* the generic kernels' launch shape the plan uses now (for its precision) */
int  fgk_fft_launch_shape(const fgk_fftplan *p, int *cols, int *rows);

/* ---- painting (fgk_paint.hip) ---- */
/* FP64 PCS paint with FP64 global atomics (reference order of operations):
 * f[...] += w (weight_sign = +1) or f[...] -= alpha*w (weight_sign = -1) */
int  fgk_paint_f64(double *f, int Ng, const double *x, const double *y,
                   const double *z, long Np, double Hinv, const double off[3],
                   double alpha, int weight_sign, fgk_stream_t s);
int  fgk_scale_f64(double *f, long n, double scale, fgk_stream_t s);
int  fgk_avg_f64(double *f, const double *g, long n, fgk_stream_t s);  /* f = 0.5*(g+f) */
int  fgk_widen_f32_f64(double *dst, const float *src, long n, fgk_stream_t s);

/* deterministic fixed-point tile paint of float32 positions into one
 * (plain, offset 0) or two (plus dual, offset 1/2) int64 grids */
typedef struct fgk_paintplan fgk_paintplan;
int  fgk_paint_plan_create(fgk_paintplan **pp, int Ng, int tile, long Np_max,
                           int interlace);
void fgk_paint_plan_destroy(fgk_paintplan *pp);
int  fgk_paint_fixed(fgk_paintplan *pp, const float *x, const float *y,
                     const float *z, long Np, double Hinv,
                     int64_t *gridA, int64_t *gridB, fgk_stream_t s);
/* Particle mass of one unit of the int64 grids: 1 / (216 C_25), within
 * 1.1e-4 of 2^-25.  Each tile paints with the fixed point its particle
 * count allows (fgk_paint_tile_bits) and is shifted to this unit. */
double fgk_paint_fixed_scale(void);
/* fraction bits S (19..25) of a tile whose work items hold n particles */
int  fgk_paint_tile_bits(long n);
/* Memory the paint kernel clears while it paints (bytes a multiple of 16,
 * 16-byte aligned; NULL: none): spread over the paint's workgroups, for a
 * caller that alternates between two grid sets and so needs the other set
 * zero for its next paint.  Kept until changed. */
void fgk_paint_set_zero(fgk_paintplan *pp, void *ptr, size_t bytes);
int  fgk_paint_plan_tile(const fgk_paintplan *pp);
/* the paint stages separately (profiling / tests): fgk_paint_nstages()
 * stages, named by fgk_paint_stage_name() */
int  fgk_paint_nstages(void);
const char *fgk_paint_stage_name(int i);
int  fgk_paint_stage(fgk_paintplan *pp, int stage, const float *x,
                     const float *y, const float *z, long Np, double Hinv,
                     int64_t *gridA, int64_t *gridB, fgk_stream_t s);

/* the same paint for many catalogs, each with its own grid, in one launch
 * per sort stage and one paint launch per tile side (the same
 * fgk_paint_nstages() stages).  A catalog's grid must need at most
 * 256 x 128 tiles (Ng <= 512 with tiles of 16). */
typedef struct fgk_paintjob_desc {
    int Ng;
    int tile;            /* tile side (0: default) */
    double Hinv;         /* Ng / grid box */
    long Np_max;
    int interlace;
    int64_t *gridA, *gridB;   /* zeroed int64 grids (gridB: interlace only) */
    int64_t *gridA2, *gridB2; /* optional second set (NULL: one set) */
} fgk_paintjob_desc;
typedef struct fgk_paintbatch fgk_paintbatch;
int  fgk_paint_batch_create(fgk_paintbatch **pb, int njobs, const fgk_paintjob_desc *jobs);
/* the same with the sort words (the arrays proportional to the particles)
 * carved from a caller's device workspace of at least
 * fgk_paint_batch_ws_bytes(sum of Np_max) bytes, which it keeps */
int  fgk_paint_batch_create_ws(fgk_paintbatch **pb, int njobs, const fgk_paintjob_desc *jobs,
                               void *ws, size_t ws_bytes);
size_t fgk_paint_batch_ws_bytes(long Np_tot);
void fgk_paint_batch_destroy(fgk_paintbatch *pb);
/* positions of every catalog (device float32) and Np[j] <= Np_max; when
 * they differ from the last call, waits for s and uploads the job table */
int  fgk_paint_batch_set(fgk_paintbatch *pb, const float *const *x, const float *const *y,
                         const float *const *z, const long *Np, fgk_stream_t s);
int  fgk_paint_batch_stage(fgk_paintbatch *pb, int stage, fgk_stream_t s);
/* the paint stage of catalogs with tile sides 8, 12 and 16 runs as one
 * launch (default), or one launch per tile side (on) */
void fgk_paint_batch_per_side(fgk_paintbatch *pb, int on);
/* grid set the paint stage writes (1: gridA2/gridB2) and the memory its
 * launches clear meanwhile (NULL: none; bytes a multiple of 16) */
int  fgk_paint_batch_select(fgk_paintbatch *pb, int set, void *zero, size_t bytes);
int  fgk_paint_batch_tile(const fgk_paintbatch *pb, int job);

/* ---- k space (fgk_kspace.hip) ---- */
/* phase[d][j] complex table (3*Ng*2 doubles) built on the host */
int  fgk_offset_phase_f64(double *f, int Ng, const double *phase_dev, fgk_stream_t s);
int  fgk_deconv_f64(double *f, int Ng, const double *winv_dev, fgk_stream_t s);
int  fgk_P_f64(double *f, int Ng, double Vinv, double Pshot, fgk_stream_t s);
/* super-sample k filter (kind 0: sub box, tables wamp[Ng], wph[3*Ng][2];
 * kind 1: sub sphere) times L_ell(mu), and the gather of the Nsub^3
 * sub-volume corners of the configuration-space grid divided by bias */
int  fgk_dl_filter_f64(double *f, int Ng, int kind, int Nsub, const double los[3], int ell,
                       const double *wamp_dev, const double *wph_dev, fgk_stream_t s);
int  fgk_dl_gather_f64(const double *f, int Ng, int Nsub, double bias, double *out_dev,
                       fgk_stream_t s);
/* sub-volume window deconvolution of a real-space grid (kind 0: sub box of
 * Ng/Nsub cells; kind 1: sub sphere of diameter sqrt(3) Ng/Nsub cells) */
int  fgk_subvolume_winv_f64(double *f, int Ng, int kind, int Nsub, fgk_stream_t s);

/* binning.  Bins are accumulated into bins_dev (doubles) laid out as
 *   [NK][4]         Pl sums (l = 0,2,4,6)      at offset 0
 *   [NK][W]         Pmu sums                    at offset NK*4
 * and, when geometry != 0, also
 *   [NK]            K sums (count*Kamp)         at offset NK*(4+W)
 *   [NK][W]         K sums per wedge            at offset NK*(5+W)
 * where W = Nmu+1 (wedge Nmu collects mu = 1) or 0 without wedges,
 * with integer counts in counts_dev (int64): [NK] then [NK][W].
 * bins_dev/counts_dev must be zeroed by the caller. */
typedef struct fgk_binspec {
    int Ng;
    int NK;
    int Nmu;            /* 0: no wedges */
    double dKinv;       /* KF/dK */
    double los[3];      /* already normalised */
    int do_pl;          /* accumulate the Legendre sums */
    int geometry;       /* also accumulate K sums and counts */
    int bk;             /* fused kernel: k extent of a workgroup's block of
                           1024 modes, 16, 32, 64 or 128 (0: 64) */
    int fs_n1;          /* fused kernel: the spectra's i planes are in the
                           order of a four-step i pass (fgk_fft_fourstep),
                           frequency i on plane (i % fs_n1) * (Ng / fs_n1) +
                           i / fs_n1; 0: natural order */
} fgk_binspec;
size_t fgk_bins_doubles(const fgk_binspec *b);
size_t fgk_bins_counts(const fgk_binspec *b);
/* Sums of the fast path (fgk_bin_geometry, fgk_fused_bin_f32*) are exact
 * fixed point: value i of the fgk_bins_doubles layout is the int64 pair
 * (H, L) at 2i, 2i+1, worth (H 2^32 + L) 2^(e - FGK_BIN_FIX_BITS), e the
 * binary exponent (ilogb) of the catalog's Pshot for power sums,
 * FGK_BIN_K_EXP for the geometry's K sums (fgk_bin_fixed_value) */
#define FGK_BIN_FIX_BITS 24
#define FGK_BIN_K_EXP (-28)
double fgk_bin_fixed_value(const int64_t *hl, int e);
/* e of a catalog's power sums: ilogb(Pshot), 0 unless 0 < Pshot < 1e300 */
int fgk_bin_fixed_exponent(double Pshot);
/* geometry only: K sums (fixed point) and counts (b->geometry must be 1);
 * P sums untouched */
int  fgk_bin_geometry(const fgk_binspec *b, int64_t *sums_dev, int64_t *counts_dev,
                      fgk_stream_t s);
/* the fixed device scratch of the bin sums (1 MiB, one copy per device):
 * its device address on the current device and its size */
#define FGK_BIN_SCRATCH_BYTES (1u << 20)
int  fgk_bin_scratch(void **ptr, size_t *bytes);
/* bin the FP64 grid holding P(k) in the real slots (after P()) */
int  fgk_bin_pgrid_f64(const double *f, const fgk_binspec *b, double *bins_dev,
                       int64_t *counts_dev, fgk_stream_t s);
/* fused: phase (dual grid), interlace average, deconvolution, |delta|^2*norm
 * - Pshot, binning.  A = plain grid, B = dual grid (NULL: no interlacing),
 * both in the packed half-spectrum layout of fgk_fft_r2c_f32_from_fixed.
 * tables: phase (3*Ng complex double), winv (Ng double) on the device.
 * Each spectrum is normalised by norm * npart / (its DC mode, the painted
 * mass in particle units).
 * Power sums only (b->geometry must be 0; fgk_bin_geometry gives K and
 * the counts).  Two launches: modes with 0 < k < Ng/2, then the k = 0 and
 * k = Ng/2 planes. */
int  fgk_fused_bin_f32(const float *A, const float *B, const fgk_binspec *b,
                       const double *phase_dev, const double *winv_dev,
                       double norm, double Vinv, double Pshot, double npart,
                       int64_t *sums_dev, fgk_stream_t s);

/* one catalog of a batched fused bin over grids of one side: packed
 * spectra A, B (B NULL: no interlacing), its bin sums (layout as above for
 * NK bins, geometry 0), dKinv = KF/dK of its box, and the arguments of
 * fgk_fused_bin_f32 */
typedef struct fgk_binjob {
    const float *A, *B;
    int64_t *sums;         /* fixed-point pairs */
    int NK;
    double dKinv;
    double norm, Vinv, Pshot, npart;
    /* fgk_fused_bin_f32_multi only: the catalog's grid side, four-step split
     * of its spectra, its side's offset-phase and deconvolution tables, and
     * its first workgroup in the bin and plane launches */
    int Ng, fs_n1;
    const double *phase, *winv;
    unsigned blk0, pblk0;
} fgk_binjob;
/* one launch pair for njobs catalogs (table in device memory) whose grids
 * share b->Ng, Nmu, los and do_pl; NK_max = largest NK of the catalogs */
int  fgk_fused_bin_f32_batch(const fgk_binspec *b, const fgk_binjob *jobs_dev, int njobs,
                             int NK_max, const double *phase_dev, const double *winv_dev,
                             fgk_stream_t s);
/* workgroups of one catalog of side Ng in the launches of
 * fgk_fused_bin_f32_multi (bin block shape b->bk) */
void fgk_bin_multi_blocks(const fgk_binspec *b, int Ng, unsigned *blocks, unsigned *pblocks);
/* one launch pair for njobs catalogs of any grid sides (b->Ng unused; each
 * job carries Ng, fs_n1, phase, winv, blk0, pblk0 with blocks laid out by
 * fgk_bin_multi_blocks in job order), sharing Nmu, los, do_pl and the bin
 * block shape; nblocks / npblocks: the launches' workgroups */
int  fgk_fused_bin_f32_multi(const fgk_binspec *b, const fgk_binjob *jobs_dev, int njobs,
                             int NK_max, unsigned nblocks, unsigned npblocks, fgk_stream_t s);

#ifdef __cplusplus
}
#endif
#endif
