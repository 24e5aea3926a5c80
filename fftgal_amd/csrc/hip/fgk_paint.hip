// PCS mass assignment on gfx950.
//
// Two forms of the reference's fft_p2g (src/fft.c:57-131):
//
// 1. k_paint_f64 -- the plain FP64 form behind the drop-in fft_p2g: one
//    thread per particle, the reference's weights and index wrap in FP64,
//    64 FP64 global atomic adds per particle into the padded grid.
//
// 2. The deterministic fixed-point tile paint of the fast path, for float32
//    positions and the interlaced pair of grids (offset 0 and offset 1/2,
//    src/fft.c:257-276).  Particles are first grouped by paint tile (T^3
//    cells, by the plain grid's PCS base cell) with a two-level MSD
//    counting sort -- coarse bucket (<= 256 per grid) then tile within the
//    bucket (<= 128); grids with more tiles take 2048 x 1024 digits -- whose
//    scatters stage each 4096-particle chunk in LDS
//    ordered by digit, so that global writes are contiguous runs.  The
//    scatters' ranks come from a wave-level match of equal digits (one LDS
//    atomic per distinct digit per wave), so heavily clustered catalogs,
//    where most particles share a few buckets, do not serialise on one LDS
//    counter; the count passes add one LDS atomic per particle (one per
//    wave when the wave holds a single digit), which measured faster.
//    Positions travel as tile-local fixed point with 2^-19 cell resolution
//    (tiles of 16): a 64-bit main word of three 21-bit coordinates plus an
//    8-bit extension of 2 more bits per axis; between the levels the
//    extension shares a 16-bit side word with the tile-within-bucket digit,
//    which is all the level-2 histogram reads.  The paint
//    then runs one workgroup per pencil segment (the tiles along k of one
//    tile column of a coarse bucket) or per dense item (<= cap particles of
//    a tile holding more): the 4x4x4 stencils of both grids go into LDS
//    footprints ((T+4)^3 plain, (T+3)^3 dual cells, one-sided: a tile's
//    stencils reach past its upper faces only) as int32 fixed point.  The
//    footprints are stored in 4x4x4 bricks, so a cell's LDS bank is fixed
//    by its coordinates mod 4, and each lane walks its stencil in the
//    residue order given by its six low bits: at every step the 64 lanes'
//    atomics fall in 64 distinct banks (and so never on one address)
//    whatever particles the wave holds.  Non-zero footprint cells are
//    flushed into the int64 grids with global integer atomics; along a
//    pencil only the T planes a tile owns are flushed (aligned 128-byte
//    rows) and the planes above are carried in LDS into the next tile.
//    Integer addition is associative, so the grids are bit-identical run to
//    run whatever the scheduling.  Each segment or item rounds its weights
//    to the finest fixed point its particle count allows without int32
//    overflow (2^-18 particle for 16384 particles, up to 2^-25), with a
//    deterministic dither that makes every rounding unbiased, so no term is
//    lost on average however small; the grids hold one common unit.
//    Grids with more tiles than even the wide two levels address (over
//    2^21) fall back to a single-level sort whose ranks come from global
//    atomics.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <math.h>
#include <stdlib.h>
#include "../fgk.h"
#include "fgk_common.h"

// ----------------------------------------------------------- plain FP64

__device__ __forceinline__ void pcs_axis_f64(double u, int Ng, int idx[4], double w[4]) {
#pragma clang fp contract(off)
    int g1 = (int)floor(u);
    double d = u - g1;
    g1 = (g1 % Ng + Ng) % Ng;
    idx[0] = (g1 - 1 + Ng) % Ng;
    idx[1] = g1;
    idx[2] = (g1 + 1) % Ng;
    idx[3] = (g1 + 2) % Ng;
    double e = 1 - d;
    w[0] = e * e * e / 6;
    w[1] = ((3 * d - 6) * (d * d) + 4) / 6;
    w[2] = (((-3 * d + 3) * d + 3) * d + 1) / 6;
    w[3] = d * d * d / 6;
}

__global__ void __launch_bounds__(256)
k_paint_f64(double *f, int Ng, int Ng_pad, const double *__restrict__ x,
            const double *__restrict__ y, const double *__restrict__ z, long Np, double Hinv,
            double o0, double o1, double o2, double alpha, int negate) {
#pragma clang fp contract(off)
    const long stride = (long)gridDim.x * blockDim.x;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < Np; p += stride) {
        int gi[4], gj[4], gk[4];
        double wx[4], wy[4], wz[4];
        pcs_axis_f64(x[p] * Hinv - o0, Ng, gi, wx);
        pcs_axis_f64(y[p] * Hinv - o1, Ng, gj, wy);
        pcs_axis_f64(z[p] * Hinv - o2, Ng, gk, wz);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double *row = f + ((long)gi[i] * Ng + gj[j]) * Ng_pad;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    double v = negate ? alpha * wx[i] * wy[j] * wz[k] : wx[i] * wy[j] * wz[k];
                    atomicAdd(row + gk[k], negate ? -v : v);
                }
            }
    }
}

__global__ void k_scale_f64(double *f, long n, double s) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) f[i] *= s;
}

__global__ void k_avg_f64(double *f, const double *g, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) f[i] = 0.5 * (g[i] + f[i]);
}

__global__ void k_widen(double *d, const float *s, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) d[i] = (double)s[i];
}

static unsigned stream_blocks(long n) {
    long b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (unsigned)b;
}

extern "C" int fgk_paint_f64(double *f, int Ng, const double *x, const double *y, const double *z,
                             long Np, double Hinv, const double off[3], double alpha,
                             int weight_sign, fgk_stream_t s) {
    if (Np <= 0) return 0;
    const int Ng_pad = 2 * (Ng / 2 + 1);
    const long nb = (Np + 255) / 256;    // grid-stride beyond 2^28 work-items
    hipLaunchKernelGGL(k_paint_f64, dim3((unsigned)(nb < (1l << 20) ? nb : (1l << 20))),
                       dim3(256), 0, (hipStream_t)s, f,
                       Ng, Ng_pad, x, y, z, Np, Hinv, off[0], off[1], off[2], alpha,
                       weight_sign < 0 ? 1 : 0);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_scale_f64(double *f, long n, double scale, fgk_stream_t s) {
    hipLaunchKernelGGL(k_scale_f64, dim3(stream_blocks(n)), dim3(256), 0, (hipStream_t)s, f, n,
                       scale);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_avg_f64(double *f, const double *g, long n, fgk_stream_t s) {
    hipLaunchKernelGGL(k_avg_f64, dim3(stream_blocks(n)), dim3(256), 0, (hipStream_t)s, f, g, n);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_widen_f32_f64(double *dst, const float *src, long n, fgk_stream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_widen, dim3(stream_blocks(n)), dim3(256), 0, (hipStream_t)s, dst, src,
                       n);
    FGK_LAUNCHED();
    return 0;
}

// ------------------------------------------------------ fixed-point tiles

#define PAINT_THREADS 1024
#define SORT_THREADS 256
#define SORT_CHUNK 4096                  // particles per sort workgroup
#define SORT_PER_THREAD (SORT_CHUNK / SORT_THREADS)
#define SCAN_THREADS 1024
#define RB 32                            // sort chunks per scan block
#define MAX_D1 256                       // coarse buckets
#define MAX_D2 128                       // tiles per coarse bucket
#define WIDE_D1 2048                     // the same for grids with more tiles
#define WIDE_D2 1024
// A particle travels through the sort as its tile-local position in fixed
// point, li + d with F fraction bits per axis (F = 23 - log2 T: 2^-19 cell
// for tiles of 16): the top POS_MAIN bits of each axis in a 64-bit main word
// (x, y, z from the top), the low POS_EXT bits of the three axes in an 8-bit
// extension.  Between the two sort levels the extension shares a 16-bit
// side word with the tile-within-bucket digit (digit << 6 | extension, at
// most 10 digit bits), which the level-2 histogram reads alone.
#define POS_MAIN 21
#define POS_EXT 2
#define POS_MASK ((1u << POS_MAIN) - 1u)
#define AUX_SHIFT (3 * POS_EXT)

struct KeyGeo {
    double Hinv;
    int Ng, T, nt, F;    // F: fraction bits of the packed tile-local positions
    int cg0, cg1, cg2;   // tiles per coarse bucket per axis (powers of two)
    int nc0, nc1, nc2;   // coarse buckets per axis
    int lg0, lg1, lg2;   // log2 of cg0, cg1, cg2
    uint64_t Tmul;       // ceil(2^32 / T): i / T == (i * Tmul) >> 32 for 0 <= i < 2^24
    int Tshift;          // log2 T when T is a power of two, else -1
};

struct fgk_paintplan {
    int Ng, T, nt, interlace, F, cap;
    int two_level;
    int wide;            // two levels of WIDE_D1 x WIDE_D2 digits
    uint32_t *bitem, *ibase, *bscratch;   // [nb1+1] wide mode: items per bucket, their scan
    KeyGeo g;            // Hinv filled per call
    int nb1, nb2, bits1, bits2;
    long nslots;         // nb1 * nb2 tile slots
    long Np_max;
    long nw1max, nw2max; // sort chunks (multiples of RB)
    uint32_t *H1, *O1, *S1, *B1, *T1;
    uint32_t *H2, *O2, *S2, *B2;
    uint32_t *seg1;      // {0, blocks} for the single level-1 segment
    uint32_t *base1;     // [nb1+1]
    uint32_t *chunk0;    // [nb1+1] first level-2 chunk of each bucket (multiple of RB)
    uint32_t *blk0;      // [nb1+1] chunk0 / RB
    uint32_t *T2;        // [nslots] particles per tile
    uint32_t *tilebase;  // [nslots+1]
    uint32_t *items;     // [nslots+1] paint work items, exclusive prefix
    uint32_t *rank32;    // [Np_max] single-level path
    uint64_t *zero_ptr;  // cleared by the paint kernel (fgk_paint_set_zero), 8-byte words
    long zero_words;
    uint64_t *sorted1, *sorted2;   // main words after level 1, level 2
    uint16_t *aux1;                // level-2 digit and extension after level 1
    uint8_t *ext2;                 // extension after level 2
};

// One catalog of a batched paint (fgk_paintbatch): its positions, grid
// geometry and grids, and where its work sits in the batch's launches.
struct PaintJob {
    const float *x, *y, *z;
    long Np;
    KeyGeo g;
    unsigned long long *GA, *GB;
    unsigned long long *GA2, *GB2;   // second grid set (set 1 of the paint launch)
    int nb1;             // coarse buckets (slots: nb1 * MAX_D2, MAX_D2 per bucket)
    int interlace;
};

// job owning entry e of a batch's per-job ranges off[0..njobs] (uniform
// operands: scalar loads)
__device__ __forceinline__ int batch_job(long e, const long *__restrict__ off, int njobs) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Tile and tile-local position of one coordinate.  u = x / H is taken in
// FP64 as the reference does (src/fft.c:68-76).  A particle belongs to the
// tile holding the base node of its dual-grid stencil, floor(u - 3/2): its
// dual stencil then covers tile-local nodes 0 .. T + 2 and its plain stencil
// (base floor(u) - 1 = floor(u - 3/2 + 1/2)) nodes 0 .. T + 3, so a tile's
// footprints extend past its upper faces only.  The stored position is
// v = u - 3/2 - T tc in [0, T), fixed point with F fraction bits (at most 26
// bits in all): the dual stencil has base floor(v) and fraction frac(v), the
// plain one base floor(v + 1/2) and fraction frac(v + 1/2).
__device__ __forceinline__ void axis_key(float xf, const KeyGeo &g, int &tc, uint32_t &q) {
#pragma clang fp contract(off)
    const double u = (double)xf * g.Hinv - 1.5;   // the product rounded, then an exact subtraction
    const double fl = floor(u);
    const double d = u - fl;
    int gi = (int)fl;
    // positions in [0, L) give gi in [-2, Ng): one conditional add wraps the
    // first cells' stencil bases; other positions (folded grids, catalogs
    // outside the box) take the remainder
    gi = gi < 0 && gi >= -g.Ng ? gi + g.Ng : gi;
    if ((unsigned)gi >= (unsigned)g.Ng) {
        gi %= g.Ng;
        if (gi < 0) gi += g.Ng;
    }
    tc = g.Tshift >= 0 ? gi >> g.Tshift : (int)(((uint64_t)(uint32_t)gi * g.Tmul) >> 32);
    const int li = gi - tc * g.T;
    int qq = (li << g.F) + (int)rint(d * (double)(1 << g.F));
    const int qmax = (g.T << g.F) - 1;
    q = (uint32_t)(qq < qmax ? qq : qmax);
}

// coarse digit (bucket), fine digit (tile within the bucket), main word and
// extension of one particle
__device__ __forceinline__ void particle_key(float xv, float yv, float zv, const KeyGeo &g,
                                             uint32_t &coarse, uint32_t &fine, uint64_t &pk,
                                             uint32_t &ext) {
    int tx, ty, tz;
    uint32_t qx, qy, qz;
    axis_key(xv, g, tx, qx);
    axis_key(yv, g, ty, qy);
    axis_key(zv, g, tz, qz);
    const int cx = tx >> g.lg0, cy = ty >> g.lg1, cz = tz >> g.lg2;
    coarse = (uint32_t)((cx * g.nc1 + cy) * g.nc2 + cz);
    fine = (uint32_t)((((tx & (g.cg0 - 1)) << g.lg1 | (ty & (g.cg1 - 1))) << g.lg2) |
                      (tz & (g.cg2 - 1)));
    pk = ((uint64_t)(qx >> POS_EXT) << (2 * POS_MAIN)) | ((uint64_t)(qy >> POS_EXT) << POS_MAIN) |
         (uint64_t)(qz >> POS_EXT);
    constexpr uint32_t em = (1u << POS_EXT) - 1u;
    ext = ((qx & em) << (2 * POS_EXT)) | ((qy & em) << POS_EXT) | (qz & em);
}

// Rank of this lane among the valid lanes of its wave holding the same
// NB-bit digit, added to the workgroup histogram with one LDS atomic per
// distinct digit per wave, so that heavily clustered catalogs, where most
// particles share a few digits, do not serialise on one LDS counter.  All
// lanes of the wave must call it together.
template <int NB>
__device__ __forceinline__ uint32_t wave_rank(uint32_t d, bool valid, uint32_t *hist) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(valid && bit);
        peers &= bit ? bal : ~bal;
    }
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)peers) - 1;
    const uint32_t below = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
    uint32_t base = 0;
    if (valid && lane == leader) base = atomicAdd(&hist[d], (uint32_t)__popcll(peers));
    base = __shfl(base, leader < 0 ? 0 : leader);
    return base + below;
}

// The digit histogram only (no ranks): one LDS atomic per valid lane, or
// one per wave when all its valid lanes hold the same digit.  All lanes of
// the wave must call it together.
__device__ __forceinline__ void wave_count(uint32_t d, bool valid, uint32_t *hist) {
    const uint64_t vm = __ballot(valid);
    const int lead = __ffsll((unsigned long long)vm) - 1;
    const uint32_t dl = __shfl(d, lead < 0 ? 0 : lead);
    if (__ballot(valid && d == dl) == vm) {
        if (__lane_id() == lead) atomicAdd(&hist[dl], (uint32_t)__popcll(vm));
    } else if (valid) {
        atomicAdd(&hist[d], 1u);
    }
}

// exclusive scan of hist[0..n) (n <= NMAX <= 2048) into off[], by the first
// wave.  NMAX <= 256: each lane holds 4 entries in registers; larger: each
// lane sums a contiguous run of NMAX/64 entries, the wave scans the sums,
// and each lane writes its run's offsets
template <int NMAX>
__device__ __forceinline__ void small_exclusive_scan(const uint32_t *hist, uint32_t *off, int n) {
    if (threadIdx.x >= 64) return;
    const int lane = threadIdx.x;
    if constexpr (NMAX <= 256) {
        uint32_t v[4], s = 0;
        for (int q = 0; q < 4; ++q) {
            const int i = lane * 4 + q;
            v[q] = i < n ? hist[i] : 0u;
            s += v[q];
        }
        uint32_t incl = s;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        uint32_t run = incl - s;
        for (int q = 0; q < 4; ++q) {
            const int i = lane * 4 + q;
            if (i < n) off[i] = run;
            run += v[q];
        }
    } else {
        constexpr int PER = NMAX / 64;
        const int i0 = lane * PER, i1 = i0 + PER < n ? i0 + PER : n;
        uint32_t s = 0;
        for (int i = i0; i < i1; ++i) s += hist[i];
        uint32_t incl = s;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        uint32_t run = incl - s;
        for (int i = i0; i < i1; ++i) {
            const uint32_t c = hist[i];
            off[i] = run;
            run += c;
        }
    }
}

// level-2 chunk j -> its coarse bucket and particle range in sorted1
__device__ __forceinline__ bool chunk_range2(long j, const uint32_t *chunk0, const uint32_t *base1,
                                             const uint32_t *T1, int nb1, long &start, int &n) {
    if (j >= (long)chunk0[nb1]) return false;
    int lo = 0, hi = nb1 - 1;              // last bucket with chunk0[b] <= j
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((long)chunk0[mid] <= j) lo = mid; else hi = mid - 1;
    }
    const long jl = j - chunk0[lo];
    start = (long)base1[lo] + jl * SORT_CHUNK;
    const long end = (long)base1[lo] + T1[lo];
    const long m = end - start;
    n = m <= 0 ? 0 : (m < SORT_CHUNK ? (int)m : SORT_CHUNK);
    return true;
}

// LEVEL: 1 reads float positions (digit = coarse bucket) and writes main
// words (out) and side words (out_aux: fine digit << 6 | extension); 2
// reads them (digit = tile within bucket, from the side word alone for the
// histogram) and writes main words and extensions (out_ext).  COUNT:
// histogram only.
// Batched level 1 (jobs != nullptr): chunk j belongs to the job whose chunk
// range jc1[job] <= j < jc1[job+1] holds it, whose positions and geometry
// replace x, y, z, Np and g.
// NB: digit bits (8 / 7 for MAX_D1 / MAX_D2 digits, 11 / 10 for WIDE_D1 /
// WIDE_D2); nd <= 2^NB digits in use.
template <int LEVEL, int COUNT, int NB>
__global__ void __launch_bounds__(SORT_THREADS)
k_sort(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
       long Np, KeyGeo g, const uint64_t *__restrict__ in, const uint16_t *__restrict__ in_aux,
       const uint32_t *__restrict__ chunk0, const uint32_t *__restrict__ base1,
       const uint32_t *__restrict__ T1, int nb1, int nd, uint32_t *__restrict__ H,
       const uint32_t *__restrict__ O, uint64_t *__restrict__ out, uint16_t *__restrict__ out_aux,
       uint8_t *__restrict__ out_ext, const PaintJob *__restrict__ jobs,
       const long *__restrict__ jc1, int njobs) {
    constexpr int ND = 1 << NB;
    using dig_t = typename std::conditional<(NB > 8), uint16_t, uint8_t>::type;
    __shared__ uint32_t hist[ND];
    __shared__ uint32_t off[ND];                 // staged offsets, then global bases
    __shared__ uint64_t stage[COUNT ? 1 : SORT_CHUNK];
    __shared__ dig_t sdig[COUNT ? 1 : SORT_CHUNK];
    __shared__ long s_start;
    __shared__ int s_n;
    const long j = blockIdx.x;
    long cj = j;                                 // chunk within its catalog
    if (LEVEL == 1 && jobs) {
        const int jb = batch_job(j, jc1, njobs);
        const PaintJob &J = jobs[jb];
        x = J.x;
        y = J.y;
        z = J.z;
        Np = J.Np;
        g = J.g;
        cj = j - jc1[jb];
    }
    if (threadIdx.x == 0) {
        long st = 0;
        int n = 0;
        if (LEVEL == 1) {
            st = cj * SORT_CHUNK;
            const long m = Np - st;
            n = m <= 0 ? 0 : (m < SORT_CHUNK ? (int)m : SORT_CHUNK);
        } else if (!chunk_range2(j, chunk0, base1, T1, nb1, st, n)) {
            n = 0;
        }
        s_start = st;
        s_n = n;
    }
    for (int i = threadIdx.x; i < nd; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const long st = s_start;
    const int n = s_n;
    uint32_t dk[SORT_PER_THREAD], rk[SORT_PER_THREAD], ak[SORT_PER_THREAD];
    uint64_t vk[COUNT ? 1 : SORT_PER_THREAD];
    float xs[LEVEL == 1 ? SORT_PER_THREAD : 1], ys[LEVEL == 1 ? SORT_PER_THREAD : 1],
        zs[LEVEL == 1 ? SORT_PER_THREAD : 1];
    // all loads first, so that they are in flight together
#pragma unroll
    for (int k = 0; k < SORT_PER_THREAD; ++k) {
        const int i = k * SORT_THREADS + threadIdx.x;
        if (i < n) {
            if (LEVEL == 1) {
                xs[k] = x[st + i];
                ys[k] = y[st + i];
                zs[k] = z[st + i];
            } else {
                ak[k] = in_aux[st + i];
                if constexpr (!COUNT) vk[k] = in[st + i];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < SORT_PER_THREAD; ++k) {
        const int i = k * SORT_THREADS + threadIdx.x;
        const bool valid = i < n;
        uint32_t d = 0, a = 0;
        uint64_t pk = 0;
        if (valid) {
            if (LEVEL == 1) {
                uint32_t fine, ext;
                particle_key(xs[k], ys[k], zs[k], g, d, fine, pk, ext);
                a = (fine << AUX_SHIFT) | ext;
            } else {
                a = ak[k];
                d = a >> AUX_SHIFT;
                if constexpr (!COUNT) pk = vk[k];
            }
        }
        if constexpr (COUNT) {
            wave_count(d, valid, hist);
            rk[k] = 0;
        } else {
            rk[k] = wave_rank<NB>(d, valid, hist);
            vk[k] = pk;
        }
        dk[k] = d;
        ak[k] = LEVEL == 1 ? a : (a & ((1u << AUX_SHIFT) - 1u));
    }
    __syncthreads();
    if (COUNT) {
        for (int i = threadIdx.x; i < nd; i += blockDim.x) H[j * nd + i] = hist[i];
        return;
    }
    small_exclusive_scan<ND>(hist, off, nd);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SORT_PER_THREAD; ++k) {
        const int i = k * SORT_THREADS + threadIdx.x;
        if (i < n) {
            const uint32_t pos = off[dk[k]] + rk[k];
            stage[pos] = vk[k];
            sdig[pos] = (dig_t)dk[k];
            rk[k] = pos;                         // staged position, for the side words
        }
    }
    __syncthreads();
    // global destination of staged position pos with digit d: off[d] + pos
    for (int i = threadIdx.x; i < nd; i += blockDim.x) off[i] = O[j * nd + i] - off[i];
    __syncthreads();
    for (int pos = threadIdx.x; pos < n; pos += blockDim.x)
        out[off[sdig[pos]] + pos] = stage[pos];
    __syncthreads();
    // the side words (level 1) or extensions (level 2), staged the same way
    uint32_t *stage32 = (uint32_t *)stage;
#pragma unroll
    for (int k = 0; k < SORT_PER_THREAD; ++k) {
        const int i = k * SORT_THREADS + threadIdx.x;
        if (i < n) stage32[rk[k]] = ak[k];
    }
    __syncthreads();
    if (LEVEL == 1) {
        for (int pos = threadIdx.x; pos < n; pos += blockDim.x)
            out_aux[off[sdig[pos]] + pos] = (uint16_t)stage32[pos];
    } else {
        for (int pos = threadIdx.x; pos < n; pos += blockDim.x)
            out_ext[off[sdig[pos]] + pos] = (uint8_t)stage32[pos];
    }
}

// S[b][d] = sum of H[j][d] over the RB chunks j of block b
__global__ void k_blocksum(const uint32_t *__restrict__ H, long nblk, int nd,
                           uint32_t *__restrict__ S) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= nblk * nd) return;
    const long b = t / nd;
    const int d = (int)(t - b * nd);
    uint32_t s = 0;
    for (int r = 0; r < RB; ++r) s += H[(b * RB + r) * nd + d];
    S[t] = s;
}

// per segment s and digit d: exclusive prefix of S over the segment's
// blocks [blk0[s], blk0[s+1]) into B, and the segment total into T
__global__ void k_segscan(const uint32_t *__restrict__ S, const uint32_t *__restrict__ blk0,
                          int nseg, int nd, uint32_t *__restrict__ B, uint32_t *__restrict__ T) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nseg * nd) return;
    const int s = t / nd, d = t - s * nd;
    uint32_t run = 0;
    for (long b = blk0[s]; b < (long)blk0[s + 1]; ++b) {
        B[b * nd + d] = run;
        run += S[b * nd + d];
    }
    T[t] = run;
}

// the same for one long segment (all blocks [0, nblk)): one workgroup per
// digit d scans S[.][d] 256 blocks at a time, carrying the running total
__global__ void __launch_bounds__(256)
k_segscan_wide(const uint32_t *__restrict__ S, long nblk, int nd, uint32_t *__restrict__ B,
               uint32_t *__restrict__ T) {
    __shared__ uint32_t wsum[4];
    const int d = blockIdx.x;
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (long b0 = 0; b0 < nblk; b0 += 256) {
        const long b = b0 + threadIdx.x;
        const uint32_t v = b < nblk ? S[b * nd + d] : 0u;
        uint32_t incl = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        uint32_t before = carry;
        for (int w = 0; w < wv; ++w) before += wsum[w];
        if (b < nblk) B[b * nd + d] = before + incl - v;
        carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (threadIdx.x == 0) T[d] = carry;
}

// paint work items of a tile of c particles: the dense items of a tile of
// more than cap particles (cap particles each, the last fewer); tiles of at
// most cap particles are painted by their pencil segment
__host__ __device__ __forceinline__ uint32_t dense_items(uint32_t c, int cap) {
    return c > (uint32_t)cap ? (c + cap - 1) / cap : 0u;
}

// paint workgroups of one catalog: its pencil segments and an upper bound of
// its dense items (a dense tile of c > cap particles has ceil(c / cap) <
// 2 c / cap items, so all of them have fewer than 2 Np / cap)
static long paint_work(long nb1, const KeyGeo &g, long Np, int cap) {
    return nb1 * g.cg0 * g.cg1 + 2 * ((Np + cap - 1) / cap);
}

// single-workgroup exclusive scan of v[0..n) into base[0..n] and of an
// auxiliary per-entry count into aux[0..n]:
//   mode 1: aux = chunks of RB-aligned level-2 work, ceil(v/SORT_CHUNK) rounded up to RB;
//           also aux_blk = aux / RB
//   mode 2: aux = dense paint items (dense_items)
__global__ void __launch_bounds__(SCAN_THREADS)
k_digitscan(const uint32_t *__restrict__ v, long n, int mode, int cap,
            uint32_t *__restrict__ base, uint32_t *__restrict__ aux,
            uint32_t *__restrict__ aux_blk) {
    constexpr int NW = SCAN_THREADS / 64;
    __shared__ uint32_t sa[NW], sb[NW];
    __shared__ uint32_t ta, tb;                  // totals
    const long per = (n + SCAN_THREADS - 1) / SCAN_THREADS;
    const long i0 = threadIdx.x * per;
    const long i1 = i0 + per < n ? i0 + per : n;
    auto aux_of = [mode, cap](uint32_t c) -> uint32_t {
        return mode == 1 ? ((c + SORT_CHUNK - 1) / SORT_CHUNK + RB - 1) / RB * RB
                         : dense_items(c, cap);
    };
    // up to RP entries per thread: loaded into registers together (one
    // memory latency rather than one per entry), 16 bytes at a time when the
    // thread's run is whole and 16-byte aligned
    constexpr int RP = 32;
    uint32_t vals[RP];
    const bool inreg = per <= RP;
    const bool vec = inreg && per % 4 == 0 && i1 - i0 == per && ((uintptr_t)(v + i0) & 15) == 0;
    uint32_t ca = 0, cb = 0;
    if (vec) {
#pragma unroll
        for (int q = 0; q < RP; q += 4) {
            uint4 w = make_uint4(0u, 0u, 0u, 0u);
            if (q < per) w = *(const uint4 *)(v + i0 + q);
            vals[q] = w.x;
            vals[q + 1] = w.y;
            vals[q + 2] = w.z;
            vals[q + 3] = w.w;
        }
    } else if (inreg) {
#pragma unroll
        for (int q = 0; q < RP; ++q) vals[q] = i0 + q < i1 ? v[i0 + q] : 0u;
    }
    if (inreg) {
#pragma unroll
        for (int q = 0; q < RP; ++q) {
            ca += vals[q];
            cb += i0 + q < i1 ? aux_of(vals[q]) : 0u;
        }
    } else {
        for (long i = i0; i < i1; ++i) {
            const uint32_t c = v[i];
            ca += c;
            cb += aux_of(c);
        }
    }
    // block-wide exclusive scan of the threads' (ca, cb): within each wave by
    // shuffles, then across the waves' totals
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t ia = ca, ib = cb;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t xa = __shfl_up(ia, o), xb = __shfl_up(ib, o);
        if (lane >= o) {
            ia += xa;
            ib += xb;
        }
    }
    if (lane == 63) {
        sa[wv] = ia;
        sb[wv] = ib;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t wa = lane < NW ? sa[lane] : 0u, wb = lane < NW ? sb[lane] : 0u;
        const uint32_t ea = wa, eb = wb;
#pragma unroll
        for (int o = 1; o < NW; o <<= 1) {
            const uint32_t xa = __shfl_up(wa, o), xb = __shfl_up(wb, o);
            if (lane >= o) {
                wa += xa;
                wb += xb;
            }
        }
        if (lane < NW) {
            sa[lane] = wa - ea;                  // exclusive prefix of the waves
            sb[lane] = wb - eb;
        }
        if (lane == NW - 1) {
            ta = wa;
            tb = wb;
        }
    }
    __syncthreads();
    uint32_t ra = sa[wv] + ia - ca, rb = sb[wv] + ib - cb;
    const bool vst = vec && aux_blk == nullptr &&
                     (((uintptr_t)(base + i0) | (uintptr_t)(aux + i0)) & 15) == 0;
    if (vst) {
#pragma unroll
        for (int q = 0; q < RP; q += 4) {
            if (q >= per) break;
            uint4 bo, ao;
            bo.x = ra; ao.x = rb; ra += vals[q]; rb += aux_of(vals[q]);
            bo.y = ra; ao.y = rb; ra += vals[q + 1]; rb += aux_of(vals[q + 1]);
            bo.z = ra; ao.z = rb; ra += vals[q + 2]; rb += aux_of(vals[q + 2]);
            bo.w = ra; ao.w = rb; ra += vals[q + 3]; rb += aux_of(vals[q + 3]);
            *(uint4 *)(base + i0 + q) = bo;
            *(uint4 *)(aux + i0 + q) = ao;
        }
    } else if (inreg) {
#pragma unroll
        for (int q = 0; q < RP; ++q) {
            if (i0 + q >= i1) break;
            base[i0 + q] = ra;
            aux[i0 + q] = rb;
            if (aux_blk) aux_blk[i0 + q] = rb / RB;
            ra += vals[q];
            rb += aux_of(vals[q]);
        }
    } else {
        for (long i = i0; i < i1; ++i) {
            base[i] = ra;
            aux[i] = rb;
            if (aux_blk) aux_blk[i] = rb / RB;
            const uint32_t c = v[i];
            ra += c;
            rb += aux_of(c);
        }
    }
    if (threadIdx.x == SCAN_THREADS - 1) {
        base[n] = ta;
        aux[n] = tb;
        if (aux_blk) aux_blk[n] = tb / RB;
    }
}

// O[j][d] = base[bidx(s, d)] + B[b][d] + sum of H[j'][d] for the earlier
// chunks j' of block b (s = segment of block b), with
// bidx(s, d) = (s / spj) * jst + (s % spj) * nd + d: spj segments per job,
// jst base entries per job (one job: spj = nseg)
__global__ void k_colscan(const uint32_t *__restrict__ H, const uint32_t *__restrict__ blk0,
                          int nseg, long nblk, int nd, const uint32_t *__restrict__ base,
                          const uint32_t *__restrict__ B, uint32_t *__restrict__ O, int spj,
                          long jst) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= nblk * nd) return;
    const long b = t / nd;
    const int d = (int)(t - b * nd);
    if (b >= (long)blk0[nseg]) return;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((long)blk0[mid] <= b) lo = mid; else hi = mid - 1;
    }
    uint32_t run = base[(long)(lo / spj) * jst + (long)(lo % spj) * nd + d] + B[t];
    for (int r = 0; r < RB; ++r) {
        const long j = b * RB + r;
        O[j * nd + d] = run;
        run += H[j * nd + d];
    }
}

// Wide two-level sort, level-2 slots: per coarse bucket b (one workgroup),
// exclusive scans of its tile counts T2[b*nd + s], s < nd, into tilebase
// (from the bucket's first sorted position base1[b]) and paint items
// (dense_items, from 0), and the bucket's item total into bitem[b]
__global__ void __launch_bounds__(SCAN_THREADS)
k_bucket_slots(const uint32_t *__restrict__ T2, const uint32_t *__restrict__ base1, int nb1,
               int nd, int cap, uint32_t *__restrict__ tilebase, uint32_t *__restrict__ items,
               uint32_t *__restrict__ bitem) {
    __shared__ uint32_t sa[SCAN_THREADS], sb[SCAN_THREADS];
    const int b = blockIdx.x;
    const uint32_t *v = T2 + (long)b * nd;
    uint32_t *base = tilebase + (long)b * nd, *aux = items + (long)b * nd;
    const uint32_t b0 = base1[b];
    const int per = (nd + SCAN_THREADS - 1) / SCAN_THREADS;
    const int i0 = threadIdx.x * per, i1 = i0 + per < nd ? i0 + per : nd;
    uint32_t ca = 0, cb = 0;
    for (int i = i0; i < i1; ++i) {
        const uint32_t c = v[i];
        ca += c;
        cb += dense_items(c, cap);
    }
    sa[threadIdx.x] = ca;
    sb[threadIdx.x] = cb;
    __syncthreads();
    for (int o = 1; o < SCAN_THREADS; o <<= 1) {
        const uint32_t va = threadIdx.x >= o ? sa[threadIdx.x - o] : 0;
        const uint32_t vb = threadIdx.x >= o ? sb[threadIdx.x - o] : 0;
        __syncthreads();
        sa[threadIdx.x] += va;
        sb[threadIdx.x] += vb;
        __syncthreads();
    }
    uint32_t ra = b0 + sa[threadIdx.x] - ca, rb = sb[threadIdx.x] - cb;
    for (int i = i0; i < i1; ++i) {
        base[i] = ra;
        aux[i] = rb;
        const uint32_t c = v[i];
        ra += c;
        rb += dense_items(c, cap);
    }
    if (threadIdx.x == SCAN_THREADS - 1) {
        bitem[b] = sb[SCAN_THREADS - 1];
        if (b == nb1 - 1) tilebase[(long)nb1 * nd] = b0 + sa[SCAN_THREADS - 1];
    }
}

// items[e] += ibase[e / nd] for the nb1 * nd slots; items[nb1 * nd] = total
__global__ void k_add_item_base(uint32_t *__restrict__ items, const uint32_t *__restrict__ ibase,
                                int nb1, int nd) {
    const long n = (long)nb1 * nd;
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e < n) items[e] += ibase[e / nd];
    if (e == n) items[n] = ibase[nb1];
}

// the single level-1 segment: blocks [0, nblk)
__global__ void k_set_seg1(uint32_t *seg, uint32_t nblk) {
    if (threadIdx.x == 0) {
        seg[0] = 0;
        seg[1] = nblk;
    }
}

// single-level path (grids with more tiles than two levels address): the
// rank of a particle within its tile comes from a global counter per tile,
// one atomic per distinct tile per wave (the wave-level match keeps
// clustered catalogs from serialising on a few counters); every lane of a
// wave calls wave_rank, valid or not
__global__ void __launch_bounds__(256)
k_count_global(const float *__restrict__ x, const float *__restrict__ y,
               const float *__restrict__ z, long Np, KeyGeo g, uint32_t *__restrict__ tot,
               uint32_t *__restrict__ rank) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const bool valid = p < Np;
    uint32_t t = 0, fine, ext;
    uint64_t pk;
    if (valid) particle_key(x[p], y[p], z[p], g, t, fine, pk, ext);
    const uint32_t r = wave_rank<24>(t, valid, tot);
    if (valid) rank[p] = r;
}

__global__ void k_scatter_global(const float *__restrict__ x, const float *__restrict__ y,
                                 const float *__restrict__ z, long Np, KeyGeo g,
                                 const uint32_t *__restrict__ base,
                                 const uint32_t *__restrict__ rank, uint64_t *__restrict__ out,
                                 uint8_t *__restrict__ out_ext) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= Np) return;
    uint32_t t, fine, ext;
    uint64_t pk;
    particle_key(x[p], y[p], z[p], g, t, fine, pk, ext);
    const uint32_t q = base[t] + rank[p];
    out[q] = pk;
    out_ext[q] = (uint8_t)ext;
}

// Six times the PCS weights of src/fft.c:83-88 in FP32 (they sum to 6).
// A term of the fixed-point paint is (6wx)(6wy)(6wz) times the tile's
// integer unit C_S (below), so a particle paints 216 C_S units: the three
// factors 1/6 never round, and the unit is exactly known.
__device__ __forceinline__ void pcs_w6_f32(float d, float w[4]) {
    const float e = 1.0f - d;
    const float d2 = d * d;
    w[0] = e * e * e;
    w[1] = (3.0f * d - 6.0f) * d2 + 4.0f;
    w[2] = ((-3.0f * d + 3.0f) * d + 3.0f) * d + 1.0f;
    w[3] = d2 * d;
}

// o[q] = w[(q + r) & 3] as a two-level barrel shift with static indices
// (a select on a lane-dependent index would be lowered to a scratch array)
__device__ __forceinline__ void rot4(const float w[4], int r, float o[4]) {
    const bool r0 = r & 1, r1 = r & 2;
    const float t0 = r0 ? w[1] : w[0], t1 = r0 ? w[2] : w[1];
    const float t2 = r0 ? w[3] : w[2], t3 = r0 ? w[0] : w[3];
    o[0] = r1 ? t2 : t0;
    o[1] = r1 ? t3 : t1;
    o[2] = r1 ? t0 : t2;
    o[3] = r1 ? t1 : t3;
}

// Fixed point of the tile paint.  An LDS footprint that takes terms from at
// most n particles paints in units of 1 / (216 C_S) particle,
// C_S = 2427 * 2^(S - 19) (2427 = floor(2^19 / 216)), with the largest
// S <= 25 for which n * (64 C_S + 1) < 2^31: a footprint cell gets at most
// one term of at most 64 C_S + 1 units per particle, so the int32 LDS sums
// cannot overflow.  n is at most 2 cap = 16384 (a pencil segment's two
// neighbouring tiles): S = 18 there, 19 for a full dense item of 8192
// particles, 25 for at most 216 particles.  The int64 grids hold units of
// 1 / (216 C_25) particle; a footprint is shifted left by 25 - S when it is
// flushed, exactly.  S depends only on particle counts, so the grids stay
// bit-identical run to run.
constexpr int FIX_S0 = 19, FIX_SMIN = 13, FIX_SMAX = 25;
constexpr float FIX_C0 = 2427.0f;
constexpr long FIX_TERM0 = 64l * 2427l;   // 64 C_19 = 2^6 * 2427: 64 C_S is exact for S >= 13
__host__ __device__ constexpr long fix_limit(int S) {
    return (2147483647l - 1) /
           ((S >= FIX_S0 ? FIX_TERM0 << (S - FIX_S0) : FIX_TERM0 >> (FIX_S0 - S)) + 1);
}

__host__ __device__ __forceinline__ int tile_bits(long n) {
    int S = FIX_SMAX;
    while (S > FIX_SMIN && n > fix_limit(S)) --S;
    return S;
}

// floor(a*b + u) for a*b >= 0, u in [0, 1), a*b + 1 < 2^24: one FMA and
// one truncating convert (truncation is floor for non-negative values).
// With u uniform and independent of a*b this rounds a*b up with probability
// frac(a*b) and down otherwise, so the expected rounded value is a*b
// exactly: no term, however small, is lost on average.
__device__ __forceinline__ int fixed_dither(float a, float b, float u) {
    return (int)__builtin_fmaf(a, b, u);
}

// Dither of one stencil row (the 4 cells along k of one (i, j) node pair):
// frac(u0 + 97 wab) with u0 the particle's dither base and wab the row's
// x-y weight product (at most 16), so the rows of a particle get different,
// effectively independent offsets from a function of the particle and the
// row alone (not of the lane or the visiting order).  Sharing one offset
// among all 64 terms instead makes their rounding errors add coherently:
// about 17 units of noise per particle at low k against 4.9 per row.
__device__ __forceinline__ float row_dither(float u0, float wab) {
    return __builtin_amdgcn_fractf(__builtin_fmaf(wab, 97.0f, u0));
}

// Dither bases of one particle for its plain and dual stencils: two 24-bit
// uniforms in [0, 1) from a 32-bit finaliser hash (MurmurHash3's fmix32,
// Appleby 2011) of its packed sort word and its tile slot (the word holds
// only tile-local coordinates).  They depend only on the particle's
// quantised position, so the grids stay bit-identical run to run.
__device__ __forceinline__ void dither_pair(uint64_t pk, uint32_t ext, uint32_t tile, float &ua,
                                            float &ub) {
    uint32_t h = (uint32_t)pk ^ ((uint32_t)(pk >> 32) * 0x9E3779B9u) ^ (tile * 0x7FEB352Du) ^
                 (ext * 0x2545F491u);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    const uint32_t h2 = (h ^ (h >> 15)) * 0x2C1B3C6Du;
    ua = (float)(h >> 8) * 0x1p-24f;
    ub = (float)(h2 >> 8) * 0x1p-24f;
}

// LDS footprints are stored in 4x4x4 bricks: cell (i,j,k) of a footprint
// with NBK bricks per side sits at word
//   (((i>>2)*NBK + (j>>2))*NBK + (k>>2))*64 + 16(i&3) + 4(j&3) + (k&3),
// so its LDS bank (word mod 64) is 16(i&3) + 4(j&3) + (k&3): the 64 cells
// of any 4x4x4 stencil lie in 64 distinct banks, wherever its base.
constexpr int brick_n(int side) { return (side + 3) / 4; }
constexpr int brick_words(int side) { return brick_n(side) * brick_n(side) * brick_n(side) * 64; }

// k lanes per row of a footprint flush: whole aligned 128-byte rows of 16
// int64 cells for tiles of 12 and 16, halves and quarters for smaller ones
constexpr int flush_klanes(int side) { return side >= 15 ? 16 : side >= 8 ? 8 : 4; }

// Flush the planes k < nz of a bricked LDS footprint of SIDE^2 rows along k,
// whose cell (0,0,0) sits at grid cell (oi, oj, ok), into the int64 grid
// with one integer atomic per non-zero cell.  (Plain stores of the cells no
// other tile reaches were measured twice as slow: they share cache lines
// with the atomics.)  A wave step covers RPG = 64 / KL rows (fj) of one i
// plane (fi) and KL consecutive k cells, so with KL = T = 16 and ok a
// multiple of 16 every row is one aligned 128-byte line (the memory-side
// atomic unit charges per line as well as per atomic); planes past the
// first KL take a second step.  The plane and row group of a wave step are
// wave-uniform (scalar), and since RPG is a multiple of 4 the rows of a
// group start on a brick boundary: a lane's LDS word is a scalar base plus
// a per-lane constant, and only the j and k wraps are per lane.  Sums are
// non-negative (every term is), so the shift to the common unit is a
// zero-extended one.
template <int SIDE>
__device__ void flush_planes(const int *__restrict__ fp, int nz, int oi, int oj, int ok, int Ng,
                             int shift, unsigned long long *__restrict__ G) {
    constexpr int KL = flush_klanes(SIDE), RPG = 64 / KL;
    static_assert(RPG % 4 == 0, "row groups start on brick boundaries");
    constexpr int NBK = brick_n(SIDE);
    constexpr int NJG = (SIDE + RPG - 1) / RPG;   // row groups per i plane
    typedef __attribute__((address_space(1))) unsigned long long gu64;
    const int lane = __lane_id();
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwv = blockDim.x >> 6;
    const int ro = lane / KL, kl = lane - ro * KL;
    // a footprint ends at most SIDE cells past the grid's last index; grids
    // smaller than that wrap more than once
    auto wrap = [Ng](int v) {
        v = v >= Ng ? v - Ng : v;
        if ((unsigned)v >= (unsigned)Ng) v = ((v % Ng) + Ng) % Ng;
        return v;
    };
    const bool one_wrap = Ng >= SIDE;
    const int lane_row = (ro >> 2) * NBK * 64 + (ro & 3) * 4;
    for (int kb = 0; kb < nz; kb += KL) {
        const int fk = kb + kl;
        const bool kin = fk < nz;
        const int gk = one_wrap ? (ok + fk >= Ng ? ok + fk - Ng : ok + fk) : wrap(ok + fk);
        const int *fpl = fp + lane_row + (fk >> 2) * 64 + (fk & 3);
        gu64 *gcol = (gu64 *)G + gk;
        for (int p = wv; p < SIDE * NJG; p += nwv) {
            const int pu = __builtin_amdgcn_readfirstlane(p);   // wave-uniform
            const int fi = pu / NJG, s = pu - fi * NJG;
            const int fj = s * RPG + ro;
            if (!kin || fj >= SIDE) continue;
            const int v = fpl[((fi >> 2) * NBK + s * (RPG / 4)) * NBK * 64 + (fi & 3) * 16];
            if (!v) continue;
            const int gi = wrap(oi + fi);
            const int gj = one_wrap ? (oj + fj >= Ng ? oj + fj - Ng : oj + fj) : wrap(oj + fj);
            // Ng < 2^16: the row index fits 32 bits
            const uint32_t row = (uint32_t)gi * (uint32_t)Ng + (uint32_t)gj;
            __hip_atomic_fetch_add(gcol + (uint64_t)row * (uint32_t)Ng,
                                   (unsigned long long)(uint32_t)v << shift, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// After a tile of a pencil is flushed (planes k < T): move its planes
// T .. SIDE-1, the last brick layer (T is a multiple of 4), to layer 0, where
// they overlap the next tile along k, and clear the other layers.
template <int SIDE, int T>
__device__ void carry_down(int *__restrict__ fp) {
    constexpr int NBK = brick_n(SIDE), L = T / 4;
    static_assert(T % 4 == 0 && NBK == L + 1, "carry layer");
    for (int w = threadIdx.x; w < NBK * NBK * 64; w += blockDim.x) {
        int *col = fp + (w >> 6) * NBK * 64 + (w & 63);   // brick column (i, j), element
        col[0] = col[L * 64];
#pragma unroll
        for (int l = 1; l <= L; ++l) col[l * 64] = 0;
    }
}

// paint work item -> tile slot and particle range [p0, p1)
__device__ __forceinline__ bool item_range(long it, const uint32_t *__restrict__ tilebase,
                                           const uint32_t *__restrict__ items, long nslots, int cap,
                                           long &t, long &p0, long &p1) {
    if (it >= (long)items[nslots]) return false;
    long lo = 0, hi = nslots - 1;          // last slot with items[t] <= it
    while (lo < hi) {
        const long mid = (lo + hi + 1) >> 1;
        if ((long)items[mid] <= it) lo = mid; else hi = mid - 1;
    }
    t = lo;
    p0 = (long)tilebase[t] + (it - (long)items[t]) * cap;
    const long pe = tilebase[t + 1];
    p1 = p0 + cap < pe ? p0 + cap : pe;
    return true;
}

// Offsets of the 4 cells of one stencil axis, in the order a lane visits
// them: step q goes to the cell of residue r = (q + rl) & 3 (rl: the lane's
// 2 bits of this axis), i.e. the first cell >= base congruent to r, which
// lies in the base's brick when r >= base & 3 and in the next one otherwise.
// w[] is rotated to match: out[q] = w[(r - base) & 3].
__device__ __forceinline__ void brick_axis(int base, int rl, int stride, int unit,
                                           const float w[4], int off[4], float out[4]) {
    const int b0 = (base >> 2) * stride, s = base & 3;
    rot4(w, (rl - base) & 3, out);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = (q + rl) & 3;
        off[q] = b0 + (r < s ? stride : 0) + r * unit;
    }
}

// One particle's 4x4x4 stencil with base (bi,bj,bk) into a bricked LDS
// footprint of NBK bricks per side.  Lane l = 16ra + 4rb + rc adds at step
// (aa,bb,cc) to the stencil cell with residues ((aa+ra)&3, (bb+rb)&3,
// (cc+rc)&3): the bank 16((aa+ra)&3) + 4((bb+rb)&3) + ((cc+rc)&3) is a
// bijection of the lane's bits, so every step's 64 atomics fall in 64
// distinct banks, for any mix of particles in the wave.
template <int NBK>
__device__ __forceinline__ void stencil_lds(int *__restrict__ fp, int bi, int bj, int bk,
                                            const float wx[4], const float wy[4],
                                            const float wzs[4], float u, int lane) {
    float xr[4], yr[4], zr[4];
    int ox[4], oy[4], oz[4];
    brick_axis(bi, (lane >> 4) & 3, NBK * NBK * 64, 16, wx, ox, xr);
    brick_axis(bj, (lane >> 2) & 3, NBK * 64, 4, wy, oy, yr);
    brick_axis(bk, lane & 3, 64, 1, wzs, oz, zr);
#pragma unroll
    for (int aa = 0; aa < 4; ++aa) {
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const float wab = xr[aa] * yr[bb];
            const float ur = row_dither(u, wab);
            int *row = fp + ox[aa] + oy[bb];
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) atomicAdd(row + oz[cc], fixed_dither(wab, zr[cc], ur));
        }
    }
}

// tile-local cell li and fraction d of one axis of a packed particle
// (main word and extension, F fraction bits; exact in FP32 for F <= 24)
__device__ __forceinline__ void unpack_axis(uint32_t q, int F, float fsc, int &li, float &d) {
    li = (int)(q >> F);
    d = (float)(q & ((1u << F) - 1u)) * fsc;
}

__device__ __forceinline__ uint32_t pk_axis(uint64_t pk, uint32_t ext, int a) {
    const uint32_t hi = (uint32_t)(pk >> ((2 - a) * POS_MAIN)) & POS_MASK;
    return (hi << POS_EXT) | ((ext >> ((2 - a) * POS_EXT)) & ((1u << POS_EXT) - 1u));
}

// one grid's stencil of one packed particle word into an LDS footprint whose
// cell (0,0,0) is the tile's origin: DUAL = 1, base floor(v), fraction
// frac(v); DUAL = 0 (plain grid), base floor(v + 1/2), fraction frac(v + 1/2)
template <int NBK, int DUAL>
__device__ __forceinline__ void paint_grid(int *__restrict__ fp, uint64_t pk, uint32_t ext, int F,
                                           float fsc, float unit, float u, int lane) {
    int li[3];
    float d[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        uint32_t q = pk_axis(pk, ext, a);
        if (!DUAL) q += 1u << (F - 1);
        unpack_axis(q, F, fsc, li[a], d[a]);
    }
    float wx[4], wy[4], wz[4], wzs[4];
    pcs_w6_f32(d[0], wx);
    pcs_w6_f32(d[1], wy);
    pcs_w6_f32(d[2], wz);
#pragma unroll
    for (int c = 0; c < 4; ++c) wzs[c] = wz[c] * unit;
    stencil_lds<NBK>(fp, li[0], li[1], li[2], wx, wy, wzs, u, lane);
}

// particles [p0, p1) of tile slot t into the plain (fa) and dual (fb) LDS
// footprints, one particle per thread per step (coalesced loads)
template <int SA, int SB>
__device__ __forceinline__ void paint_range(const uint64_t *__restrict__ sorted,
                                            const uint8_t *__restrict__ sorted_ext, long p0,
                                            long p1, long t, int *fa, int *fb, int interlace,
                                            int F, float unit) {
    const float fsc = 1.0f / (float)(1 << F);
    const int lane = __lane_id();
    for (long p = p0 + threadIdx.x; p < p1; p += PAINT_THREADS) {
        const uint64_t pk = sorted[p];
        const uint32_t ext = sorted_ext[p];
        float ua, ub;
        dither_pair(pk, ext, (uint32_t)t, ua, ub);
        paint_grid<brick_n(SA), 0>(fa, pk, ext, F, fsc, unit, ua, lane);
        if (interlace) paint_grid<brick_n(SB), 1>(fb, pk, ext, F, fsc, unit, ub, lane);
    }
}

__device__ __forceinline__ float fix_unit(int S) { return ldexpf(FIX_C0, S - FIX_S0); }

// Paint of the sorted particles into the int64 grids.  The plain footprint
// of a tile of side T has T + 4 cells per side, the dual one T + 3, both
// starting at the tile's origin (see axis_key); they are int32 fixed point
// in LDS, flushed with integer atomics.  A launch holds two kinds of work:
//   * pencil segments (work index w < nb1 * cg0 * cg1): the cg2 tiles along
//     k of one coarse bucket and (i, j) tile column, in order.  After each
//     tile only its planes k < T are flushed (whole aligned rows); the planes
//     above are carried down in LDS into the next tile's footprint, and the
//     segment's last tile flushes its whole footprints.  Tiles of more than
//     cap particles are left to the second kind (the segment flushes only
//     what it carries into them).  All tiles of a segment share one fixed
//     point, chosen from the largest particle count of two neighbouring
//     tiles: a footprint cell gets terms from particles of its tile and of
//     the one below it.
//   * dense items (w >= that count): <= cap particles of a tile with more
//     than cap particles (items: exclusive prefix of those items over the
//     tile slots); whole footprints flushed; fixed point from min(count, cap).
// zbuf (one plan): every launched workgroup first clears its share of
// zwords words there.
// zbuf (batch): the launches of one paint stage clear consecutive pieces.
// Batched (jobs != nullptr): workgroup b does work ji0[ja] + b of the batch;
// the job owning it (ji0[job] <= work < ji0[job+1], ja <= job < jb) supplies
// the geometry, grids (set 0: GA/GB, set 1: GA2/GB2) and tile tables
// (tilebase, items + job * sl, MAX_D2 slots per coarse bucket).
// (k_paint_tiles<T> runs this body; k_paint_tiles_multi runs the body of
// each batch catalog's tile side in one launch)
template <int T>
__device__ __forceinline__ void
paint_tiles_body(const uint64_t *__restrict__ sorted, const uint8_t *__restrict__ sorted_ext,
                 const uint32_t *__restrict__ tilebase,
                 const uint32_t *__restrict__ items, long nslots, KeyGeo g, int nb2,
                 int cap, int interlace, unsigned long long *__restrict__ GA,
                 unsigned long long *__restrict__ GB, const PaintJob *__restrict__ jobs,
                 const long *__restrict__ ji0, int ja, int jb, long sl, int set,
                 uint64_t *__restrict__ zbuf, long zwords) {
    constexpr int SA = T + 4, SB = T + 3;
    constexpr bool CARRY = T % 4 == 0;
    if (zbuf) {   // this workgroup's slice of the memory to clear (16-byte stores)
        const long per = ((zwords + gridDim.x - 1) / gridDim.x + 1) & ~1l;
        const long z0 = blockIdx.x * per, z1 = z0 + per < zwords ? z0 + per : zwords;
        for (long q = z0 + 2 * threadIdx.x; q + 1 < z1; q += 2 * blockDim.x)
            *(ulonglong2 *)(zbuf + q) = make_ulonglong2(0ull, 0ull);
    }
    constexpr int NA = brick_words(SA), NB = brick_words(SB);
    extern __shared__ int lds[];
    long w = blockIdx.x;
    if (jobs) {
        const long e = ji0[ja] + blockIdx.x;
        const int job = ja + batch_job(e, ji0 + ja, jb - ja);
        const PaintJob &J = jobs[job];
        w = e - ji0[job];
        tilebase += job * sl;
        items += job * sl;
        nslots = (long)J.nb1 * MAX_D2;
        g = J.g;
        nb2 = MAX_D2;
        interlace = J.interlace;
        GA = set ? J.GA2 : J.GA;
        GB = set ? J.GB2 : J.GB;
    }
    const int nl = NA + (interlace ? NB : 0);
    int *fa = lds, *fb = lds + NA;
    const int pen = g.cg0 * g.cg1;
    const long nseg = nslots / nb2 * pen;
    // grid origin of the tile in slot t
    auto tile_origin = [&g, nb2](long t, int &oi, int &oj, int &ok) {
        const long b1 = t / nb2, f = t - b1 * nb2;
        const int cx = (int)(b1 / ((long)g.nc1 * g.nc2)), cy = (int)((b1 / g.nc2) % g.nc1),
                  cz = (int)(b1 % g.nc2);
        const int fx = (int)(f / (g.cg1 * g.cg2)), fy = (int)((f / g.cg2) % g.cg1),
                  fz = (int)(f % g.cg2);
        oi = (cx * g.cg0 + fx) * T;
        oj = (cy * g.cg1 + fy) * T;
        ok = (cz * g.cg2 + fz) * T;
    };
    if (w < nseg) {
        const long b1 = w / pen;
        const long s0 = b1 * nb2 + ((w - b1 * pen) << g.lg2);   // the segment's first slot
        int oi, oj, ok0;
        tile_origin(s0, oi, oj, ok0);
        if (oi >= g.nt * T || oj >= g.nt * T) return;   // slots past the grid: empty
        const int nz = g.cg2 < g.nt - ok0 / T ? g.cg2 : g.nt - ok0 / T;
        long m = 0, prev = 0;
        for (int q = 0; q < nz; ++q) {
            long c = (long)tilebase[s0 + q + 1] - (long)tilebase[s0 + q];
            if (c > cap) c = 0;
            if (prev + c > m) m = prev + c;
            prev = c;
        }
        if (m == 0) return;   // nothing to paint here (dense tiles are items)
        const int S = tile_bits(m);
        const float unit = fix_unit(S);
        for (int i = threadIdx.x; i < nl; i += blockDim.x) lds[i] = 0;
        __syncthreads();
        bool carry = false;   // the carried planes may be non-zero
        for (int q = 0; q < nz; ++q) {
            const long t = s0 + q;
            const long p0 = tilebase[t], c = (long)tilebase[t + 1] - p0;
            const bool painted = c > 0 && c <= cap;
            if (!painted && !carry) continue;   // footprints all zero
            if (painted)
                paint_range<SA, SB>(sorted, sorted_ext, p0, p0 + c, t, fa, fb, interlace, g.F,
                                    unit);
            __syncthreads();
            const int ok = ok0 + q * T;
            if (!CARRY || q == nz - 1) {
                flush_planes<SA>(fa, SA, oi, oj, ok, g.Ng, FIX_SMAX - S, GA);
                if (interlace) flush_planes<SB>(fb, SB, oi, oj, ok, g.Ng, FIX_SMAX - S, GB);
                if (q == nz - 1) break;
                __syncthreads();
                for (int i = threadIdx.x; i < nl; i += blockDim.x) lds[i] = 0;
            } else {
                flush_planes<SA>(fa, T, oi, oj, ok, g.Ng, FIX_SMAX - S, GA);
                if (interlace) flush_planes<SB>(fb, T, oi, oj, ok, g.Ng, FIX_SMAX - S, GB);
                __syncthreads();
                if constexpr (CARRY) {
                    carry_down<SA, T>(fa);
                    if (interlace) carry_down<SB, T>(fb);
                }
            }
            __syncthreads();
            carry = CARRY && painted;
        }
        return;
    }
    long t, p0, p1;
    if (!item_range(w - nseg, tilebase, items, nslots, cap, t, p0, p1)) return;
    for (int i = threadIdx.x; i < nl; i += blockDim.x) lds[i] = 0;
    __syncthreads();
    // the tile's fixed point, from its particle count (all its items agree)
    const long ntile = (long)tilebase[t + 1] - (long)tilebase[t];
    const int S = tile_bits(ntile < cap ? ntile : cap);
    paint_range<SA, SB>(sorted, sorted_ext, p0, p1, t, fa, fb, interlace, g.F, fix_unit(S));
    __syncthreads();
    int oi, oj, ok;
    tile_origin(t, oi, oj, ok);
    flush_planes<SA>(fa, SA, oi, oj, ok, g.Ng, FIX_SMAX - S, GA);
    if (interlace) flush_planes<SB>(fb, SB, oi, oj, ok, g.Ng, FIX_SMAX - S, GB);
}

template <int T>
__global__ void __launch_bounds__(PAINT_THREADS, 8)
k_paint_tiles(const uint64_t *__restrict__ sorted, const uint8_t *__restrict__ sorted_ext,
              const uint32_t *__restrict__ tilebase,
              const uint32_t *__restrict__ items, long nslots, KeyGeo g, int nb2,
              int cap, int interlace, unsigned long long *__restrict__ GA,
              unsigned long long *__restrict__ GB, const PaintJob *__restrict__ jobs,
              const long *__restrict__ ji0, int ja, int jb, long sl, int set,
              uint64_t *__restrict__ zbuf, long zwords) {
    paint_tiles_body<T>(sorted, sorted_ext, tilebase, items, nslots, g, nb2, cap, interlace, GA,
                        GB, jobs, ji0, ja, jb, sl, set, zbuf, zwords);
}

// One paint launch for every catalog of a batch whatever its tile side:
// workgroup b does work b of the batch (ji0[0] = 0), and the side of the
// catalog owning it (its KeyGeo) selects the compile-time body; the
// dynamic LDS is that of the largest side in the launch.
__global__ void __launch_bounds__(PAINT_THREADS, 8)
k_paint_tiles_multi(const uint64_t *__restrict__ sorted, const uint8_t *__restrict__ sorted_ext,
                    const uint32_t *__restrict__ tilebase, const uint32_t *__restrict__ items,
                    int cap, const PaintJob *__restrict__ jobs, const long *__restrict__ ji0,
                    int njobs, long sl, int set, uint64_t *__restrict__ zbuf, long zwords) {
    const int job = batch_job(blockIdx.x, ji0, njobs);
    const KeyGeo g0 = jobs[job].g;
    switch (g0.T) {
#define PAINT_MULTI_T(TT)                                                                     \
    case TT:                                                                                  \
        paint_tiles_body<TT>(sorted, sorted_ext, tilebase, items, 0l, g0, MAX_D2, cap, 0,     \
                             (unsigned long long *)nullptr, (unsigned long long *)nullptr,    \
                             jobs, ji0, 0, njobs, sl, set, zbuf, zwords);                     \
        break;
        PAINT_MULTI_T(16)
        PAINT_MULTI_T(12)
        PAINT_MULTI_T(8)
#undef PAINT_MULTI_T
    default:
        break;
    }
}

// ------------------------------------------------------------------ host

// the bricked LDS footprints of one work item
static int paint_lds_bytes(int T, int interlace) {
    return (brick_words(T + 4) + (interlace ? brick_words(T + 3) : 0)) * 4;
}

static int ceil_log2(long v) {
    int b = 0;
    while ((1l << b) < v) ++b;
    return b;
}

static long round_up(long v, long m) { return (v + m - 1) / m * m; }

// tile sides with a compiled paint kernel
static const int kSides[] = {16, 12, 8, 4, 2, 1};

static int paint_side(int Ng, int tile) {
    int T = tile > 0 ? tile : 16;
    if (T > Ng) T = Ng;
    for (int i = 0; i < 6; ++i)
        if (kSides[i] <= T) return kSides[i];
    return 1;
}

// fixed-point fraction bits S of the weights for a catalog whose work
// items hold at most n particles (n = min(cap, Np)): the largest S <= 25
// with n * ((2/3)^3 2^S + 1) < 2^31 (a footprint cell gets at most one
// dithered term of at most (2/3)^3 2^S + 1 per particle, so the int32 LDS
// sums cannot overflow) and (2/3)^3 2^S + 1 < 2^24 (FP32 holds every
// dithered sum exactly).  S = 19 for full 8192-particle items, 25 for a
// single particle.
// particle mass of one int64 grid unit: 1 / (216 C_25), exact in FP64
static double fixed_scale(void) { return 1.0 / (216.0 * (double)FIX_C0 * (double)(1 << (FIX_SMAX - FIX_S0))); }

// Tile geometry of a grid of side Ng painted in tiles of side T: coarse
// buckets group up to d2 tiles, doubling the group along the axis with the
// most buckets, so that small grids get few, large buckets (few level-2
// segments, long level-1 runs).  mode 1: two levels when at most MAX_D1
// buckets of MAX_D2 tiles cover the grid; mode 2 (allow_wide): two levels of
// WIDE_D1 x WIDE_D2; mode 0: every tile is its own bucket (single level).
static KeyGeo tile_geometry(int Ng, int T, int allow_wide, int &mode) {
    const int nt = (Ng + T - 1) / T;
    int cg[3] = {1, 1, 1}, nc[3] = {nt, nt, nt};
    mode = 0;
    for (int m = 1; m <= (allow_wide ? 2 : 1) && !mode; ++m) {
        const int d1 = m == 1 ? MAX_D1 : WIDE_D1, d2 = m == 1 ? MAX_D2 : WIDE_D2;
        cg[0] = cg[1] = cg[2] = 1;
        nc[0] = nc[1] = nc[2] = nt;
        while (cg[0] * cg[1] * cg[2] < d2 && (nc[0] > 1 || nc[1] > 1 || nc[2] > 1)) {
            int ax = 2;
            if (nc[1] > nc[ax]) ax = 1;
            if (nc[0] > nc[ax]) ax = 0;
            cg[ax] *= 2;
            nc[ax] = (nt + cg[ax] - 1) / cg[ax];
        }
        if ((long)nc[0] * nc[1] * nc[2] <= d1) mode = m;
    }
    if (!mode) {
        cg[0] = cg[1] = cg[2] = 1;
        nc[0] = nc[1] = nc[2] = nt;
    }
    const int F = POS_MAIN + POS_EXT - ceil_log2(T) < 24 ? POS_MAIN + POS_EXT - ceil_log2(T) : 24;
    // (the digit and extension share 16 bits between the sort levels)
    static_assert(AUX_SHIFT + 10 <= 16, "side word");
    KeyGeo g = {0.0,   Ng,    T,     nt,    F,
                cg[0], cg[1], cg[2], nc[0], nc[1],
                nc[2], ceil_log2(cg[0]), ceil_log2(cg[1]), ceil_log2(cg[2]),
                ((1ull << 32) + (uint64_t)T - 1) / (uint64_t)T,
                (T & (T - 1)) == 0 ? ceil_log2(T) : -1};
    return g;
}

extern "C" int fgk_paint_plan_create(fgk_paintplan **out, int Ng, int tile, long Np_max,
                                     int interlace) {
    *out = NULL;
    if (Ng < 2 || Ng > 65535) { fgk_set_error("paint: Ng must be in [2, 65535]"); return -2; }
    if (Np_max < 1) Np_max = 1;
    if (Np_max >= (1l << 32) - 1) { fgk_set_error("paint: at most 2^32-2 particles"); return -2; }
    const int T = paint_side(Ng, tile);
    fgk_paintplan *pp = (fgk_paintplan *)calloc(1, sizeof(fgk_paintplan));
    if (!pp) { fgk_set_error("out of host memory"); return -1; }
    pp->Ng = Ng;
    pp->T = T;
    pp->interlace = interlace ? 1 : 0;
    pp->cap = 8192;
    pp->Np_max = Np_max;
    int mode = 0;
    pp->g = tile_geometry(Ng, T, 1, mode);
    pp->two_level = mode != 0;
    pp->wide = mode == 2;
    pp->nt = pp->g.nt;
    pp->F = pp->g.F;
    if (!pp->two_level && (long)pp->nt * pp->nt * pp->nt >= (1l << 24)) {
        fgk_set_error("paint: %d^3 tiles of side %d exceed the 2^24 the sort addresses", pp->nt,
                      T);
        free(pp);
        return -2;
    }
    const int nc[3] = {pp->g.nc0, pp->g.nc1, pp->g.nc2}, cg[3] = {pp->g.cg0, pp->g.cg1, pp->g.cg2};
    pp->nb1 = nc[0] * nc[1] * nc[2];
    pp->nb2 = cg[0] * cg[1] * cg[2];
    pp->bits1 = ceil_log2(pp->nb1);
    pp->bits2 = ceil_log2(pp->nb2);
    pp->nslots = (long)pp->nb1 * pp->nb2;
    // one paint workgroup per work item; a dispatch has fewer than 2^32 work-items
    if ((paint_work(pp->nb1, pp->g, Np_max, pp->cap) + 8) * (long)PAINT_THREADS >= (1l << 32)) {
        fgk_set_error("paint: %ld tiles of side %d and %ld particles exceed one dispatch",
                      pp->nslots, T, Np_max);
        free(pp);
        return -2;
    }
    pp->nw1max = round_up((Np_max + SORT_CHUNK - 1) / SORT_CHUNK, RB);
    pp->nw2max = pp->nw1max + (long)pp->nb1 * RB;
    int rc = 0;
    const size_t u = sizeof(uint32_t);
    if (pp->two_level) {
        rc |= fgk_malloc((void **)&pp->H1, (size_t)pp->nw1max * pp->nb1 * u);
        rc |= fgk_malloc((void **)&pp->O1, (size_t)pp->nw1max * pp->nb1 * u);
        rc |= fgk_malloc((void **)&pp->S1, (size_t)(pp->nw1max / RB) * pp->nb1 * u);
        rc |= fgk_malloc((void **)&pp->B1, (size_t)(pp->nw1max / RB) * pp->nb1 * u);
        rc |= fgk_malloc((void **)&pp->T1, (size_t)pp->nb1 * u);
        rc |= fgk_malloc((void **)&pp->H2, (size_t)pp->nw2max * pp->nb2 * u);
        rc |= fgk_malloc((void **)&pp->O2, (size_t)pp->nw2max * pp->nb2 * u);
        rc |= fgk_malloc((void **)&pp->S2, (size_t)(pp->nw2max / RB) * pp->nb2 * u);
        rc |= fgk_malloc((void **)&pp->B2, (size_t)(pp->nw2max / RB) * pp->nb2 * u);
        rc |= fgk_malloc((void **)&pp->seg1, 2 * u);
        rc |= fgk_malloc((void **)&pp->base1, (size_t)(pp->nb1 + 1) * u);
        rc |= fgk_malloc((void **)&pp->chunk0, (size_t)(pp->nb1 + 1) * u);
        rc |= fgk_malloc((void **)&pp->blk0, (size_t)(pp->nb1 + 1) * u);
        if (pp->wide) {
            rc |= fgk_malloc((void **)&pp->bitem, (size_t)(pp->nb1 + 1) * u);
            rc |= fgk_malloc((void **)&pp->ibase, (size_t)(pp->nb1 + 1) * u);
            rc |= fgk_malloc((void **)&pp->bscratch, (size_t)(pp->nb1 + 1) * u);
        }
    } else {
        rc |= fgk_malloc((void **)&pp->rank32, (size_t)Np_max * u);
    }
    rc |= fgk_malloc((void **)&pp->T2, (size_t)pp->nslots * u);
    rc |= fgk_malloc((void **)&pp->tilebase, (size_t)(pp->nslots + 1) * u);
    rc |= fgk_malloc((void **)&pp->items, (size_t)(pp->nslots + 1) * u);
    if (pp->two_level) {
        rc |= fgk_malloc((void **)&pp->sorted1, (size_t)Np_max * sizeof(uint64_t));
        rc |= fgk_malloc((void **)&pp->aux1, (size_t)Np_max * sizeof(uint16_t));
    }
    rc |= fgk_malloc((void **)&pp->sorted2, (size_t)Np_max * sizeof(uint64_t));
    rc |= fgk_malloc((void **)&pp->ext2, (size_t)Np_max * sizeof(uint8_t));
    if (rc) {
        fgk_paint_plan_destroy(pp);
        return -1;
    }
    *out = pp;
    return 0;
}

extern "C" void fgk_paint_plan_destroy(fgk_paintplan *pp) {
    if (!pp) return;
    uint32_t *bufs[] = {pp->H1, pp->O1, pp->S1, pp->B1, pp->T1, pp->H2, pp->O2, pp->S2, pp->B2,
                        pp->seg1, pp->base1, pp->chunk0, pp->blk0, pp->T2, pp->tilebase,
                        pp->items, pp->rank32, pp->bitem, pp->ibase, pp->bscratch};
    for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; ++i) fgk_free(bufs[i]);
    fgk_free(pp->sorted1);
    fgk_free(pp->sorted2);
    fgk_free(pp->aux1);
    fgk_free(pp->ext2);
    free(pp);
}

extern "C" double fgk_paint_fixed_scale(void) { return fixed_scale(); }

extern "C" int fgk_paint_tile_bits(long n) { return tile_bits(n); }
extern "C" int fgk_paint_plan_tile(const fgk_paintplan *pp) { return pp->T; }

static const char *const kPaintStages[] = {"sort1_count", "sort1_scan", "sort1_scatter",
                                           "sort2_count", "sort2_scan", "sort2_scatter",
                                           "paint_tiles"};

extern "C" int fgk_paint_nstages(void) { return (int)(sizeof kPaintStages / sizeof kPaintStages[0]); }
extern "C" const char *fgk_paint_stage_name(int i) {
    return (i >= 0 && i < fgk_paint_nstages()) ? kPaintStages[i] : "";
}

static unsigned blocks_for(long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// scans of one level: H[nw][nd] histograms of nw chunks in nseg segments
// (segment s = blocks [blk0[s], blk0[s+1])) -> O[j][d] output offsets
static int level_scans(hipStream_t s, const uint32_t *H, long nw, int nd, uint32_t *S, uint32_t *B,
                       const uint32_t *blk0, int nseg, uint32_t *T, uint32_t *base, int mode,
                       int cap, uint32_t *aux, uint32_t *aux_blk, uint32_t *O, long nslots,
                       int spj, long jst) {
    const long nblk = nw / RB;
    hipLaunchKernelGGL(k_blocksum, dim3(blocks_for(nblk * nd, 256)), dim3(256), 0, s, H, nblk, nd,
                       S);
    FGK_LAUNCHED();
    if (nseg == 1)
        hipLaunchKernelGGL(k_segscan_wide, dim3(nd), dim3(256), 0, s, S, nblk, nd, B, T);
    else
        hipLaunchKernelGGL(k_segscan, dim3(blocks_for((long)nseg * nd, 256)), dim3(256), 0, s,
                           S, blk0, nseg, nd, B, T);
    FGK_LAUNCHED();
    hipLaunchKernelGGL(k_digitscan, dim3(1), dim3(SCAN_THREADS), 0, s, T, nslots, mode, cap, base,
                       aux, aux_blk);
    FGK_LAUNCHED();
    hipLaunchKernelGGL(k_colscan, dim3(blocks_for(nblk * nd, 256)), dim3(256), 0, s, H, blk0, nseg,
                       nblk, nd, base, B, O, spj, jst);
    FGK_LAUNCHED();
    return 0;
}

// k_sort of one level over nw chunks with the plan's digit width (standard:
// 8 / 7 bits, wide: 11 / 10 bits)
#define SORT_LAUNCH(LEVEL, COUNT, nw, ...)                                                    \
    do {                                                                                      \
        constexpr int nbs_ = LEVEL == 1 ? 8 : 7, nbw_ = LEVEL == 1 ? 11 : 10;                 \
        if (pp->wide)                                                                         \
            hipLaunchKernelGGL((k_sort<LEVEL, COUNT, nbw_>), dim3((unsigned)(nw)),            \
                               dim3(SORT_THREADS), 0, s, __VA_ARGS__,                         \
                               (const PaintJob *)nullptr, (const long *)nullptr, 0);          \
        else                                                                                  \
            hipLaunchKernelGGL((k_sort<LEVEL, COUNT, nbs_>), dim3((unsigned)(nw)),            \
                               dim3(SORT_THREADS), 0, s, __VA_ARGS__,                         \
                               (const PaintJob *)nullptr, (const long *)nullptr, 0);          \
        FGK_LAUNCHED();                                                                       \
    } while (0)

extern "C" int fgk_paint_stage(fgk_paintplan *pp, int stage, const float *x, const float *y,
                               const float *z, long Np, double Hinv, int64_t *gridA,
                               int64_t *gridB, fgk_stream_t st) {
    hipStream_t s = (hipStream_t)st;
    if (Np > pp->Np_max) {
        fgk_set_error("paint: Np=%ld exceeds the plan's Np_max=%ld", Np, pp->Np_max);
        return -2;
    }
    KeyGeo g = pp->g;
    g.Hinv = Hinv;
    const long nw1 = round_up((Np + SORT_CHUNK - 1) / SORT_CHUNK, RB);
    if (!pp->two_level) {
        switch (stage) {
        case 0:
            if (fgk_fill_zero(pp->T2, (size_t)pp->nslots * sizeof(uint32_t), st)) return -1;
            if (Np > 0) {
                hipLaunchKernelGGL(k_count_global, dim3(blocks_for(Np, 256)), dim3(256), 0, s, x,
                                   y, z, Np, g, pp->T2, pp->rank32);
                FGK_LAUNCHED();
            }
            return 0;
        case 1:
            hipLaunchKernelGGL(k_digitscan, dim3(1), dim3(SCAN_THREADS), 0, s, pp->T2, pp->nslots,
                               2, pp->cap, pp->tilebase, pp->items, (uint32_t *)nullptr);
            FGK_LAUNCHED();
            return 0;
        case 2:
            if (Np > 0) {
                hipLaunchKernelGGL(k_scatter_global, dim3(blocks_for(Np, 256)), dim3(256), 0, s,
                                   x, y, z, Np, g, pp->tilebase, pp->rank32, pp->sorted2,
                                   pp->ext2);
                FGK_LAUNCHED();
            }
            return 0;
        case 3:
        case 4:
        case 5:
            return 0;
        default:
            break;
        }
    } else {
        switch (stage) {
        case 0:   // level 1 histograms
            SORT_LAUNCH(1, 1, nw1, x, y, z, Np, g, (const uint64_t *)nullptr,
                        (const uint16_t *)nullptr, (const uint32_t *)nullptr,
                        (const uint32_t *)nullptr, (const uint32_t *)nullptr, pp->nb1, pp->nb1,
                        pp->H1, (const uint32_t *)nullptr, (uint64_t *)nullptr,
                        (uint16_t *)nullptr, (uint8_t *)nullptr);
            return 0;
        case 1:   // level 1 offsets and the level 2 chunk layout
            hipLaunchKernelGGL(k_set_seg1, dim3(1), dim3(64), 0, s, pp->seg1,
                               (uint32_t)(nw1 / RB));
            FGK_LAUNCHED();
            return level_scans(s, pp->H1, nw1, pp->nb1, pp->S1, pp->B1, pp->seg1, 1, pp->T1,
                               pp->base1, 1, pp->cap, pp->chunk0, pp->blk0, pp->O1, pp->nb1, 1,
                               (long)pp->nb1);
        case 2:
            SORT_LAUNCH(1, 0, nw1, x, y, z, Np, g, (const uint64_t *)nullptr,
                        (const uint16_t *)nullptr, (const uint32_t *)nullptr,
                        (const uint32_t *)nullptr, (const uint32_t *)nullptr, pp->nb1, pp->nb1,
                        (uint32_t *)nullptr, pp->O1, pp->sorted1, pp->aux1, (uint8_t *)nullptr);
            return 0;
        case 3: { // level 2 histograms (every capacity chunk; empty ones write zeros)
            const long nw2 = nw1 + (long)pp->nb1 * RB;
            SORT_LAUNCH(2, 1, nw2, (const float *)nullptr, (const float *)nullptr,
                        (const float *)nullptr, Np, g, pp->sorted1, pp->aux1, pp->chunk0,
                        pp->base1, pp->T1, pp->nb1, pp->nb2, pp->H2, (const uint32_t *)nullptr,
                        (uint64_t *)nullptr, (uint16_t *)nullptr, (uint8_t *)nullptr);
            return 0;
        }
        case 4: {
            const long nw2 = nw1 + (long)pp->nb1 * RB;
            if (!pp->wide)
                return level_scans(s, pp->H2, nw2, pp->nb2, pp->S2, pp->B2, pp->blk0, pp->nb1,
                                   pp->T2, pp->tilebase, 2, pp->cap, pp->items,
                                   (uint32_t *)nullptr, pp->O2, pp->nslots, pp->nb1, 0l);
            // wide: the tile scan runs per bucket in parallel, then the
            // buckets' item bases are added
            const long nblk = nw2 / RB;
            hipLaunchKernelGGL(k_blocksum, dim3(blocks_for(nblk * pp->nb2, 256)), dim3(256), 0, s,
                               pp->H2, nblk, pp->nb2, pp->S2);
            FGK_LAUNCHED();
            hipLaunchKernelGGL(k_segscan, dim3(blocks_for((long)pp->nb1 * pp->nb2, 256)),
                               dim3(256), 0, s, pp->S2, pp->blk0, pp->nb1, pp->nb2, pp->B2, pp->T2);
            FGK_LAUNCHED();
            hipLaunchKernelGGL(k_bucket_slots, dim3((unsigned)pp->nb1), dim3(SCAN_THREADS), 0, s,
                               pp->T2, pp->base1, pp->nb1, pp->nb2, pp->cap, pp->tilebase,
                               pp->items, pp->bitem);
            FGK_LAUNCHED();
            hipLaunchKernelGGL(k_digitscan, dim3(1), dim3(SCAN_THREADS), 0, s, pp->bitem,
                               (long)pp->nb1, 2, pp->cap, pp->ibase, pp->bscratch,
                               (uint32_t *)nullptr);
            FGK_LAUNCHED();
            hipLaunchKernelGGL(k_add_item_base, dim3(blocks_for(pp->nslots + 1, 256)), dim3(256), 0,
                               s, pp->items, pp->ibase, pp->nb1, pp->nb2);
            FGK_LAUNCHED();
            hipLaunchKernelGGL(k_colscan, dim3(blocks_for(nblk * pp->nb2, 256)), dim3(256), 0, s,
                               pp->H2, pp->blk0, pp->nb1, nblk, pp->nb2, pp->tilebase, pp->B2,
                               pp->O2, pp->nb1, 0l);
            FGK_LAUNCHED();
            return 0;
        }
        case 5: {
            const long nw2 = nw1 + (long)pp->nb1 * RB;
            SORT_LAUNCH(2, 0, nw2, (const float *)nullptr, (const float *)nullptr,
                        (const float *)nullptr, Np, g, pp->sorted1, pp->aux1, pp->chunk0,
                        pp->base1, pp->T1, pp->nb1, pp->nb2, (uint32_t *)nullptr, pp->O2,
                        pp->sorted2, (uint16_t *)nullptr, pp->ext2);
            return 0;
        }
        default:
            break;
        }
    }
    if (stage == 6) {
        const long max_items = paint_work(pp->nb1, g, Np, pp->cap);
        const size_t lds = (size_t)paint_lds_bytes(pp->T, pp->interlace);
        const dim3 grid((unsigned)max_items), block(PAINT_THREADS);
        unsigned long long *A = (unsigned long long *)gridA, *B = (unsigned long long *)gridB;
#define PAINT_T(TT)                                                                           \
    hipLaunchKernelGGL((k_paint_tiles<TT>), grid, block, lds, s, pp->sorted2, pp->ext2,       \
                       pp->tilebase, pp->items, pp->nslots, g, pp->nb2, pp->cap,              \
                       pp->interlace, A, B, (const PaintJob *)nullptr, (const long *)nullptr, \
                       0, 0, 0l, 0, pp->zero_ptr, pp->zero_words)
        switch (pp->T) {
        case 16: PAINT_T(16); break;
        case 12: PAINT_T(12); break;
        case 8: PAINT_T(8); break;
        case 4: PAINT_T(4); break;
        case 2: PAINT_T(2); break;
        case 1: PAINT_T(1); break;
        default:
            fgk_set_error("paint: no kernel instance for tile side %d", pp->T);
            return -2;
        }
#undef PAINT_T
        FGK_LAUNCHED();
        return 0;
    }
    fgk_set_error("paint: unknown stage %d", stage);
    return -2;
}

extern "C" void fgk_paint_set_zero(fgk_paintplan *pp, void *ptr, size_t bytes) {
    pp->zero_ptr = (uint64_t *)ptr;
    pp->zero_words = ptr ? (long)(bytes / 16) * 2 : 0;
}

extern "C" int fgk_paint_fixed(fgk_paintplan *pp, const float *x, const float *y, const float *z,
                               long Np, double Hinv, int64_t *gridA, int64_t *gridB,
                               fgk_stream_t s) {
    for (int stage = 0; stage < fgk_paint_nstages(); ++stage)
        if (fgk_paint_stage(pp, stage, x, y, z, Np, Hinv, gridA, gridB, s)) return -1;
    return 0;
}

// ------------------------------------------------------- batched catalogs
//
// Many catalogs, each with its own grid (side, box, tile side), painted by
// one launch per sort stage and one paint launch per tile side.  Level 1
// runs over the concatenated 4096-particle chunks of all catalogs (each
// catalog's chunk range a multiple of RB, its digits its own <= 256 coarse
// buckets); level 2 over the (catalog, bucket) segments; the tile tables of
// catalog j sit at j * BATCH_SL with MAX_D2 slots per bucket.  The sorted
// words of catalog j follow those of catalogs 0..j-1, because the level-1
// offsets are one exclusive scan over (catalog, bucket).

#define BATCH_SL ((long)MAX_D1 * MAX_D2 + 1)   // tile-table entries per catalog

// per catalog j (one workgroup each): exclusive scans of its slot counts
// T2[j][s], s < nb1_j * MAX_D2, into tilebase (starting at the catalog's
// first sorted position base1[j * MAX_D1]) and dense paint items (dense_items)
__global__ void __launch_bounds__(SCAN_THREADS)
k_digitscan_jobs(const uint32_t *__restrict__ T2, const uint32_t *__restrict__ base1,
                 const PaintJob *__restrict__ jobs, int cap, uint32_t *__restrict__ tilebase,
                 uint32_t *__restrict__ items) {
    __shared__ uint32_t sa[SCAN_THREADS], sb[SCAN_THREADS];
    const int job = blockIdx.x;
    const long n = (long)jobs[job].nb1 * MAX_D2;
    const uint32_t *v = T2 + (long)job * MAX_D1 * MAX_D2;
    uint32_t *base = tilebase + job * BATCH_SL, *aux = items + job * BATCH_SL;
    const uint32_t b0 = base1[(long)job * MAX_D1];
    const long per = (n + SCAN_THREADS - 1) / SCAN_THREADS;
    const long i0 = threadIdx.x * per;
    const long i1 = i0 + per < n ? i0 + per : n;
    uint32_t ca = 0, cb = 0;
    for (long i = i0; i < i1; ++i) {
        const uint32_t c = v[i];
        ca += c;
        cb += dense_items(c, cap);
    }
    sa[threadIdx.x] = ca;
    sb[threadIdx.x] = cb;
    __syncthreads();
    for (int o = 1; o < SCAN_THREADS; o <<= 1) {
        const uint32_t va = threadIdx.x >= o ? sa[threadIdx.x - o] : 0;
        const uint32_t vb = threadIdx.x >= o ? sb[threadIdx.x - o] : 0;
        __syncthreads();
        sa[threadIdx.x] += va;
        sb[threadIdx.x] += vb;
        __syncthreads();
    }
    uint32_t ra = b0 + sa[threadIdx.x] - ca, rb = sb[threadIdx.x] - cb;
    for (long i = i0; i < i1; ++i) {
        base[i] = ra;
        aux[i] = rb;
        const uint32_t c = v[i];
        ra += c;
        rb += dense_items(c, cap);
    }
    if (threadIdx.x == SCAN_THREADS - 1) {
        base[n] = b0 + sa[SCAN_THREADS - 1];
        aux[n] = sb[SCAN_THREADS - 1];
    }
}

struct fgk_paintbatch {
    int njobs, cap;
    long NW1, NW2, Np_tot;
    int nsides, side[6], sja[7];   // jobs [sja[i], sja[i+1]) paint with tile side side[i]
    int *perm;                     // internal job -> caller's job
    PaintJob *hjobs, *djobs;
    long *hNpmax;                  // per internal job
    long *hc1, *hi0, *dc1, *di0;   // [njobs+1] chunk and paint-item offsets
    int interlace_any;
    uint32_t *blk1;                // [njobs+1] level-1 segments in RB-chunk blocks
    uint32_t *H1, *O1, *S1, *B1, *T1;
    uint32_t *base1, *chunk0, *blk2;
    uint32_t *H2, *O2, *S2, *B2, *T2;
    uint32_t *tilebase, *items;
    uint64_t *sorted1, *sorted2;
    uint16_t *aux1;
    uint8_t *ext2;
    int table_ok;                  // djobs holds hjobs
    int sets;                      // grid sets of every job (1 or 2)
    int set;                       // set the paint stage writes
    uint64_t *zero;                // cleared by the paint stage's launches
    long zero_words;
    int words_external;            // sorted1, sorted2, aux1, ext2 in a caller's workspace
    int per_side;                  // paint launches per tile side (fgk_paint_batch_per_side)
};

extern "C" void fgk_paint_batch_destroy(fgk_paintbatch *pb) {
    if (!pb) return;
    uint32_t *bufs[] = {pb->blk1, pb->H1,    pb->O1,     pb->S1,       pb->B1,   pb->T1,
                        pb->base1, pb->chunk0, pb->blk2, pb->H2,       pb->O2,   pb->S2,
                        pb->B2,   pb->T2,    pb->tilebase, pb->items};
    for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; ++i) fgk_free(bufs[i]);
    if (!pb->words_external) {
        fgk_free(pb->sorted1);
        fgk_free(pb->sorted2);
        fgk_free(pb->aux1);
        fgk_free(pb->ext2);
    }
    fgk_free(pb->djobs);
    fgk_free(pb->dc1);
    fgk_free(pb->di0);
    free(pb->perm);
    free(pb->hjobs);
    free(pb->hNpmax);
    free(pb->hc1);
    free(pb->hi0);
    free(pb);
}

// the sort words of Np_tot particles (two main words, the side word and the
// extension), each array 256-byte aligned, as carved from a workspace
static size_t words_layout(long Np_tot, size_t off[4]) {
    const size_t sz[4] = {(size_t)Np_tot * 8, (size_t)Np_tot * 8, (size_t)Np_tot * 2,
                          (size_t)Np_tot};
    size_t o = 0;
    for (int i = 0; i < 4; ++i) {
        off[i] = o;
        o += (sz[i] + 255) & ~(size_t)255;
    }
    return o;
}

extern "C" size_t fgk_paint_batch_ws_bytes(long Np_tot) {
    size_t off[4];
    return words_layout(Np_tot, off);
}

extern "C" int fgk_paint_batch_create(fgk_paintbatch **out, int njobs,
                                      const fgk_paintjob_desc *jobs) {
    return fgk_paint_batch_create_ws(out, njobs, jobs, NULL, 0);
}

extern "C" int fgk_paint_batch_create_ws(fgk_paintbatch **out, int njobs,
                                         const fgk_paintjob_desc *jobs, void *ws,
                                         size_t ws_bytes) {
    *out = NULL;
    if (njobs < 1 || njobs > 65535) {
        fgk_set_error("paint batch: need 1..65535 catalogs");
        return -2;
    }
    fgk_paintbatch *pb = (fgk_paintbatch *)calloc(1, sizeof(fgk_paintbatch));
    if (!pb) { fgk_set_error("out of host memory"); return -1; }
    pb->njobs = njobs;
    pb->cap = 8192;
    pb->perm = (int *)calloc((size_t)njobs, sizeof(int));
    pb->hjobs = (PaintJob *)calloc((size_t)njobs, sizeof(PaintJob));
    pb->hNpmax = (long *)calloc((size_t)njobs, sizeof(long));
    pb->hc1 = (long *)calloc((size_t)njobs + 1, sizeof(long));
    pb->hi0 = (long *)calloc((size_t)njobs + 1, sizeof(long));
    if (!pb->perm || !pb->hjobs || !pb->hNpmax || !pb->hc1 || !pb->hi0) {
        fgk_set_error("out of host memory");
        fgk_paint_batch_destroy(pb);
        return -1;
    }
    // internal order: by tile side (largest first), so that each side's
    // paint launch covers one contiguous range of work items
    int n = 0;
    for (int si = 0; si < 6; ++si) {
        const int T = kSides[si];
        int first = n;
        for (int j = 0; j < njobs; ++j) {
            const fgk_paintjob_desc &d = jobs[j];
            if (d.Ng < 2 || d.Ng > 65535 || !(d.Hinv > 0) || d.Np_max < 1 || !d.gridA ||
                (d.interlace && !d.gridB)) {
                fgk_set_error("paint batch: invalid catalog %d", j);
                fgk_paint_batch_destroy(pb);
                return -2;
            }
            if (paint_side(d.Ng, d.tile) == T) pb->perm[n++] = j;
        }
        if (n > first) {
            pb->side[pb->nsides] = T;
            pb->sja[pb->nsides] = first;
            ++pb->nsides;
        }
    }
    pb->sja[pb->nsides] = n;
    long Np_tot = 0, c1 = 0, i0 = 0;
    long nb1_tot = 0;
    for (int q = 0; q < njobs; ++q) {
        const fgk_paintjob_desc &d = jobs[pb->perm[q]];
        const int T = paint_side(d.Ng, d.tile);
        int two_level = 0;
        KeyGeo g = tile_geometry(d.Ng, T, 0, two_level);
        if (!two_level) {
            fgk_set_error("paint batch: catalog %d, Ng=%d with tiles of side %d has more tiles "
                          "than a batch addresses (%d x %d); paint it with its own plan",
                          pb->perm[q], d.Ng, T, MAX_D1, MAX_D2);
            fgk_paint_batch_destroy(pb);
            return -2;
        }
        g.Hinv = d.Hinv;
        PaintJob &J = pb->hjobs[q];
        J.g = g;
        J.GA = (unsigned long long *)d.gridA;
        J.GB = (unsigned long long *)(d.interlace ? d.gridB : nullptr);
        J.GA2 = (unsigned long long *)d.gridA2;
        J.GB2 = (unsigned long long *)(d.interlace ? d.gridB2 : nullptr);
        if (q == 0) pb->sets = d.gridA2 ? 2 : 1;
        if ((d.gridA2 != nullptr) != (pb->sets == 2) || (d.gridA2 && d.interlace && !d.gridB2)) {
            fgk_set_error("paint batch: catalog %d: every catalog needs the same number of "
                          "grid sets", pb->perm[q]);
            fgk_paint_batch_destroy(pb);
            return -2;
        }
        J.nb1 = g.nc0 * g.nc1 * g.nc2;
        J.interlace = d.interlace ? 1 : 0;
        pb->interlace_any |= J.interlace;
        pb->hNpmax[q] = d.Np_max;
        Np_tot += d.Np_max;
        nb1_tot += J.nb1;
        pb->hc1[q] = c1;
        c1 += round_up((d.Np_max + SORT_CHUNK - 1) / SORT_CHUNK, RB);
        pb->hi0[q] = i0;
        i0 += paint_work(J.nb1, g, d.Np_max, pb->cap);
    }
    pb->hc1[njobs] = c1;
    pb->hi0[njobs] = i0;
    pb->Np_tot = Np_tot;
    pb->NW1 = c1;
    pb->NW2 = c1 + nb1_tot * RB;
    if (Np_tot >= (1l << 32) - 1) {
        fgk_set_error("paint batch: at most 2^32-2 particles over all catalogs (%ld)", Np_tot);
        fgk_paint_batch_destroy(pb);
        return -2;
    }
    if (pb->NW2 >= (1l << 32) / SORT_THREADS || i0 * (long)PAINT_THREADS >= (1l << 32)) {
        fgk_set_error("paint batch: too much work for one dispatch per stage");
        fgk_paint_batch_destroy(pb);
        return -2;
    }
    const long nseg2 = (long)njobs * MAX_D1;
    const size_t u = sizeof(uint32_t);
    int rc = 0;
    rc |= fgk_malloc((void **)&pb->djobs, (size_t)njobs * sizeof(PaintJob));
    rc |= fgk_malloc((void **)&pb->dc1, (size_t)(njobs + 1) * sizeof(long));
    rc |= fgk_malloc((void **)&pb->di0, (size_t)(njobs + 1) * sizeof(long));
    rc |= fgk_malloc((void **)&pb->blk1, (size_t)(njobs + 1) * u);
    rc |= fgk_malloc((void **)&pb->H1, (size_t)pb->NW1 * MAX_D1 * u);
    rc |= fgk_malloc((void **)&pb->O1, (size_t)pb->NW1 * MAX_D1 * u);
    rc |= fgk_malloc((void **)&pb->S1, (size_t)(pb->NW1 / RB) * MAX_D1 * u);
    rc |= fgk_malloc((void **)&pb->B1, (size_t)(pb->NW1 / RB) * MAX_D1 * u);
    rc |= fgk_malloc((void **)&pb->T1, (size_t)nseg2 * u);
    rc |= fgk_malloc((void **)&pb->base1, (size_t)(nseg2 + 1) * u);
    rc |= fgk_malloc((void **)&pb->chunk0, (size_t)(nseg2 + 1) * u);
    rc |= fgk_malloc((void **)&pb->blk2, (size_t)(nseg2 + 1) * u);
    rc |= fgk_malloc((void **)&pb->H2, (size_t)pb->NW2 * MAX_D2 * u);
    rc |= fgk_malloc((void **)&pb->O2, (size_t)pb->NW2 * MAX_D2 * u);
    rc |= fgk_malloc((void **)&pb->S2, (size_t)(pb->NW2 / RB) * MAX_D2 * u);
    rc |= fgk_malloc((void **)&pb->B2, (size_t)(pb->NW2 / RB) * MAX_D2 * u);
    rc |= fgk_malloc((void **)&pb->T2, (size_t)nseg2 * MAX_D2 * u);
    rc |= fgk_malloc((void **)&pb->tilebase, (size_t)njobs * BATCH_SL * u);
    rc |= fgk_malloc((void **)&pb->items, (size_t)njobs * BATCH_SL * u);
    if (ws) {
        size_t off[4];
        if (words_layout(Np_tot, off) > ws_bytes) {
            fgk_set_error("paint batch: workspace of %zu bytes below the %zu the sort needs",
                          ws_bytes, words_layout(Np_tot, off));
            fgk_paint_batch_destroy(pb);
            return -2;
        }
        pb->words_external = 1;
        pb->sorted1 = (uint64_t *)((char *)ws + off[0]);
        pb->sorted2 = (uint64_t *)((char *)ws + off[1]);
        pb->aux1 = (uint16_t *)((char *)ws + off[2]);
        pb->ext2 = (uint8_t *)((char *)ws + off[3]);
    } else {
        rc |= fgk_malloc((void **)&pb->sorted1, (size_t)Np_tot * sizeof(uint64_t));
        rc |= fgk_malloc((void **)&pb->sorted2, (size_t)Np_tot * sizeof(uint64_t));
        rc |= fgk_malloc((void **)&pb->aux1, (size_t)Np_tot * sizeof(uint16_t));
        rc |= fgk_malloc((void **)&pb->ext2, (size_t)Np_tot * sizeof(uint8_t));
    }
    uint32_t *hb = (uint32_t *)malloc((size_t)(njobs + 1) * u);
    if (!rc && !hb) { fgk_set_error("out of host memory"); rc = -1; }
    if (!rc) {
        for (int q = 0; q <= njobs; ++q) hb[q] = (uint32_t)(pb->hc1[q] / RB);
        rc = fgk_h2d(pb->blk1, hb, (size_t)(njobs + 1) * u, NULL);
    }
    if (!rc) rc = fgk_h2d(pb->dc1, pb->hc1, (size_t)(njobs + 1) * sizeof(long), NULL);
    if (!rc) rc = fgk_h2d(pb->di0, pb->hi0, (size_t)(njobs + 1) * sizeof(long), NULL);
    if (!rc) rc = fgk_stream_sync(NULL);
    free(hb);
    if (rc) {
        fgk_paint_batch_destroy(pb);
        return -1;
    }
    *out = pb;
    return 0;
}

extern "C" int fgk_paint_batch_select(fgk_paintbatch *pb, int set, void *zero, size_t bytes) {
    if (set < 0 || set >= pb->sets || (zero && (bytes % 16 || ((uintptr_t)zero & 15)))) {
        fgk_set_error("paint batch: grid set %d of %d, or clear region not 16-byte aligned", set,
                      pb->sets);
        return -2;
    }
    pb->set = set;
    pb->zero = (uint64_t *)zero;
    pb->zero_words = zero ? (long)(bytes / 8) : 0;
    return 0;
}


extern "C" int fgk_paint_batch_tile(const fgk_paintbatch *pb, int job) {
    for (int q = 0; q < pb->njobs; ++q)
        if (pb->perm[q] == job) return pb->hjobs[q].g.T;
    return -1;
}

extern "C" int fgk_paint_batch_set(fgk_paintbatch *pb, const float *const *x,
                                   const float *const *y, const float *const *z, const long *Np,
                                   fgk_stream_t s) {
    int same = pb->table_ok;
    for (int q = 0; q < pb->njobs; ++q) {
        const int j = pb->perm[q];
        if (Np[j] < 0 || Np[j] > pb->hNpmax[q]) {
            fgk_set_error("paint batch: catalog %d has Np=%ld outside [0, Np_max=%ld]", j, Np[j],
                          pb->hNpmax[q]);
            return -2;
        }
        if (Np[j] > 0 && (!x[j] || !y[j] || !z[j])) {
            fgk_set_error("paint batch: catalog %d has NULL positions", j);
            return -2;
        }
        PaintJob &J = pb->hjobs[q];
        if (J.x != x[j] || J.y != y[j] || J.z != z[j] || J.Np != Np[j]) same = 0;
        J.x = x[j];
        J.y = y[j];
        J.z = z[j];
        J.Np = Np[j];
    }
    if (same) return 0;
    // the table is read by kernels already queued: wait for them, copy, wait
    if (fgk_stream_sync(s)) return -1;
    if (fgk_h2d(pb->djobs, pb->hjobs, (size_t)pb->njobs * sizeof(PaintJob), s)) return -1;
    if (fgk_stream_sync(s)) return -1;
    pb->table_ok = 1;
    return 0;
}

extern "C" void fgk_paint_batch_per_side(fgk_paintbatch *pb, int on) { pb->per_side = on ? 1 : 0; }

extern "C" int fgk_paint_batch_stage(fgk_paintbatch *pb, int stage, fgk_stream_t st) {
    hipStream_t s = (hipStream_t)st;
    if (!pb->table_ok) {
        fgk_set_error("paint batch: catalogs not set");
        return -2;
    }
    const int nj = pb->njobs;
    const long nseg2 = (long)nj * MAX_D1;
    const KeyGeo g0 = pb->hjobs[0].g;   // unused by batched launches
    switch (stage) {
    case 0:
        hipLaunchKernelGGL((k_sort<1, 1, 8>), dim3((unsigned)pb->NW1), dim3(SORT_THREADS), 0, s,
                           (const float *)nullptr, (const float *)nullptr,
                           (const float *)nullptr, 0l, g0, (const uint64_t *)nullptr,
                           (const uint16_t *)nullptr, (const uint32_t *)nullptr,
                           (const uint32_t *)nullptr, (const uint32_t *)nullptr, 0, MAX_D1,
                           pb->H1, (const uint32_t *)nullptr, (uint64_t *)nullptr,
                           (uint16_t *)nullptr, (uint8_t *)nullptr,
                           (const PaintJob *)pb->djobs, (const long *)pb->dc1, nj);
        FGK_LAUNCHED();
        return 0;
    case 1: {
        const long nblk = pb->NW1 / RB;
        hipLaunchKernelGGL(k_blocksum, dim3(blocks_for(nblk * MAX_D1, 256)), dim3(256), 0, s,
                           pb->H1, nblk, MAX_D1, pb->S1);
        FGK_LAUNCHED();
        hipLaunchKernelGGL(k_segscan, dim3(blocks_for((long)nj * MAX_D1, 256)), dim3(256), 0, s,
                           pb->S1, pb->blk1, nj, MAX_D1, pb->B1, pb->T1);
        FGK_LAUNCHED();
        hipLaunchKernelGGL(k_digitscan, dim3(1), dim3(SCAN_THREADS), 0, s, pb->T1, nseg2, 1,
                           pb->cap, pb->base1, pb->chunk0, pb->blk2);
        FGK_LAUNCHED();
        hipLaunchKernelGGL(k_colscan, dim3(blocks_for(nblk * MAX_D1, 256)), dim3(256), 0, s,
                           pb->H1, pb->blk1, nj, nblk, MAX_D1, pb->base1, pb->B1, pb->O1, 1,
                           (long)MAX_D1);
        FGK_LAUNCHED();
        return 0;
    }
    case 2:
        hipLaunchKernelGGL((k_sort<1, 0, 8>), dim3((unsigned)pb->NW1), dim3(SORT_THREADS), 0, s,
                           (const float *)nullptr, (const float *)nullptr,
                           (const float *)nullptr, 0l, g0, (const uint64_t *)nullptr,
                           (const uint16_t *)nullptr, (const uint32_t *)nullptr,
                           (const uint32_t *)nullptr, (const uint32_t *)nullptr, 0, MAX_D1,
                           (uint32_t *)nullptr, pb->O1, pb->sorted1, pb->aux1,
                           (uint8_t *)nullptr, (const PaintJob *)pb->djobs,
                           (const long *)pb->dc1, nj);
        FGK_LAUNCHED();
        return 0;
    case 3:
        hipLaunchKernelGGL((k_sort<2, 1, 7>), dim3((unsigned)pb->NW2), dim3(SORT_THREADS), 0, s,
                           (const float *)nullptr, (const float *)nullptr,
                           (const float *)nullptr, 0l, g0, pb->sorted1, pb->aux1, pb->chunk0,
                           pb->base1, pb->T1, (int)nseg2, MAX_D2, pb->H2,
                           (const uint32_t *)nullptr, (uint64_t *)nullptr, (uint16_t *)nullptr,
                           (uint8_t *)nullptr, (const PaintJob *)nullptr, (const long *)nullptr,
                           0);
        FGK_LAUNCHED();
        return 0;
    case 4: {
        const long nblk = pb->NW2 / RB;
        hipLaunchKernelGGL(k_blocksum, dim3(blocks_for(nblk * MAX_D2, 256)), dim3(256), 0, s,
                           pb->H2, nblk, MAX_D2, pb->S2);
        FGK_LAUNCHED();
        hipLaunchKernelGGL(k_segscan, dim3(blocks_for(nseg2 * MAX_D2, 256)), dim3(256), 0, s,
                           pb->S2, pb->blk2, (int)nseg2, MAX_D2, pb->B2, pb->T2);
        FGK_LAUNCHED();
        hipLaunchKernelGGL(k_digitscan_jobs, dim3((unsigned)nj), dim3(SCAN_THREADS), 0, s, pb->T2,
                           pb->base1, (const PaintJob *)pb->djobs, pb->cap, pb->tilebase,
                           pb->items);
        FGK_LAUNCHED();
        hipLaunchKernelGGL(k_colscan, dim3(blocks_for(nblk * MAX_D2, 256)), dim3(256), 0, s,
                           pb->H2, pb->blk2, (int)nseg2, nblk, MAX_D2, pb->tilebase, pb->B2,
                           pb->O2, MAX_D1, BATCH_SL);
        FGK_LAUNCHED();
        return 0;
    }
    case 5:
        hipLaunchKernelGGL((k_sort<2, 0, 7>), dim3((unsigned)pb->NW2), dim3(SORT_THREADS), 0, s,
                           (const float *)nullptr, (const float *)nullptr,
                           (const float *)nullptr, 0l, g0, pb->sorted1, pb->aux1, pb->chunk0,
                           pb->base1, pb->T1, (int)nseg2, MAX_D2, (uint32_t *)nullptr, pb->O2,
                           pb->sorted2, (uint16_t *)nullptr, pb->ext2, (const PaintJob *)nullptr,
                           (const long *)nullptr, 0);
        FGK_LAUNCHED();
        return 0;
    case 6: {
        bool multi = pb->nsides > 1 && !pb->per_side;
        int smax = 0;
        for (int i = 0; i < pb->nsides; ++i) {
            const int t = pb->side[i];
            if (t != 16 && t != 12 && t != 8) multi = false;
            if (t > smax) smax = t;
        }
        if (multi) {   // one launch for every catalog, whatever its tile side
            const long nitems = pb->hi0[pb->njobs];
            if (nitems <= 0) return 0;
            const size_t lds = (size_t)paint_lds_bytes(smax, pb->interlace_any);
            hipLaunchKernelGGL(k_paint_tiles_multi, dim3((unsigned)nitems), dim3(PAINT_THREADS),
                               lds, s, pb->sorted2, pb->ext2, pb->tilebase, pb->items, pb->cap,
                               (const PaintJob *)pb->djobs, (const long *)pb->di0, pb->njobs,
                               BATCH_SL, pb->set, pb->zero, pb->zero ? pb->zero_words : 0l);
            FGK_LAUNCHED();
            return 0;
        }
        // per tile side: the region to clear, in even-word pieces proportional
        // to each launch's work items
        const long pairs = pb->zero_words / 2, all = pb->hi0[pb->sja[pb->nsides]];
        for (int i = 0; i < pb->nsides; ++i) {
            const int ja = pb->sja[i], jb = pb->sja[i + 1];
            const long nitems = pb->hi0[jb] - pb->hi0[ja];
            const long w0 = 2 * (long)((double)pairs * pb->hi0[ja] / all),
                       w1 = i + 1 == pb->nsides ? 2 * pairs
                                                : 2 * (long)((double)pairs * pb->hi0[jb] / all);
            uint64_t *zb = pb->zero && w1 > w0 ? pb->zero + w0 : nullptr;
            if (nitems <= 0) continue;
            const size_t lds = (size_t)paint_lds_bytes(pb->side[i], pb->interlace_any);
            const dim3 grid((unsigned)nitems), block(PAINT_THREADS);
#define PAINT_TB(TT)                                                                          \
    hipLaunchKernelGGL((k_paint_tiles<TT>), grid, block, lds, s, pb->sorted2, pb->ext2,      \
                       pb->tilebase,                                                          \
                       pb->items, 0l, g0, MAX_D2, pb->cap, 0,                                 \
                       (unsigned long long *)nullptr, (unsigned long long *)nullptr,          \
                       (const PaintJob *)pb->djobs, (const long *)pb->di0, ja, jb, BATCH_SL,  \
                       pb->set, zb, zb ? w1 - w0 : 0l)
            switch (pb->side[i]) {
            case 16: PAINT_TB(16); break;
            case 12: PAINT_TB(12); break;
            case 8: PAINT_TB(8); break;
            case 4: PAINT_TB(4); break;
            case 2: PAINT_TB(2); break;
            case 1: PAINT_TB(1); break;
            default:
                fgk_set_error("paint: no kernel instance for tile side %d", pb->side[i]);
                return -2;
            }
#undef PAINT_TB
            FGK_LAUNCHED();
        }
        return 0;
    }
    default:
        break;
    }
    fgk_set_error("paint batch: unknown stage %d", stage);
    return -2;
}
