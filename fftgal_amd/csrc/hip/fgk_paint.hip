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
//    src/fft.c:257-276):
//      count   -- each particle's tile (T^3 cells, by the plain grid's PCS
//                 base cell); a per-workgroup tile histogram in LDS whose
//                 LDS atomic return value is the particle's rank in
//                 (workgroup, tile)
//      scan    -- column sums and prefix sums over the (workgroup x tile)
//                 histogram give every (workgroup, tile) its output range,
//                 and per tile the number of paint work items
//      scatter -- each particle written once, as a packed 64-bit tile-local
//                 position (3 x 21-bit fixed point), to its sorted slot
//      paint   -- one workgroup per work item (<= cap particles of one tile)
//                 accumulates the 4x4x4 stencils of both grids into LDS
//                 footprints ((T+3)^3 plain, (T+4)^3 dual) as int32 fixed
//                 point, then flushes the non-zero footprint cells into the
//                 int64 grids with global integer atomics.
//    Integer addition is associative, so the grids are bit-identical run to
//    run whatever the scheduling.  Weights are rounded to 2^-S per particle
//    and cell, S chosen so that cap particles cannot overflow int32.
#include <hip/hip_runtime.h>
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
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= Np) return;
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
    hipLaunchKernelGGL(k_paint_f64, dim3(fgk_ceil_div(Np, 256)), dim3(256), 0, (hipStream_t)s, f,
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

#define PAINT_THREADS 256
#define COUNT_THREADS 512
#define SCAN_THREADS 1024
#define LDS_HIST_MAX_TILES 32768   // packed uint16 pairs: 64 KB of LDS

struct fgk_paintplan {
    int Ng;
    int T;           // tile side in cells
    int nt;          // tiles per dimension
    int interlace;
    int F;           // fraction bits of the packed local position
    int S;           // fixed-point bits of the painted weights
    int cap;         // particles per paint work item
    int lds_hist;    // 1: per-workgroup LDS histograms; 0: global atomics
    long ntiles;
    long Np_max;
    long chunk;      // particles per count/scatter workgroup
    long nwg;        // count/scatter workgroups at Np_max
    uint16_t *H;     // [nwg][ntiles] histogram (lds_hist)
    uint32_t *O;     // [nwg][ntiles] output offsets (lds_hist)
    uint32_t *tot;   // [ntiles] per tile counts
    uint32_t *base;  // [ntiles+1] exclusive prefix of tot
    uint32_t *items; // [ntiles+1] exclusive prefix of work items per tile
    void *rank;      // [Np_max] uint16 (lds_hist) or uint32
    uint64_t *sorted;// [Np_max] packed tile-local positions
};

struct TileGeo {
    int Ng, T, nt, F;
    double Hinv;
};

// plain-grid PCS base cell (wrapped) and the particle's tile, plus the
// packed tile-local position li + d in 21-bit fixed point per axis
__device__ __forceinline__ void tile_axis(float xf, const TileGeo &g, int &tcoord,
                                          uint32_t &q) {
    double u = (double)xf * g.Hinv;
    double fl = floor(u);
    double d = u - fl;
    long gi = (long)fl;
    int gw = (int)(((gi % g.Ng) + g.Ng) % g.Ng);
    tcoord = gw / g.T;
    int li = gw - tcoord * g.T;
    const double sc = (double)(1u << g.F);
    long qq = (long)li * (1l << g.F) + (long)rint(d * sc);
    long qmax = (long)g.T * (1l << g.F) - 1;
    if (qq > qmax) qq = qmax;
    q = (uint32_t)qq;
}

__device__ __forceinline__ void particle_tile(const float *x, const float *y, const float *z,
                                              long p, const TileGeo &g, int &t, uint64_t &pk) {
    int ti, tj, tk;
    uint32_t qi, qj, qk;
    tile_axis(x[p], g, ti, qi);
    tile_axis(y[p], g, tj, qj);
    tile_axis(z[p], g, tk, qk);
    t = (ti * g.nt + tj) * g.nt + tk;
    pk = ((uint64_t)qi << 42) | ((uint64_t)qj << 21) | (uint64_t)qk;
}

__global__ void __launch_bounds__(COUNT_THREADS)
k_count_lds(const float *__restrict__ x, const float *__restrict__ y,
            const float *__restrict__ z, long Np, TileGeo g, int ntiles, long chunk,
            uint16_t *__restrict__ H, uint16_t *__restrict__ rank) {
    extern __shared__ uint32_t hist[];
    const int nw = (ntiles + 1) >> 1;
    for (int i = threadIdx.x; i < nw; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const long p0 = blockIdx.x * chunk;
    const long p1 = p0 + chunk < Np ? p0 + chunk : Np;
    for (long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        int t;
        uint64_t pk;
        particle_tile(x, y, z, p, g, t, pk);
        const uint32_t sh = (t & 1) << 4;
        uint32_t old = atomicAdd(&hist[t >> 1], 1u << sh);
        rank[p] = (uint16_t)((old >> sh) & 0xffffu);
    }
    __syncthreads();
    uint16_t *row = H + (long)blockIdx.x * ntiles;
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x)
        row[i] = (uint16_t)((hist[i >> 1] >> ((i & 1) << 4)) & 0xffffu);
}

__global__ void k_count_global(const float *__restrict__ x, const float *__restrict__ y,
                               const float *__restrict__ z, long Np, TileGeo g,
                               uint32_t *__restrict__ tot, uint32_t *__restrict__ rank) {
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= Np) return;
    int t;
    uint64_t pk;
    particle_tile(x, y, z, p, g, t, pk);
    rank[p] = atomicAdd(&tot[t], 1u);
}

// tot[t] = sum over workgroups of H[w][t]
__global__ void k_colsum(const uint16_t *__restrict__ H, long nwg, int ntiles, uint32_t *tot) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    uint32_t s = 0;
    for (long w = 0; w < nwg; ++w) s += H[w * ntiles + t];
    tot[t] = s;
}

// O[w][t] = base[t] + sum_{w' < w} H[w'][t]
__global__ void k_colscan(const uint16_t *__restrict__ H, long nwg, int ntiles,
                          const uint32_t *__restrict__ base, uint32_t *__restrict__ O) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    uint32_t s = base[t];
    for (long w = 0; w < nwg; ++w) {
        O[w * ntiles + t] = s;
        s += H[w * ntiles + t];
    }
}

// single-workgroup exclusive scans of the tile counts and of the number of
// paint items per tile (ceil(count / cap)); writes n+1 entries each
__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_tiles(const uint32_t *__restrict__ tot, long n, int cap, uint32_t *__restrict__ base,
             uint32_t *__restrict__ items) {
    __shared__ uint32_t sa[SCAN_THREADS], sb[SCAN_THREADS];
    const long per = (n + SCAN_THREADS - 1) / SCAN_THREADS;
    const long b0 = threadIdx.x * per;
    const long b1 = b0 + per < n ? b0 + per : n;
    uint32_t ca = 0, cb = 0;
    for (long i = b0; i < b1; ++i) {
        uint32_t c = tot[i];
        ca += c;
        cb += (c + cap - 1) / cap;
    }
    sa[threadIdx.x] = ca;
    sb[threadIdx.x] = cb;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        uint32_t va = threadIdx.x >= off ? sa[threadIdx.x - off] : 0;
        uint32_t vb = threadIdx.x >= off ? sb[threadIdx.x - off] : 0;
        __syncthreads();
        sa[threadIdx.x] += va;
        sb[threadIdx.x] += vb;
        __syncthreads();
    }
    uint32_t ra = sa[threadIdx.x] - ca, rb = sb[threadIdx.x] - cb;
    for (long i = b0; i < b1; ++i) {
        base[i] = ra;
        items[i] = rb;
        uint32_t c = tot[i];
        ra += c;
        rb += (c + cap - 1) / cap;
    }
    if (threadIdx.x == SCAN_THREADS - 1) {
        base[n] = sa[SCAN_THREADS - 1];
        items[n] = sb[SCAN_THREADS - 1];
    }
}

template <int LDSH>
__global__ void __launch_bounds__(COUNT_THREADS)
k_scatter(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
          long Np, TileGeo g, int ntiles, long chunk, const uint32_t *__restrict__ off,
          const void *__restrict__ rank, uint64_t *__restrict__ sorted) {
    const long p0 = blockIdx.x * chunk;
    const long p1 = p0 + chunk < Np ? p0 + chunk : Np;
    const uint32_t *offrow = LDSH ? off + (long)blockIdx.x * ntiles : off;
    for (long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        int t;
        uint64_t pk;
        particle_tile(x, y, z, p, g, t, pk);
        uint32_t r = LDSH ? (uint32_t)((const uint16_t *)rank)[p] : ((const uint32_t *)rank)[p];
        sorted[offrow[t] + r] = pk;
    }
}

__device__ __forceinline__ void pcs_w_f32(float d, float w[4]) {
    float e = 1.0f - d;
    w[0] = e * e * e * (1.0f / 6.0f);
    w[1] = ((3.0f * d - 6.0f) * d * d + 4.0f) * (1.0f / 6.0f);
    w[2] = (((-3.0f * d + 3.0f) * d + 3.0f) * d + 1.0f) * (1.0f / 6.0f);
    w[3] = d * d * d * (1.0f / 6.0f);
}

__device__ __forceinline__ int wrapi(int v, int Ng) {
    v %= Ng;
    return v < 0 ? v + Ng : v;
}

__device__ void flush_footprint(const int *__restrict__ fp, int side, int oi, int oj, int ok,
                                int Ng, unsigned long long *__restrict__ G) {
    const int n = side * side * side;
    for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
        int v = fp[idx];
        if (!v) continue;
        int fi = idx / (side * side);
        int r = idx - fi * side * side;
        int fj = r / side;
        int fk = r - fj * side;
        long gi = wrapi(oi + fi, Ng), gj = wrapi(oj + fj, Ng), gk = wrapi(ok + fk, Ng);
        atomicAdd(G + (gi * Ng + gj) * Ng + gk, (unsigned long long)(long long)v);
    }
}

__global__ void __launch_bounds__(PAINT_THREADS)
k_paint_tiles(const uint64_t *__restrict__ sorted, const uint32_t *__restrict__ base,
              const uint32_t *__restrict__ items, int ntiles, int nt, int T, int Ng, int F,
              float wscale, int cap, int interlace, unsigned long long *__restrict__ GA,
              unsigned long long *__restrict__ GB) {
    extern __shared__ int lds[];
    const int it = blockIdx.x;
    if (it >= (int)items[ntiles]) return;
    int lo = 0, hi = ntiles - 1;          // last tile with items[t] <= it
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if ((int)items[mid] <= it) lo = mid; else hi = mid - 1;
    }
    const int t = lo;
    const long p0 = (long)base[t] + (long)(it - (int)items[t]) * cap;
    const long pe = base[t + 1];
    const long p1 = p0 + cap < pe ? p0 + cap : pe;

    const int sa = T + 3, sb = T + 4;
    const int na = sa * sa * sa, nb = interlace ? sb * sb * sb : 0;
    int *fa = lds, *fb = lds + na;
    for (int i = threadIdx.x; i < na + nb; i += blockDim.x) lds[i] = 0;
    __syncthreads();

    const uint32_t fmask = (1u << F) - 1u;
    const float fsc = 1.0f / (float)(1u << F);
    for (long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        uint64_t pk = sorted[p];
        uint32_t q[3] = {(uint32_t)(pk >> 42) & 0x1fffffu, (uint32_t)(pk >> 21) & 0x1fffffu,
                         (uint32_t)pk & 0x1fffffu};
        int li[3];
        float d[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            li[a] = (int)(q[a] >> F);
            d[a] = (float)(q[a] & fmask) * fsc;
        }
        float wx[4], wy[4], wz[4];
        pcs_w_f32(d[0], wx);
        pcs_w_f32(d[1], wy);
        pcs_w_f32(d[2], wz);
        // plain grid: nodes li-1..li+2, footprint origin (tile origin - 1)
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                int *row = fa + ((li[0] + a) * sa + (li[1] + b)) * sa + li[2];
                const float wab = wx[a] * wy[b];
#pragma unroll
                for (int c = 0; c < 4; ++c) atomicAdd(row + c, (int)rintf(wab * wz[c] * wscale));
            }
        if (interlace) {
            // dual grid (offset 1/2): base li or li-1, footprint origin (tile origin - 2)
            int ob[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (d[a] >= 0.5f) { d[a] -= 0.5f; ob[a] = li[a] + 1; }
                else { d[a] += 0.5f; ob[a] = li[a]; }
            }
            pcs_w_f32(d[0], wx);
            pcs_w_f32(d[1], wy);
            pcs_w_f32(d[2], wz);
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    int *row = fb + ((ob[0] + a) * sb + (ob[1] + b)) * sb + ob[2];
                    const float wab = wx[a] * wy[b];
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        atomicAdd(row + c, (int)rintf(wab * wz[c] * wscale));
                }
        }
    }
    __syncthreads();
    const int ti = t / (nt * nt), tj = (t / nt) % nt, tk = t % nt;
    flush_footprint(fa, sa, ti * T - 1, tj * T - 1, tk * T - 1, Ng, GA);
    if (interlace) flush_footprint(fb, sb, ti * T - 2, tj * T - 2, tk * T - 2, Ng, GB);
}

// ------------------------------------------------------------------ host

static int paint_lds_bytes(int T, int interlace) {
    int a = (T + 3) * (T + 3) * (T + 3);
    int b = interlace ? (T + 4) * (T + 4) * (T + 4) : 0;
    return (a + b) * 4;
}

extern "C" int fgk_paint_plan_create(fgk_paintplan **out, int Ng, int tile, long Np_max,
                                     int interlace) {
    *out = NULL;
    if (Ng < 2) { fgk_set_error("paint: Ng must be >= 2"); return -2; }
    if (Np_max < 1) Np_max = 1;
    if (Np_max >= (1l << 32) - 1) { fgk_set_error("paint: at most 2^32-2 particles"); return -2; }
    int T = tile > 0 ? tile : 16;
    if (T > Ng) T = Ng;
    if (T > 32) T = 32;
    while (T > 1 && paint_lds_bytes(T, interlace) > 64 * 1024) --T;
    fgk_paintplan *pp = (fgk_paintplan *)calloc(1, sizeof(fgk_paintplan));
    if (!pp) { fgk_set_error("out of host memory"); return -1; }
    pp->Ng = Ng;
    pp->T = T;
    pp->nt = (Ng + T - 1) / T;
    pp->interlace = interlace ? 1 : 0;
    int ib = 0;
    while ((1 << ib) < T) ++ib;
    pp->F = 21 - ib;
    pp->cap = 8192;
    // largest S with cap * (2/3)^3 * 2^S < 2^31
    pp->S = 0;
    while ((double)pp->cap * (8.0 / 27.0) * ldexp(1.0, pp->S + 1) < 2147483647.0) ++pp->S;
    pp->ntiles = (long)pp->nt * pp->nt * pp->nt;
    pp->Np_max = Np_max;
    pp->lds_hist = pp->ntiles <= LDS_HIST_MAX_TILES;
    pp->chunk = 65535;
    pp->nwg = (Np_max + pp->chunk - 1) / pp->chunk;
    int rc = 0;
    if (pp->lds_hist) {
        rc |= fgk_malloc((void **)&pp->H, (size_t)pp->nwg * pp->ntiles * sizeof(uint16_t));
        rc |= fgk_malloc((void **)&pp->O, (size_t)pp->nwg * pp->ntiles * sizeof(uint32_t));
        rc |= fgk_malloc(&pp->rank, (size_t)Np_max * sizeof(uint16_t));
    } else {
        rc |= fgk_malloc(&pp->rank, (size_t)Np_max * sizeof(uint32_t));
    }
    rc |= fgk_malloc((void **)&pp->tot, (size_t)pp->ntiles * sizeof(uint32_t));
    rc |= fgk_malloc((void **)&pp->base, (size_t)(pp->ntiles + 1) * sizeof(uint32_t));
    rc |= fgk_malloc((void **)&pp->items, (size_t)(pp->ntiles + 1) * sizeof(uint32_t));
    rc |= fgk_malloc((void **)&pp->sorted, (size_t)Np_max * sizeof(uint64_t));
    if (rc) {
        fgk_paint_plan_destroy(pp);
        return -1;
    }
    *out = pp;
    return 0;
}

extern "C" void fgk_paint_plan_destroy(fgk_paintplan *pp) {
    if (!pp) return;
    fgk_free(pp->H);
    fgk_free(pp->O);
    fgk_free(pp->rank);
    fgk_free(pp->tot);
    fgk_free(pp->base);
    fgk_free(pp->items);
    fgk_free(pp->sorted);
    free(pp);
}

extern "C" double fgk_paint_fixed_scale(const fgk_paintplan *pp) { return ldexp(1.0, -pp->S); }
extern "C" int fgk_paint_plan_tile(const fgk_paintplan *pp) { return pp->T; }

// stage 0: count, 1: scan, 2: scatter, 3: paint
extern "C" int fgk_paint_stage(fgk_paintplan *pp, int stage, const float *x, const float *y,
                               const float *z, long Np, double Hinv, int64_t *gridA,
                               int64_t *gridB, fgk_stream_t st) {
    hipStream_t s = (hipStream_t)st;
    if (Np > pp->Np_max) {
        fgk_set_error("paint: Np=%ld exceeds the plan's Np_max=%ld", Np, pp->Np_max);
        return -2;
    }
    TileGeo g = {pp->Ng, pp->T, pp->nt, pp->F, Hinv};
    const int ntiles = (int)pp->ntiles;
    const long nwg = (Np + pp->chunk - 1) / pp->chunk;
    switch (stage) {
    case 0:
        if (pp->lds_hist) {
            if (nwg > 0) {
                size_t lds = (size_t)((ntiles + 1) / 2) * 4;
                hipLaunchKernelGGL(k_count_lds, dim3((unsigned)nwg), dim3(COUNT_THREADS), lds, s,
                                   x, y, z, Np, g, ntiles, pp->chunk, pp->H,
                                   (uint16_t *)pp->rank);
                FGK_LAUNCHED();
            }
        } else {
            if (fgk_fill_zero(pp->tot, (size_t)ntiles * 4, st)) return -1;
            if (Np > 0) {
                hipLaunchKernelGGL(k_count_global, dim3(fgk_ceil_div(Np, 256)), dim3(256), 0, s,
                                   x, y, z, Np, g, pp->tot, (uint32_t *)pp->rank);
                FGK_LAUNCHED();
            }
        }
        return 0;
    case 1:
        if (pp->lds_hist) {
            hipLaunchKernelGGL(k_colsum, dim3(fgk_ceil_div(ntiles, 256)), dim3(256), 0, s, pp->H,
                               nwg, ntiles, pp->tot);
            FGK_LAUNCHED();
        }
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, s, pp->tot,
                           (long)ntiles, pp->cap, pp->base, pp->items);
        FGK_LAUNCHED();
        if (pp->lds_hist) {
            hipLaunchKernelGGL(k_colscan, dim3(fgk_ceil_div(ntiles, 256)), dim3(256), 0, s, pp->H,
                               nwg, ntiles, pp->base, pp->O);
            FGK_LAUNCHED();
        }
        return 0;
    case 2:
        if (Np <= 0) return 0;
        if (pp->lds_hist) {
            hipLaunchKernelGGL(k_scatter<1>, dim3((unsigned)nwg), dim3(COUNT_THREADS), 0, s, x, y,
                               z, Np, g, ntiles, pp->chunk, pp->O, pp->rank, pp->sorted);
        } else {
            hipLaunchKernelGGL(k_scatter<0>, dim3(fgk_ceil_div(Np, COUNT_THREADS)),
                               dim3(COUNT_THREADS), 0, s, x, y, z, Np, g, ntiles,
                               (long)COUNT_THREADS, pp->base, pp->rank, pp->sorted);
        }
        FGK_LAUNCHED();
        return 0;
    case 3: {
        long max_items = pp->ntiles + (Np + pp->cap - 1) / pp->cap;
        size_t lds = (size_t)paint_lds_bytes(pp->T, pp->interlace);
        hipLaunchKernelGGL(k_paint_tiles, dim3((unsigned)max_items), dim3(PAINT_THREADS), lds, s,
                           pp->sorted, pp->base, pp->items, ntiles, pp->nt, pp->T, pp->Ng, pp->F,
                           (float)ldexp(1.0, pp->S), pp->cap, pp->interlace,
                           (unsigned long long *)gridA, (unsigned long long *)gridB);
        FGK_LAUNCHED();
        return 0;
    }
    default:
        fgk_set_error("paint: unknown stage %d", stage);
        return -2;
    }
}

extern "C" int fgk_paint_fixed(fgk_paintplan *pp, const float *x, const float *y, const float *z,
                               long Np, double Hinv, int64_t *gridA, int64_t *gridB,
                               fgk_stream_t s) {
    for (int stage = 0; stage < 4; ++stage)
        if (fgk_paint_stage(pp, stage, x, y, z, Np, Hinv, gridA, gridB, s)) return -1;
    return 0;
}
