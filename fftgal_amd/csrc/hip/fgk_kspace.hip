// k-space kernels on gfx950.
//
// Per-function FP64 kernels behind the drop-in API:
//   k_offset_phase_f64  src/fft.c:196-233 (separable table, 4-term product)
//   k_deconv_f64        src/fft.c:236-254 (PCS window W^-1 per axis)
//   k_P_f64             src/power.c:12-27 (|delta|^2/V - Pshot, DC = 0)
// and one binning kernel used two ways:
//   MODE 0: bins P(k) held in the FP64 grid (src/power.c:52-67, 125-138)
//   MODE 1: the fused fast-path kernel -- reads the FP32 spectra of the plain
//           and dual grids, applies the dual grid's offset phase, the
//           interlace average (src/fft.c:270-271), the deconvolution, forms
//           |delta|^2/V - Pshot in FP64 and bins it, without writing a grid.
// Binning: each workgroup owns a 4 x 4 x 64 block of modes, whose |k| spans
// a short window of k bins; the Legendre (l = 0,2,4,6) and mu-wedge sums of
// that window are accumulated in LDS (FP64) and flushed with one global
// FP64 atomic per bin and quantity.  Bin and wedge indices are computed with
// the reference's FP64 expressions (no contraction, IEEE sqrt and division),
// so mode counts match the reference exactly.
#include <hip/hip_runtime.h>
#include <math.h>
#include "../fgk.h"
#include "fgk_common.h"

// ----------------------------------------------------------- FP64 per-fn

__device__ __forceinline__ void mode_ijk(long m, int Ng, int &i, int &j, int &k) {
    const int Nk = Ng / 2 + 1;
    long r = m / Nk;
    k = (int)(m - r * Nk);
    j = (int)(r % Ng);
    i = (int)(r / Ng);
}

__global__ void k_offset_phase_f64(double *f, int Ng, const double *__restrict__ ph) {
#pragma clang fp contract(off)
    const long nmodes = (long)Ng * Ng * (Ng / 2 + 1);
    long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (m >= nmodes) return;
    int i, j, k;
    mode_ijk(m, Ng, i, j, k);
    const double *p0 = ph + 2 * i, *p1 = ph + 2 * (Ng + j), *p2 = ph + 2 * (2 * Ng + k);
    double pr = p0[0] * p1[0] * p2[0] - p0[0] * p1[1] * p2[1] - p0[1] * p1[0] * p2[1]
              - p0[1] * p1[1] * p2[0];
    double pi = p0[1] * p1[0] * p2[0] + p0[0] * p1[1] * p2[0] + p0[0] * p1[0] * p2[1]
              - p0[1] * p1[1] * p2[1];
    double *c = f + 2 * m;
    double fr = c[0], fi = c[1];
    c[0] = fr * pr - fi * pi;
    c[1] = fr * pi + fi * pr;
}

__global__ void k_deconv_f64(double *f, int Ng, const double *__restrict__ winv) {
#pragma clang fp contract(off)
    const long nmodes = (long)Ng * Ng * (Ng / 2 + 1);
    long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (m >= nmodes) return;
    int i, j, k;
    mode_ijk(m, Ng, i, j, k);
    double w3 = winv[i] * winv[j] * winv[k];
    f[2 * m] *= w3;
    f[2 * m + 1] *= w3;
}

__global__ void k_P_f64(double *f, int Ng, double Vinv, double Pshot) {
#pragma clang fp contract(off)
    const long nmodes = (long)Ng * Ng * (Ng / 2 + 1);
    long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (m >= nmodes) return;
    if (m == 0) {
        f[0] = 0;
        f[1] = 0;
        return;
    }
    double re = f[2 * m], im = f[2 * m + 1];
    double d2 = re * re + im * im;
    f[2 * m] = d2 * Vinv - Pshot;
    f[2 * m + 1] = 0;
}

extern "C" int fgk_offset_phase_f64(double *f, int Ng, const double *phase_dev, fgk_stream_t s) {
    long n = (long)Ng * Ng * (Ng / 2 + 1);
    hipLaunchKernelGGL(k_offset_phase_f64, dim3(fgk_ceil_div(n, 256)), dim3(256), 0,
                       (hipStream_t)s, f, Ng, phase_dev);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_deconv_f64(double *f, int Ng, const double *winv_dev, fgk_stream_t s) {
    long n = (long)Ng * Ng * (Ng / 2 + 1);
    hipLaunchKernelGGL(k_deconv_f64, dim3(fgk_ceil_div(n, 256)), dim3(256), 0, (hipStream_t)s, f,
                       Ng, winv_dev);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_P_f64(double *f, int Ng, double Vinv, double Pshot, fgk_stream_t s) {
    long n = (long)Ng * Ng * (Ng / 2 + 1);
    hipLaunchKernelGGL(k_P_f64, dim3(fgk_ceil_div(n, 256)), dim3(256), 0, (hipStream_t)s, f, Ng,
                       Vinv, Pshot);
    FGK_LAUNCHED();
    return 0;
}

// ------------------------------------------------------------- binning

#define BIN_THREADS 256
#define BI 4
#define BJ 4
#define BK 64
#define BIN_LDS_BYTES (12 * 1024)

extern "C" size_t fgk_bins_doubles(const fgk_binspec *b) {
    size_t W = b->Nmu > 0 ? (size_t)b->Nmu + 1 : 0;
    size_t n = (size_t)b->NK * (4 + W);
    if (b->geometry) n += (size_t)b->NK * (1 + W);
    return n;
}

extern "C" size_t fgk_bins_counts(const fgk_binspec *b) {
    size_t W = b->Nmu > 0 ? (size_t)b->Nmu + 1 : 0;
    return b->geometry ? (size_t)b->NK * (1 + W) : 0;
}

struct BinArgs {
    int Ng;
    int NK;
    int Nmu;
    int do_pl;
    int geometry;
    int nvals;        // doubles per bin in LDS
    int ncnt;         // counts per bin in LDS
    int wmax;         // LDS window capacity in bins
    double dKinv;
    double l0, l1, l2;
};

// signed frequency of index i with the Nyquist index carrying -Ng/2
// (src/power.c:46-51)
__device__ __forceinline__ int kval(int i, int Ng) { return (2 * i < Ng) ? i : i - Ng; }

// smallest and largest |frequency| over an index range [a, b) of one axis
__device__ __forceinline__ void axis_range(int a, int b, int Ng, int &lo, int &hi) {
    int fa = kval(a, Ng), fb = kval(b - 1, Ng);
    // within a block the signed frequencies are contiguous (block edges are
    // multiples of the block size, which divides Ng/2); handle generic Ng
    int mn = fa < fb ? fa : fb, mx = fa < fb ? fb : fa;
    if (fb - fa != b - 1 - a) {  // block straddles the Nyquist wrap: be safe
        lo = 0;
        hi = Ng / 2;
        return;
    }
    if (mn <= 0 && mx >= 0) lo = 0;
    else lo = mn > 0 ? mn : -mx;
    int am = mn < 0 ? -mn : mn, ax = mx < 0 ? -mx : mx;
    hi = am > ax ? am : ax;
}

template <int MODE>
__global__ void __launch_bounds__(BIN_THREADS)
k_bin(const double *__restrict__ fP, const float *__restrict__ A, const float *__restrict__ B,
      BinArgs a, const double *__restrict__ ph, const double *__restrict__ winv, double norm,
      double Vinv, double Pshot, double *__restrict__ bins, unsigned long long *__restrict__ cnts) {
#pragma clang fp contract(off)
    extern __shared__ double sh[];
    const int Ng = a.Ng;
    const int Nk = Ng / 2 + 1;
    const int nbk = (Nk + BK - 1) / BK;
    const int nbj = (Ng + BJ - 1) / BJ;
    int blk = blockIdx.x;
    const int bk = blk % nbk;
    blk /= nbk;
    const int bj = blk % nbj;
    const int bi = blk / nbj;
    const int i0 = bi * BI, j0 = bj * BJ, k0 = bk * BK;
    const int i1 = min(i0 + BI, Ng), j1 = min(j0 + BJ, Ng), k1 = min(k0 + BK, Nk);

    // k-bin window of this block
    int lo0, hi0, lo1, hi1, lo2, hi2;
    axis_range(i0, i1, Ng, lo0, hi0);
    axis_range(j0, j1, Ng, lo1, hi1);
    axis_range(k0, k1, Ng, lo2, hi2);
    double kmin = sqrt((double)lo0 * lo0 + (double)lo1 * lo1 + (double)lo2 * lo2);
    double kmax = sqrt((double)hi0 * hi0 + (double)hi1 * hi1 + (double)hi2 * hi2);
    int blo = (int)floor(kmin * a.dKinv);
    int bhi = (int)floor(kmax * a.dKinv);
    if (blo > a.NK - 1) blo = a.NK - 1;
    if (bhi > a.NK - 1) bhi = a.NK - 1;
    const int wn = bhi - blo + 1;
    const bool windowed = wn <= a.wmax;
    const int Wm = a.Nmu > 0 ? a.Nmu + 1 : 0;

    double *sv = sh;                                         // [wmax][nvals]
    unsigned int *sc = (unsigned int *)(sh + (size_t)a.wmax * a.nvals);  // [wmax][ncnt]
    if (windowed) {
        for (int t = threadIdx.x; t < wn * a.nvals; t += blockDim.x) sv[t] = 0.0;
        for (int t = threadIdx.x; t < wn * a.ncnt; t += blockDim.x) sc[t] = 0u;
    }
    __syncthreads();

    // global layout offsets
    const size_t offPmu = (size_t)a.NK * 4;
    const size_t offK = offPmu + (size_t)a.NK * Wm;
    const size_t offKmu = offK + (size_t)a.NK;

    const int kk = threadIdx.x % BK;
    const int k = k0 + kk;
    for (int r = threadIdx.x / BK; r < BI * BJ; r += BIN_THREADS / BK) {
        const int i = i0 + r / BJ, j = j0 + r % BJ;
        if (i >= i1 || j >= j1 || k >= k1) continue;
        if (i == 0 && j == 0 && k == 0) continue;
        const long m = ((long)i * Ng + j) * Nk + k;
        double P = 0.0;
        if (MODE == 0) {
            P = fP[2 * m];
        } else if (MODE == 1) {
            // fused: plain + phase * dual, averaged, deconvolved
            const float2 av = ((const float2 *)A)[m];
            double dr = (double)av.x * norm, di = (double)av.y * norm;
            if (B) {
                const double *p0 = ph + 2 * i, *p1 = ph + 2 * (Ng + j), *p2 = ph + 2 * (2 * Ng + k);
                double pr = p0[0] * p1[0] * p2[0] - p0[0] * p1[1] * p2[1]
                          - p0[1] * p1[0] * p2[1] - p0[1] * p1[1] * p2[0];
                double pim = p0[1] * p1[0] * p2[0] + p0[0] * p1[1] * p2[0]
                           + p0[0] * p1[0] * p2[1] - p0[1] * p1[1] * p2[1];
                const float2 bv = ((const float2 *)B)[m];
                double br = (double)bv.x * norm, bim = (double)bv.y * norm;
                double er = br * pr - bim * pim, ei = br * pim + bim * pr;
                dr = 0.5 * (er + dr);
                di = 0.5 * (ei + di);
            }
            double w3 = winv[i] * winv[j] * winv[k];
            dr *= w3;
            di *= w3;
            P = (dr * dr + di * di) * Vinv - Pshot;
        }
        const int count = 1 + ((2 * k) % Ng > 0);
        const double d2 = P * count;
        const double fi = (double)kval(i, Ng), fj = (double)kval(j, Ng), fk = (double)kval(k, Ng);
        const double Kamp = sqrt(fi * fi + fj * fj + fk * fk);
        const int iK = (int)floor(Kamp * a.dKinv);
        if (iK >= a.NK) continue;   // cannot happen for NK of src/power.c:36; keeps LDS in bounds
        const double dot = fi * a.l0 + fj * a.l1 + fk * a.l2;
        const double mu = dot / Kamp;
        int imu = 0;
        if (a.Nmu > 0) imu = (int)floor(fabs(mu * a.Nmu));
        double pl[4];
        const bool sums = MODE != 2;
        if (a.do_pl) {
            const double mu2 = mu * mu;
            pl[0] = d2;
            pl[1] = d2 * (1.5 * mu2 - 0.5);
            pl[2] = d2 * ((4.375 * mu2 - 3.75) * mu2 + 0.375);
            pl[3] = d2 * (((14.4375 * mu2 - 19.6875) * mu2 + 6.5625) * mu2 - 0.3125);
        }
        if (windowed) {
            double *v = sv + (size_t)(iK - blo) * a.nvals;
            int o = 0;
            if (a.do_pl) {
                if (sums)
                    for (int l = 0; l < 4; ++l) atomicAdd(v + l, pl[l]);
                o = 4;
            }
            if (Wm && sums) atomicAdd(v + o + imu, d2);
            if (a.geometry) {
                const double kc = count * Kamp;
                atomicAdd(v + o + Wm, kc);
                if (Wm) atomicAdd(v + o + Wm + 1 + imu, kc);
                unsigned int *c = sc + (size_t)(iK - blo) * a.ncnt;
                atomicAdd(c, (unsigned)count);
                if (Wm) atomicAdd(c + 1 + imu, (unsigned)count);
            }
        } else {
            if (a.do_pl && sums)
                for (int l = 0; l < 4; ++l) atomicAdd(bins + (size_t)iK * 4 + l, pl[l]);
            if (Wm && sums) atomicAdd(bins + offPmu + (size_t)iK * Wm + imu, d2);
            if (a.geometry) {
                const double kc = count * Kamp;
                atomicAdd(bins + offK + iK, kc);
                if (Wm) atomicAdd(bins + offKmu + (size_t)iK * Wm + imu, kc);
                atomicAdd(cnts + iK, (unsigned long long)count);
                if (Wm) atomicAdd(cnts + a.NK + (size_t)iK * Wm + imu, (unsigned long long)count);
            }
        }
    }
    if (!windowed) return;
    __syncthreads();
    for (int t = threadIdx.x; t < wn * a.nvals; t += blockDim.x) {
        const double v = sv[t];
        if (v == 0.0) continue;
        const int w = t / a.nvals, q = t - w * a.nvals;
        const int iK = blo + w;
        int qq = q;
        if (a.do_pl) {
            if (qq < 4) { atomicAdd(bins + (size_t)iK * 4 + qq, v); continue; }
            qq -= 4;
        }
        if (qq < Wm) { atomicAdd(bins + offPmu + (size_t)iK * Wm + qq, v); continue; }
        qq -= Wm;
        if (qq == 0) { atomicAdd(bins + offK + iK, v); continue; }
        atomicAdd(bins + offKmu + (size_t)iK * Wm + (qq - 1), v);
    }
    if (a.geometry) {
        for (int t = threadIdx.x; t < wn * a.ncnt; t += blockDim.x) {
            const unsigned v = sc[t];
            if (!v) continue;
            const int w = t / a.ncnt, q = t - w * a.ncnt;
            const int iK = blo + w;
            if (q == 0) atomicAdd(cnts + iK, (unsigned long long)v);
            else atomicAdd(cnts + a.NK + (size_t)iK * Wm + (q - 1), (unsigned long long)v);
        }
    }
}

static BinArgs make_args(const fgk_binspec *b) {
    BinArgs a;
    a.Ng = b->Ng;
    a.NK = b->NK;
    a.Nmu = b->Nmu;
    a.do_pl = b->do_pl;
    a.geometry = b->geometry;
    const int Wm = b->Nmu > 0 ? b->Nmu + 1 : 0;
    a.nvals = (b->do_pl ? 4 : 0) + Wm + (b->geometry ? 1 + Wm : 0);
    a.ncnt = b->geometry ? 1 + Wm : 0;
    size_t per = (size_t)a.nvals * 8 + (size_t)a.ncnt * 4;
    a.wmax = per ? (int)(BIN_LDS_BYTES / per) : 1;
    if (a.wmax < 1) a.wmax = 1;
    a.dKinv = b->dKinv;
    a.l0 = b->los[0];
    a.l1 = b->los[1];
    a.l2 = b->los[2];
    return a;
}

static unsigned bin_blocks(int Ng) {
    const int Nk = Ng / 2 + 1;
    long n = (long)((Ng + BI - 1) / BI) * ((Ng + BJ - 1) / BJ) * ((Nk + BK - 1) / BK);
    return (unsigned)n;
}

static size_t bin_lds(const BinArgs &a) {
    return (size_t)a.wmax * a.nvals * 8 + (size_t)a.wmax * a.ncnt * 4 + 16;
}

extern "C" int fgk_bin_pgrid_f64(const double *f, const fgk_binspec *b, double *bins_dev,
                                 int64_t *counts_dev, fgk_stream_t s) {
    BinArgs a = make_args(b);
    hipLaunchKernelGGL(k_bin<0>, dim3(bin_blocks(b->Ng)), dim3(BIN_THREADS), bin_lds(a),
                       (hipStream_t)s, f, (const float *)nullptr, (const float *)nullptr, a,
                       (const double *)nullptr, (const double *)nullptr, 0.0, 0.0, 0.0, bins_dev,
                       (unsigned long long *)counts_dev);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_bin_geometry(const fgk_binspec *b, double *bins_dev, int64_t *counts_dev,
                                fgk_stream_t s) {
    if (!b->geometry) {
        fgk_set_error("fgk_bin_geometry: binspec without geometry");
        return -2;
    }
    BinArgs a = make_args(b);
    hipLaunchKernelGGL(k_bin<2>, dim3(bin_blocks(b->Ng)), dim3(BIN_THREADS), bin_lds(a),
                       (hipStream_t)s, (const double *)nullptr, (const float *)nullptr,
                       (const float *)nullptr, a, (const double *)nullptr,
                       (const double *)nullptr, 0.0, 0.0, 0.0, bins_dev,
                       (unsigned long long *)counts_dev);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_fused_bin_f32(const float *A, const float *B, const fgk_binspec *b,
                                 const double *phase_dev, const double *winv_dev, double norm,
                                 double Vinv, double Pshot, double *bins_dev, int64_t *counts_dev,
                                 fgk_stream_t s) {
    BinArgs a = make_args(b);
    hipLaunchKernelGGL(k_bin<1>, dim3(bin_blocks(b->Ng)), dim3(BIN_THREADS), bin_lds(a),
                       (hipStream_t)s, (const double *)nullptr, A, B, a, phase_dev, winv_dev,
                       norm, Vinv, Pshot, bins_dev, (unsigned long long *)counts_dev);
    FGK_LAUNCHED();
    return 0;
}
