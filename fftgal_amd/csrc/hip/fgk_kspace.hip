// k-space kernels on gfx950.
//
// Per-function FP64 kernels behind the drop-in API:
//   k_offset_phase_f64  src/fft.c:196-233 (separable table, 4-term product)
//   k_deconv_f64        src/fft.c:236-254 (PCS window W^-1 per axis)
//   k_P_f64             src/power.c:12-27 (|delta|^2/V - Pshot, DC = 0)
// and one binning kernel used three ways:
//   MODE 0: bins P(k) held in the FP64 grid (src/power.c:52-67, 125-138)
//   MODE 1: the fused fast-path kernel -- reads the FP32 spectra of the plain
//           and dual grids (packed half-spectrum layout: Ng/2 columns, the
//           k = 0 column holding F(X0) + i F(XN), separated here with the
//           mirror mode), applies the dual grid's offset phase, the
//           interlace average (src/fft.c:270-271), the deconvolution, forms
//           |delta|^2/V - Pshot in FP64 and bins it, without writing a grid.
//   MODE 2: the bin geometry alone (K sums and mode counts).
// Binning: each workgroup owns a 4 x 4 x 64 block of modes, whose |k| spans
// a short window of k bins; the Legendre (l = 0,2,4,6) and mu-wedge sums of
// that window are accumulated in LDS (FP64) and flushed with one global
// FP64 atomic per bin and quantity.  Bin and wedge indices are computed with
// the reference's FP64 expressions (no contraction, IEEE sqrt and division),
// so mode counts match the reference exactly.
#include <hip/hip_runtime.h>
#include <math.h>
#include <mutex>
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

// The per-mode kernels walk the Ng^2 (Ng/2+1) modes with grid-stride loops:
// at Ng = 2048 there are more modes than one dispatch may have work-items
// (2^32).
__global__ void k_offset_phase_f64(double *f, int Ng, const double *__restrict__ ph) {
#pragma clang fp contract(off)
    const long nmodes = (long)Ng * Ng * (Ng / 2 + 1);
    const long stride = (long)gridDim.x * blockDim.x;
    for (long m = blockIdx.x * (long)blockDim.x + threadIdx.x; m < nmodes; m += stride) {
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
}

__global__ void k_deconv_f64(double *f, int Ng, const double *__restrict__ winv) {
#pragma clang fp contract(off)
    const long nmodes = (long)Ng * Ng * (Ng / 2 + 1);
    const long stride = (long)gridDim.x * blockDim.x;
    for (long m = blockIdx.x * (long)blockDim.x + threadIdx.x; m < nmodes; m += stride) {
        int i, j, k;
        mode_ijk(m, Ng, i, j, k);
        double w3 = winv[i] * winv[j] * winv[k];
        f[2 * m] *= w3;
        f[2 * m + 1] *= w3;
    }
}

__global__ void k_P_f64(double *f, int Ng, double Vinv, double Pshot) {
#pragma clang fp contract(off)
    const long nmodes = (long)Ng * Ng * (Ng / 2 + 1);
    const long stride = (long)gridDim.x * blockDim.x;
    for (long m = blockIdx.x * (long)blockDim.x + threadIdx.x; m < nmodes; m += stride) {
        if (m == 0) {
            f[0] = 0;
            f[1] = 0;
            continue;
        }
        double re = f[2 * m], im = f[2 * m + 1];
        double d2 = re * re + im * im;
        f[2 * m] = d2 * Vinv - Pshot;
        f[2 * m + 1] = 0;
    }
}

// Super-sample filter of src/app/DLsb.c:66-88 (kind 0: sub-box window with
// its phase, tables wamp[Ng] and wph[3 Ng][2] built on the host as the
// reference builds them) and src/app/DLss.c:62-82 (kind 1: spherical top
// hat of K R = Kamp pi sqrt(3) / Nsub), times L_ell(mu) for ell = 0, 2, 4;
// Re of the DC mode set to 0 and the DC mode otherwise untouched.  The
// frequencies carry the Nyquist index as -Ng/2 (the reference's Kval table).
__global__ void k_dl_filter_f64(double *f, int Ng, int kind, int Nsub, double l0, double l1,
                                double l2, int ell, const double *__restrict__ wamp,
                                const double *__restrict__ wph) {
#pragma clang fp contract(off)
    const long nmodes = (long)Ng * Ng * (Ng / 2 + 1);
    const long stride = (long)gridDim.x * blockDim.x;
    for (long m = blockIdx.x * (long)blockDim.x + threadIdx.x; m < nmodes; m += stride) {
        if (m == 0) {
            f[0] = 0;
            continue;
        }
        int i, j, k;
        mode_ijk(m, Ng, i, j, k);
        const double Ki = 2 * i < Ng ? i : i - Ng, Kj = 2 * j < Ng ? j : j - Ng,
                     Kk = 2 * k < Ng ? k : k - Ng;
        const double Kamp = sqrt(Ki * Ki + Kj * Kj + Kk * Kk);
        const double q = (Ki * l0 + Kj * l1 + Kk * l2) / Kamp;
        const double mu2 = q * q;
        const double leg = ell == 0 ? 1.0
                         : ell == 2 ? 1.5 * mu2 - 0.5
                                    : (4.375 * mu2 - 3.75) * mu2 + 0.375;
        double *c = f + 2 * m;
        const double fr = c[0], fi = c[1];
        if (kind == 0) {
            const double w3 = wamp[i] * wamp[j] * wamp[k];
            const double wr = wph[2 * (i + j + k)] * w3, wi = wph[2 * (i + j + k) + 1] * w3;
            c[0] = leg * (fr * wr - fi * wi);
            c[1] = leg * (fr * wi + fi * wr);
        } else {
            const double x = Kamp * M_PI * 1.73205080756887729352744634151 / Nsub;
            const double W = 3 * (sin(x) - x * cos(x)) / (x * x * x);
            c[0] = fr * (leg * W);
            c[1] = fi * (leg * W);
        }
    }
}

// out[(a*Nsub + b)*Nsub + c] = f[(a*Ngsub, b*Ngsub, c*Ngsub)] / bias (padded layout)
__global__ void k_dl_gather(const double *f, int Ng, int Nsub, double bias, double *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Nsub * Nsub * Nsub) return;
    const int a = t / (Nsub * Nsub), b = (t / Nsub) % Nsub, c = t % Nsub;
    const int gs = Ng / Nsub;
    const long Ng_pad = 2 * (Ng / 2 + 1);
    out[t] = f[((long)a * gs * Ng + (long)b * gs) * Ng_pad + (long)c * gs] / bias;
}

// Window deconvolution of a sub-volume correlation function on the real
// grid (padded layout), the host loops of src/app/Psb.c:74-83 (kind 0:
// lags |i|,|j|,|k| < gs, factor gs^3 / prod(gs - |lag|)) and
// src/app/Pss.c:71-84 (kind 1: lags |i|,|j|,|k| < ceil(sqrt(3) gs) with
// i^2 + j^2 + k^2 < 0.99 * 3 gs^2, factor 1 / ((1 - d)^2 (1 + d/2)),
// d = |lag| / (sqrt(3) gs)).  A lag reaches cell (i + Ng) % Ng; when two
// lags of one axis reach the same cell (the sphere's range exceeds Ng),
// the cell takes both factors, in the reference's loop order.
__global__ void k_subvolume_winv_f64(double *f, int Ng, int kind, int gs) {
#pragma clang fp contract(off)
    const long ncell = (long)Ng * Ng * Ng;
    const long stride = (long)gridDim.x * blockDim.x;
    const long Ng_pad = 2 * (Ng / 2 + 1);
    const double D2 = 3.0 * gs * gs;
    const int bnd = kind == 0 ? gs : (int)ceil(sqrt(D2));
    for (long m = blockIdx.x * (long)blockDim.x + threadIdx.x; m < ncell; m += stride) {
        const int x = (int)(m / ((long)Ng * Ng)), y = (int)((m / Ng) % Ng), z = (int)(m % Ng);
        // lags in (-bnd, bnd) congruent to the cell index, ascending
        int li[4], lj[4], lk[4], ni = 0, nj = 0, nk = 0;
        for (int q = -2; q <= 2; ++q) {
            const int a = x + q * Ng, b = y + q * Ng, c = z + q * Ng;
            if (a > -bnd && a < bnd) li[ni++] = a;
            if (b > -bnd && b < bnd) lj[nj++] = b;
            if (c > -bnd && c < bnd) lk[nk++] = c;
        }
        if (!ni || !nj || !nk) continue;
        double *cell = f + ((long)x * Ng + y) * Ng_pad + z;
        double v = *cell;
        for (int ia = 0; ia < ni; ++ia)
            for (int ib = 0; ib < nj; ++ib)
                for (int ic = 0; ic < nk; ++ic) {
                    const int i = li[ia], j = lj[ib], k = lk[ic];
                    if (kind == 0) {
                        const double g3 = (double)gs * gs * gs;
                        const double w = g3 / ((double)(gs - abs(i)) * (gs - abs(j)) *
                                               (gs - abs(k)));
                        v *= w;
                    } else {
                        const double d2 = (double)i * i + (double)j * j + (double)k * k;
                        if (d2 < 0.99 * D2) {
                            const double dD = sqrt(d2 / D2);
                            v *= 1 / ((1 - dD) * (1 - dD) * (1 + 0.5 * dD));
                        }
                    }
                }
        *cell = v;
    }
}

extern "C" int fgk_subvolume_winv_f64(double *f, int Ng, int kind, int Nsub, fgk_stream_t s);

// workgroups for a grid-stride walk over n elements with 256 threads
static unsigned stride_blocks(long n) {
    long b = (n + 255) / 256;
    const long cap = 1l << 20;          // 2^28 work-items per dispatch
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

extern "C" int fgk_offset_phase_f64(double *f, int Ng, const double *phase_dev, fgk_stream_t s) {
    long n = (long)Ng * Ng * (Ng / 2 + 1);
    hipLaunchKernelGGL(k_offset_phase_f64, dim3(stride_blocks(n)), dim3(256), 0,
                       (hipStream_t)s, f, Ng, phase_dev);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_deconv_f64(double *f, int Ng, const double *winv_dev, fgk_stream_t s) {
    long n = (long)Ng * Ng * (Ng / 2 + 1);
    hipLaunchKernelGGL(k_deconv_f64, dim3(stride_blocks(n)), dim3(256), 0, (hipStream_t)s, f,
                       Ng, winv_dev);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_dl_filter_f64(double *f, int Ng, int kind, int Nsub, const double los[3],
                                 int ell, const double *wamp_dev, const double *wph_dev,
                                 fgk_stream_t s) {
    const long n = (long)Ng * Ng * (Ng / 2 + 1);
    hipLaunchKernelGGL(k_dl_filter_f64, dim3(stride_blocks(n)), dim3(256), 0, (hipStream_t)s, f,
                       Ng, kind, Nsub, los[0], los[1], los[2], ell, wamp_dev, wph_dev);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_subvolume_winv_f64(double *f, int Ng, int kind, int Nsub, fgk_stream_t s) {
    const long n = (long)Ng * Ng * Ng;
    hipLaunchKernelGGL(k_subvolume_winv_f64, dim3(stride_blocks(n)), dim3(256), 0,
                       (hipStream_t)s, f, Ng, kind, Ng / Nsub);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_dl_gather_f64(const double *f, int Ng, int Nsub, double bias, double *out_dev,
                                 fgk_stream_t s) {
    const int n = Nsub * Nsub * Nsub;
    hipLaunchKernelGGL(k_dl_gather, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)s, f, Ng,
                       Nsub, bias, out_dev);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_P_f64(double *f, int Ng, double Vinv, double Pshot, fgk_stream_t s) {
    long n = (long)Ng * Ng * (Ng / 2 + 1);
    hipLaunchKernelGGL(k_P_f64, dim3(stride_blocks(n)), dim3(256), 0, (hipStream_t)s, f, Ng,
                       Vinv, Pshot);
    FGK_LAUNCHED();
    return 0;
}

// ------------------------------------------------------------- binning

#define BIN_THREADS 256
// block of modes per workgroup of the bin kernel: BI x BJ x BK (i, j, k),
// 1024 modes; BK is the tunable (fgk_binspec.bk), 16 by default
template <int BKT> struct BinBlock {
    static constexpr int BK = BKT;
    static constexpr int BI = BKT >= 64 ? 4 : 8;
    static constexpr int BJ = 1024 / BKT / BI;
};
#define BIN_LDS_BYTES (12 * 1024)

extern "C" size_t fgk_bins_doubles(const fgk_binspec *b) {
    size_t W = b->Nmu > 0 ? (size_t)b->Nmu + 1 : 0;
    size_t n = (size_t)b->NK * (4 + W);
    if (b->geometry) n += (size_t)b->NK * (1 + W);
    return n;
}

// (H 2^32 + L) 2^(e - FGK_BIN_FIX_BITS) of one fixed-point pair, exact up
// to the one rounding to FP64
extern "C" double fgk_bin_fixed_value(const int64_t *hl, int e) {
    const __int128 y = ((__int128)hl[0] << 32) + (__int128)(uint64_t)hl[1];
    return ldexp((double)y, e - FGK_BIN_FIX_BITS);
}

extern "C" size_t fgk_bins_counts(const fgk_binspec *b) {
    size_t W = b->Nmu > 0 ? (size_t)b->Nmu + 1 : 0;
    return b->geometry ? (size_t)b->NK * (1 + W) : 0;
}

struct BinArgs {
    int multi;        // > 0: one launch over this many catalogs of any sides (jobs)
    int Ng;
    int NK;
    int Nmu;
    int do_pl;
    int geometry;
    int nvals;        // doubles per bin in LDS
    int ncnt;         // counts per bin in LDS
    int wmax;         // LDS window capacity in bins
    int fs_n1;        // spectra planes in four-step order (0: natural)
    int fx;           // sums: 0 FP64; 1, 2 exact fixed point (see fix_int)
    double dKinv;
    double l0, l1, l2;
};

// i-plane holding frequency i of the packed spectra
__device__ __forceinline__ int iplane(const BinArgs &a, int i) {
    return a.fs_n1 ? (i % a.fs_n1) * (a.Ng / a.fs_n1) + i / a.fs_n1 : i;
}

// signed frequency of index i with the Nyquist index carrying -Ng/2
// (src/power.c:46-51)
__device__ __forceinline__ int kval(int i, int Ng) { return (2 * i < Ng) ? i : i - Ng; }

// Exact bin sums.  With a.fx set, a value v is added as the integer
// Y = rint(v * qs), qs = 2^(FGK_BIN_FIX_BITS - e), e the binary exponent of
// the catalog's Pshot (FGK_BIN_K_EXP for the geometry's K sums, in units of
// KF).  The global sums are pairs of int64 accumulators (H, L), Y adding
// H = floor(Y / 2^32) and L = Y - H 2^32 in [0, 2^32), so every sum is an
// exact integer whatever the order of the atomics and reruns give
// bit-identical bins; the host forms (H_sum 2^32 + L_sum) / qs.  The
// quantum is 2^-24 of Pshot's power of two for power sums (the rounding of
// one value, the sum of at most 4 modes, is below 3e-8 Pshot) and 2^-52 KF
// for K sums.
//   fx = 1 (fast path): the LDS window holds one int64 per value, Y from one
//     FMA against a magic constant while |Y| < 2^50 (a block adds at most
//     1024 such values to one entry, so the entry stays below 2^60), split
//     into (H, L) when the window is flushed; larger values, |v| above
//     2^26 Pshot (lognormal mocks with 1 Mpc/h seed cells reach 1e7 Pshot in
//     small boxes), go to the global pair directly.
//   fx = 2 (geometry): the LDS window holds pairs, every value split.
__host__ __device__ __forceinline__ int fix_exponent(double Pshot) {
    return (Pshot > 0.0 && Pshot < 1e300) ? ilogb(Pshot) : 0;
}

__device__ __forceinline__ double fix_scale(int e) { return ldexp(1.0, FGK_BIN_FIX_BITS - e); }

extern "C" int fgk_bin_fixed_exponent(double Pshot) { return fix_exponent(Pshot); }

// Y of any magnitude into a pair (H, L); out of line, off the common path
__device__ __attribute__((noinline)) void pair_add(unsigned long long *acc, double v, double qs) {
#pragma clang fp contract(off)
    const double Y = rint(v * qs);
    const double H = floor(Y * 0x1p-32);
    const unsigned long long h = (unsigned long long)(long long)H;
    const unsigned long long l = (unsigned long long)(unsigned)(Y - H * 0x1p32);
    if (h) atomicAdd(acc, h);
    if (l) atomicAdd(acc + 1, l);
}

// rint(v * qs) for |v * qs| < 2^50: the FMA rounds v * qs + 1.5 2^52 once to
// an integer (unit spacing there), whose bits less those of 1.5 2^52 are Y
__device__ __forceinline__ unsigned long long fix_int(double v, double qs) {
    const double M = 0x1.8p52;
    return (unsigned long long)(__double_as_longlong(__builtin_fma(v, qs, M)) -
                                __double_as_longlong(M));
}

// int64 words of one value in the LDS window
__device__ __host__ __forceinline__ int lds_words(int fx) { return fx == 2 ? 2 : 1; }

// add v to a bin's sums: entry lidx of its LDS window row lrow, or (lrow
// null) entry gidx of the global sums
__device__ __forceinline__ void put(const BinArgs &a, void *gbins, double qs, void *lrow,
                                    int lidx, size_t gidx, double v) {
    if (a.fx == 0) {
        atomicAdd(lrow ? (double *)lrow + lidx : (double *)gbins + gidx, v);
        return;
    }
    if (a.fx == 1 && lrow) {
        if (fabs(v * qs) < 0x1p50) {
            atomicAdd((unsigned long long *)lrow + lidx, fix_int(v, qs));
            return;
        }
        lrow = nullptr;
    }
    pair_add(lrow ? (unsigned long long *)lrow + 2 * lidx
                  : (unsigned long long *)gbins + 2 * gidx, v, qs);
}

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

// where one workgroup's modes are binned: LDS windows of bins (sv:
// [wtot][nvals] sums, sc: [wtot][ncnt] counts) -- window 0 holds bins
// [blo, blo + wn), window 1 (used by the k = 0 / k = Ng/2 plane kernel for
// its Nyquist modes) bins [blo2, blo2 + wtot - wn) -- or, when the windows
// do not fit, straight into the global bins
struct BinOut {
    void *sv;                  // LDS sums: FP64, or int64 pairs (a.fx)
    unsigned int *sc;
    int blo, wn, blo2, wtot;
    bool windowed;
    void *bins;                // global sums, the same form
    unsigned long long *cnts;
    double qs;                 // fixed-point scale (a.fx)
};

// bin, wedge and sums of one mode of power P, given its count (1 on the
// k = 0 and k = Ng/2 planes, else 2), |k| in units of KF (Kamp) and
// los . k (dot), as src/power.c:52-67 (Pl) and 125-143 (Pmu) compute them,
// in FP64 without contraction.  False when the mode lies past the last bin.
struct ModeVals {
    int iK, imu, count;
    double pl[4];     // Legendre-weighted count * P (do_pl and SUMS)
    double d2;        // count * P
    double kc;        // count * |k|
};

template <bool SUMS>
__device__ __forceinline__ bool mode_vals(const BinArgs &a, int count, double Kamp, double dot,
                                          double P, ModeVals &m) {
#pragma clang fp contract(off)
    m.count = count;
    m.d2 = P * count;
    m.iK = (int)floor(Kamp * a.dKinv);
    if (m.iK >= a.NK) return false;   // cannot happen for NK of src/power.c:36; keeps LDS in bounds
    const double mu = dot / Kamp;
    m.imu = 0;
    if (a.Nmu > 0) m.imu = (int)floor(fabs(mu * a.Nmu));
    if (a.do_pl && SUMS) {
        const double mu2 = mu * mu;
        m.pl[0] = m.d2;
        m.pl[1] = m.d2 * (1.5 * mu2 - 0.5);
        m.pl[2] = m.d2 * ((4.375 * mu2 - 3.75) * mu2 + 0.375);
        m.pl[3] = m.d2 * (((14.4375 * mu2 - 19.6875) * mu2 + 6.5625) * mu2 - 0.3125);
    }
    m.kc = count * Kamp;
    return true;
}

// LDS window row of bin iK (its counts in *crow), or nullptr when the bin
// lies outside the window and goes straight to the global bins
__device__ __forceinline__ void *window_row(const BinArgs &a, const BinOut &o, int iK, int win,
                                            unsigned int **crow) {
    const int slot = win ? o.wn + iK - o.blo2 : iK - o.blo;
    if (o.windowed && slot >= (win ? o.wn : 0) && slot < (win ? o.wtot : o.wn)) {
        *crow = o.sc + (size_t)slot * a.ncnt;
        return (void *)((unsigned long long *)o.sv + (size_t)slot * a.nvals * lds_words(a.fx));
    }
    return nullptr;
}

// the bin part of the sums of one mode or of a run of modes in bin iK:
// Legendre power sums (SUMS), K sum and mode count (a.geometry).  LDS row
// of a bin: [P_l x4 (do_pl)][wedge P x Wm][K][wedge K x Wm]; counts
// [N][wedge N x Wm].  win selects the LDS window (0 or 1).
template <bool SUMS>
__device__ __forceinline__ void emit_bin(const BinArgs &a, const BinOut &o, int iK,
                                         const double pl[4], double kc, int count, int win) {
    const int Wm = a.Nmu > 0 ? a.Nmu + 1 : 0;
    unsigned int *c = nullptr;
    void *v = window_row(a, o, iK, win, &c);
    const size_t offK = (size_t)a.NK * 4 + (size_t)a.NK * Wm;
    if (a.do_pl && SUMS)
        for (int l = 0; l < 4; ++l) put(a, o.bins, o.qs, v, l, (size_t)iK * 4 + l, pl[l]);
    if (a.geometry) {
        put(a, o.bins, o.qs, v, (a.do_pl ? 4 : 0) + Wm, offK + iK, kc);
        if (v) atomicAdd(c, (unsigned)count);
        else atomicAdd(o.cnts + iK, (unsigned long long)count);
    }
}

// the wedge part: power sum (SUMS), K sum and count of wedge imu of bin iK
template <bool SUMS>
__device__ __forceinline__ void emit_wedge(const BinArgs &a, const BinOut &o, int iK, int imu,
                                           double d2, double kc, int count, int win) {
    const int Wm = a.Nmu > 0 ? a.Nmu + 1 : 0;
    if (!Wm) return;
    unsigned int *c = nullptr;
    void *v = window_row(a, o, iK, win, &c);
    const int q = a.do_pl ? 4 : 0;
    const size_t offPmu = (size_t)a.NK * 4;
    const size_t offKmu = offPmu + (size_t)a.NK * Wm + (size_t)a.NK;
    if (SUMS) put(a, o.bins, o.qs, v, q + imu, offPmu + (size_t)iK * Wm + imu, d2);
    if (a.geometry) {
        put(a, o.bins, o.qs, v, q + Wm + 1 + imu, offKmu + (size_t)iK * Wm + imu, kc);
        if (v) atomicAdd(c + 1 + imu, (unsigned)count);
        else atomicAdd(o.cnts + a.NK + (size_t)iK * Wm + imu, (unsigned long long)count);
    }
}

// one mode straight into the bins
template <bool SUMS>
__device__ __forceinline__ void bin_mode(const BinArgs &a, const BinOut &o, int count,
                                         double Kamp, double dot, double P, int win) {
    ModeVals m;
    if (!mode_vals<SUMS>(a, count, Kamp, dot, P, m)) return;
    emit_bin<SUMS>(a, o, m.iK, m.pl, m.kc, m.count, win);
    emit_wedge<SUMS>(a, o, m.iK, m.imu, m.d2, m.kc, m.count, win);
}

// Sums of a thread's consecutive modes that fall in one bin (and one wedge)
// are added in registers and go to the bins once per run: along a row of
// the spectrum |k| grows with k, so a thread's consecutive modes mostly
// share a bin, and a wave's lanes then mostly hold different bins (fewer
// LDS atomics, and fewer of them on one address).
template <bool SUMS>
struct BinRun {
    int iK = -1, imu = -1, cnt = 0, wcnt = 0;
    double pl[4] = {0.0, 0.0, 0.0, 0.0}, kc = 0.0, wd2 = 0.0, wkc = 0.0;

    __device__ __forceinline__ void flush(const BinArgs &a, const BinOut &o, int win) {
        if (iK < 0) return;
        emit_bin<SUMS>(a, o, iK, pl, kc, cnt, win);
        emit_wedge<SUMS>(a, o, iK, imu, wd2, wkc, wcnt, win);
    }

    __device__ __forceinline__ void add(const BinArgs &a, const BinOut &o, const ModeVals &m,
                                        int win) {
#pragma clang fp contract(off)
        if (m.iK != iK) {
            flush(a, o, win);
            iK = m.iK;
            for (int l = 0; l < 4; ++l) pl[l] = m.pl[l];
            kc = m.kc;
            cnt = m.count;
            imu = m.imu;
            wd2 = m.d2;
            wkc = m.kc;
            wcnt = m.count;
            return;
        }
        if (a.do_pl && SUMS)
            for (int l = 0; l < 4; ++l) pl[l] += m.pl[l];
        kc += m.kc;
        cnt += m.count;
        if (m.imu != imu) {
            emit_wedge<SUMS>(a, o, iK, imu, wd2, wkc, wcnt, win);
            imu = m.imu;
            wd2 = m.d2;
            wkc = m.kc;
            wcnt = m.count;
            return;
        }
        wd2 += m.d2;
        wkc += m.kc;
        wcnt += m.count;
    }
};

// |k| of a mode in units of KF, sqrt(fi^2 + fj^2 + fk^2) as src/power.c:55
__device__ __forceinline__ double kamp(int fi, int fj, int fk) {
#pragma clang fp contract(off)
    const double a = fi, b = fj, c = fk;
    return sqrt(a * a + b * b + c * c);
}

// los . k in the reference's order of operations (src/power.c:56)
__device__ __forceinline__ double los_dot(const BinArgs &a, int fi, int fj, int fk) {
#pragma clang fp contract(off)
    return (double)fi * a.l0 + (double)fj * a.l1 + (double)fk * a.l2;
}

// per-k factors of the fused power: dual-grid phase along k and the k
// deconvolution weight
struct KFac {
    double pr, pi, w;
};

// power of one mode of the fast path from the plain (a) and dual (b)
// spectra: dual times the offset phase (src/fft.c:217-232, here as the
// product of the (i, j) phase q and the k phase), interlace average
// (src/fft.c:270-271), deconvolution (src/fft.c:249-253),
// |delta|^2 / V - Pshot (src/power.c:18-24)
__device__ __forceinline__ double fused_power(double ar, double ai, double br, double bi, bool dual,
                                              double qr, double qi, double wij, const KFac &kf,
                                              double na, double nb, double Vinv, double Pshot) {
    double dr = ar * na, di = ai * na;
    if (dual) {
        const double pr = qr * kf.pr - qi * kf.pi, pim = qr * kf.pi + qi * kf.pr;
        const double xr = br * nb, xi = bi * nb;
        const double er = xr * pr - xi * pim, ei = xr * pim + xi * pr;
        dr = 0.5 * (er + dr);
        di = 0.5 * (ei + di);
    }
    const double w3 = wij * kf.w;
    dr *= w3;
    di *= w3;
    return (dr * dr + di * di) * Vinv - Pshot;
}

// k-bin range [lo, hi] of modes with |frequency| per axis in [lo_d, hi_d]
__device__ __forceinline__ void bin_window(const BinArgs &a, int lo0, int hi0, int lo1, int hi1,
                                           int lo2, int hi2, int &lo, int &hi) {
    const double kmin = sqrt((double)lo0 * lo0 + (double)lo1 * lo1 + (double)lo2 * lo2);
    const double kmax = sqrt((double)hi0 * hi0 + (double)hi1 * hi1 + (double)hi2 * hi2);
    lo = (int)floor(kmin * a.dKinv);
    hi = (int)floor(kmax * a.dKinv);
    if (lo > a.NK - 1) lo = a.NK - 1;
    if (hi > a.NK - 1) hi = a.NK - 1;
}

__device__ __forceinline__ void window_clear(const BinArgs &a, const BinOut &o) {
    if (o.windowed) {
        for (int t = threadIdx.x; t < lds_words(a.fx) * o.wtot * a.nvals; t += blockDim.x)
            ((unsigned long long *)o.sv)[t] = 0ull;   // also FP64 +0
        for (int t = threadIdx.x; t < o.wtot * a.ncnt; t += blockDim.x) o.sc[t] = 0u;
    }
    __syncthreads();
}

// global index of entry q of LDS window row w
__device__ __forceinline__ size_t flush_index(const BinArgs &a, const BinOut &o, int w, int q) {
    const int Wm = a.Nmu > 0 ? a.Nmu + 1 : 0;
    const size_t offPmu = (size_t)a.NK * 4;
    const size_t offK = offPmu + (size_t)a.NK * Wm;
    const size_t offKmu = offK + (size_t)a.NK;
    const int iK = w < o.wn ? o.blo + w : o.blo2 + w - o.wn;
    if (a.do_pl) {
        if (q < 4) return (size_t)iK * 4 + q;
        q -= 4;
    }
    if (q < Wm) return offPmu + (size_t)iK * Wm + q;
    q -= Wm;
    if (q == 0) return offK + iK;
    return offKmu + (size_t)iK * Wm + (q - 1);
}

// one global atomic per non-zero LDS window entry (fixed point: per
// non-zero half of the pair; integer sums need no conversion)
__device__ __forceinline__ void window_flush(const BinArgs &a, const BinOut &o) {
    if (!o.windowed) return;
    __syncthreads();
    const int Wm = a.Nmu > 0 ? a.Nmu + 1 : 0;
    if (a.fx == 2) {
        const unsigned long long *sv = (const unsigned long long *)o.sv;
        unsigned long long *gb = (unsigned long long *)o.bins;
        for (int t = threadIdx.x; t < 2 * o.wtot * a.nvals; t += blockDim.x) {
            const unsigned long long v = sv[t];
            if (!v) continue;
            const int e = t >> 1, w = e / a.nvals;
            atomicAdd(gb + 2 * flush_index(a, o, w, e - w * a.nvals) + (t & 1), v);
        }
    } else if (a.fx == 1) {
        // S = H 2^32 + L, L = the low 32 bits (two's complement)
        const long long *sv = (const long long *)o.sv;
        unsigned long long *gb = (unsigned long long *)o.bins;
        for (int t = threadIdx.x; t < o.wtot * a.nvals; t += blockDim.x) {
            const long long v = sv[t];
            if (!v) continue;
            const int w = t / a.nvals;
            unsigned long long *g = gb + 2 * flush_index(a, o, w, t - w * a.nvals);
            const unsigned long long h = (unsigned long long)(v >> 32), l = (unsigned long long)v & 0xffffffffull;
            if (h) atomicAdd(g, h);
            if (l) atomicAdd(g + 1, l);
        }
    } else {
        for (int t = threadIdx.x; t < o.wtot * a.nvals; t += blockDim.x) {
            const double v = ((const double *)o.sv)[t];
            if (v == 0.0) continue;
            const int w = t / a.nvals;
            atomicAdd((double *)o.bins + flush_index(a, o, w, t - w * a.nvals), v);
        }
    }
    if (a.geometry) {
        for (int t = threadIdx.x; t < o.wtot * a.ncnt; t += blockDim.x) {
            const unsigned v = o.sc[t];
            if (!v) continue;
            const int w = t / a.ncnt, q = t - w * a.ncnt;
            const int iK = w < o.wn ? o.blo + w : o.blo2 + w - o.wn;
            if (q == 0) atomicAdd(o.cnts + iK, (unsigned long long)v);
            else atomicAdd(o.cnts + a.NK + (size_t)iK * Wm + (q - 1), (unsigned long long)v);
        }
    }
}

__device__ __forceinline__ BinOut make_out(const BinArgs &a, double *sh, int blo, int bhi,
                                           int blo2, int bhi2, void *bins,
                                           unsigned long long *cnts, double qs) {
    BinOut o;
    o.blo = blo;
    o.wn = bhi - blo + 1;
    o.blo2 = blo2;
    o.wtot = o.wn + (bhi2 >= blo2 ? bhi2 - blo2 + 1 : 0);
    o.windowed = o.wtot <= a.wmax;
    o.sv = sh;
    o.sc = (unsigned int *)(sh + (size_t)a.wmax * a.nvals * lds_words(a.fx));
    o.bins = bins;
    o.cnts = cnts;
    o.qs = qs;
    return o;
}

// MODE 0: P(k) held in the FP64 padded grid fP (k <= Ng/2);
// MODE 1: fused fast path over the packed FP32 spectra A, B, modes with
//         0 < k < Ng/2 (the k = 0 and k = Ng/2 planes: k_bin_plane);
// MODE 2: geometry only (K sums and mode counts), padded k range.
// Each workgroup owns a BI x BJ x BK block of modes, one thread per k.
// A batch of catalogs on grids of one side (jobs != nullptr): blockIdx.y
// selects the catalog, whose spectra, bins, k binning and normalisations
// replace A, B, bins, a.NK, a.dKinv, norm, Vinv, Pshot and npart.
// With a.multi > 0 (one launch over catalogs of any sides) the catalog
// owning workgroup blk is found by its first workgroup (blk0, or pblk0 for
// the plane kernel), blk becomes the workgroup within it, and the job also
// supplies the grid side, the four-step split and the side's tables.
#define BIN_JOB(jobs, blk, plane)                                              \
    do {                                                                       \
        int jx_ = blockIdx.y;                                                  \
        if (a.multi) {                                                         \
            int lo_ = 0, hi_ = a.multi - 1;                                    \
            while (lo_ < hi_) {                                                \
                const int mid_ = (lo_ + hi_ + 1) >> 1;                         \
                const unsigned b0_ = plane ? jobs[mid_].pblk0 : jobs[mid_].blk0; \
                if (b0_ <= (unsigned)blk) lo_ = mid_; else hi_ = mid_ - 1;     \
            }                                                                  \
            jx_ = lo_;                                                         \
        }                                                                      \
        const fgk_binjob &J = jobs[jx_];                                       \
        A = J.A;                                                               \
        B = J.B;                                                               \
        bins = J.sums;                                                         \
        a.NK = J.NK;                                                           \
        a.dKinv = J.dKinv;                                                     \
        norm = J.norm;                                                         \
        Vinv = J.Vinv;                                                         \
        Pshot = J.Pshot;                                                       \
        npart = J.npart;                                                       \
        if (a.multi) {                                                         \
            blk -= (int)(plane ? J.pblk0 : J.blk0);                            \
            a.Ng = J.Ng;                                                       \
            a.fs_n1 = J.fs_n1;                                                 \
            ph = J.phase;                                                      \
            winv = J.winv;                                                     \
        }                                                                      \
    } while (0)

// For BK <= 32 the kernel is held to 96 VGPRs, five waves per SIMD (no
// spills; measured 0.61 -> 0.54 ms at 2^27 / Ng 512, BK = 32); wider
// blocks keep the compiler's choice (BK = 64 is slower when held).
template <int MODE, int BKT>
__global__ void __launch_bounds__(BIN_THREADS, (BKT <= 32 ? 5 : 1))
k_bin(const double *__restrict__ fP, const float *__restrict__ A, const float *__restrict__ B,
      BinArgs a, const double *__restrict__ ph, const double *__restrict__ winv, double norm,
      double Vinv, double Pshot, double npart, void *__restrict__ bins,
      unsigned long long *__restrict__ cnts, const fgk_binjob *__restrict__ jobs) {
#pragma clang fp contract(off)
    extern __shared__ double sh[];
    a.fx = MODE == 1 ? 1 : (MODE == 2 ? 2 : 0);   // as the host sized the LDS; constant here
    int blk = blockIdx.x;
    if (MODE == 1 && jobs) BIN_JOB(jobs, blk, false);
    constexpr int BI = BinBlock<BKT>::BI, BJ = BinBlock<BKT>::BJ, BK = BinBlock<BKT>::BK;
    const int Ng = a.Ng;
    const int Mh = Ng / 2;
    const int Nk = MODE == 1 ? Mh : Mh + 1;      // k columns walked
    const int nbk = (Nk + BK - 1) / BK;
    const int nbj = (Ng + BJ - 1) / BJ;
    const int bk = blk % nbk;
    blk /= nbk;
    const int bj = blk % nbj;
    const int bi = blk / nbj;
    const int i0 = bi * BI, j0 = bj * BJ, k0 = bk * BK;
    const int i1 = min(i0 + BI, Ng), j1 = min(j0 + BJ, Ng), k1 = min(k0 + BK, Nk);

    int lo0, hi0, lo1, hi1, lo2, hi2, blo, bhi;
    axis_range(i0, i1, Ng, lo0, hi0);
    axis_range(j0, j1, Ng, lo1, hi1);
    axis_range(k0, k1, Ng, lo2, hi2);
    bin_window(a, lo0, hi0, lo1, hi1, lo2, hi2, blo, bhi);
    const BinOut o = make_out(a, sh, blo, bhi, 0, -1, bins, cnts,
                              a.fx ? fix_scale(MODE == 1 ? fix_exponent(Pshot) : FGK_BIN_K_EXP)
                                   : 0.0);
    window_clear(a, o);

    // each thread takes KPT consecutive k of the rows it visits
    constexpr int KPT = 4, TPR = BK / KPT;
    const int kb = k0 + (threadIdx.x % TPR) * KPT;
    int cnt[KPT], fk[KPT];
    bool kl[KPT];
    KFac kf[KPT];
    // norm = V / Np (src/fft.c:125-127 with the H^3 of x2k), times 2^-S
    // where the FFT read the fixed-point grids unscaled.  The fixed-point
    // paint rounds without bias (dithered), so its expected mass is Np, as
    // the reference's.
    const double na = norm, nb = norm;
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int k = kb + q;
        fk[q] = kval(k, Ng);
        cnt[q] = (k == 0 || 2 * k == Ng) ? 1 : 2;
        kl[q] = k < k1 && !(MODE == 1 && k == 0);
        kf[q] = {1.0, 0.0, 1.0};
        if (MODE == 1 && k < k1) {
            kf[q].pr = ph[2 * (2 * Ng + k)];
            kf[q].pi = ph[2 * (2 * Ng + k) + 1];
            kf[q].w = winv[k];
        }
    }
    // the packed spectra are read 4 modes (two float4) at a time when the
    // rows keep that alignment
    const bool vec = MODE == 1 && (Mh & 1) == 0 && kb + KPT <= k1;
    BinRun<MODE != 2> run;
    for (int r = threadIdx.x / TPR; r < BI * BJ; r += BIN_THREADS / TPR) {
        const int i = i0 + r / BJ, j = j0 + r % BJ;
        if (kb >= k1 || i >= i1 || j >= j1) continue;
        const int fi = kval(i, Ng), fj = kval(j, Ng);
        float2 av[KPT], bv[KPT];
        double qr = 1.0, qi = 0.0, wij = 1.0;
        if (MODE == 1) {
            const long m = ((long)iplane(a, i) * Ng + j) * Mh + kb;
            if (vec) {
                const float4 *pa = (const float4 *)((const float2 *)A + m);
                const float4 a0 = pa[0], a1 = pa[1];
                av[0] = make_float2(a0.x, a0.y);
                av[1] = make_float2(a0.z, a0.w);
                av[2] = make_float2(a1.x, a1.y);
                av[3] = make_float2(a1.z, a1.w);
                if (B) {
                    const float4 *pb = (const float4 *)((const float2 *)B + m);
                    const float4 b0 = pb[0], b1 = pb[1];
                    bv[0] = make_float2(b0.x, b0.y);
                    bv[1] = make_float2(b0.z, b0.w);
                    bv[2] = make_float2(b1.x, b1.y);
                    bv[3] = make_float2(b1.z, b1.w);
                }
            } else {
#pragma unroll
                for (int q = 0; q < KPT; ++q) {
                    av[q] = bv[q] = make_float2(0.0f, 0.0f);
                    if (kb + q < k1) {
                        av[q] = ((const float2 *)A)[m + q];
                        if (B) bv[q] = ((const float2 *)B)[m + q];
                    }
                }
            }
            if (!B)
                for (int q = 0; q < KPT; ++q) bv[q] = make_float2(0.0f, 0.0f);
            // (i, j) parts of the dual phase and the deconvolution
            const double *p0 = ph + 2 * i, *p1 = ph + 2 * (Ng + j);
            qr = p0[0] * p1[0] - p0[1] * p1[1];
            qi = p0[0] * p1[1] + p0[1] * p1[0];
            wij = winv[i] * winv[j];
        }
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            if (!kl[q]) continue;
            if (i == 0 && j == 0 && kb + q == 0) continue;
            double P = 0.0;
            if (MODE == 0) {
                P = fP[2 * (((long)i * Ng + j) * Nk + kb + q)];
            } else if (MODE == 1) {
                P = fused_power(av[q].x, av[q].y, bv[q].x, bv[q].y, B != nullptr, qr, qi, wij,
                                kf[q], na, nb, Vinv, Pshot);
            }
            ModeVals mv;
            if (mode_vals<MODE != 2>(a, cnt[q], kamp(fi, fj, fk[q]), los_dot(a, fi, fj, fk[q]), P,
                                     mv))
                run.add(a, o, mv, 0);
        }
    }
    run.flush(a, o, 0);
    window_flush(a, o);
}

// The k = 0 column of the packed spectra holds Z = F(X0) + i F(XN); this
// kernel separates the k = 0 and k = Ng/2 planes with the mirror mode,
// X0 = (Z + conj Z(-i,-j)) / 2, XN = (Z - conj Z(-i,-j)) / 2i, and bins both
// (window 0: k = 0 modes, window 1: k = Ng/2 modes).  One thread per (i, j)
// of a PI x PJ block.
#define PI 16
#define PJ 16
__global__ void __launch_bounds__(PI * PJ)
k_bin_plane(const float *__restrict__ A, const float *__restrict__ B, BinArgs a,
            const double *__restrict__ ph, const double *__restrict__ winv, double norm,
            double Vinv, double Pshot, double npart, void *__restrict__ bins,
            const fgk_binjob *__restrict__ jobs) {
#pragma clang fp contract(off)
    extern __shared__ double sh[];
    a.fx = 1;   // as the host sized the LDS; constant here
    int blk = blockIdx.x;
    if (jobs) BIN_JOB(jobs, blk, true);
    const int Ng = a.Ng;
    const int Mh = Ng / 2;
    const int nbj = (Ng + PJ - 1) / PJ;
    const int i0 = (blk / nbj) * PI, j0 = (blk % nbj) * PJ;
    const int i1 = min(i0 + PI, Ng), j1 = min(j0 + PJ, Ng);
    int lo0, hi0, lo1, hi1, blo, bhi, blo2, bhi2;
    axis_range(i0, i1, Ng, lo0, hi0);
    axis_range(j0, j1, Ng, lo1, hi1);
    bin_window(a, lo0, hi0, lo1, hi1, 0, 0, blo, bhi);
    bin_window(a, lo0, hi0, lo1, hi1, Mh, Mh, blo2, bhi2);
    const BinOut o = make_out(a, sh, blo, bhi, blo2, bhi2, bins, nullptr,
                              fix_scale(fix_exponent(Pshot)));
    window_clear(a, o);
    const int i = i0 + threadIdx.x / PJ, j = j0 + threadIdx.x % PJ;
    const double na = norm, nb = norm;
    if (i < i1 && j < j1) {
        const int im = i ? Ng - i : 0, jm = j ? Ng - j : 0;
        const long m = ((long)iplane(a, i) * Ng + j) * Mh,
                   mm = ((long)iplane(a, im) * Ng + jm) * Mh;
        const float2 av = ((const float2 *)A)[m], am = ((const float2 *)A)[mm];
        float2 bv = make_float2(0.0f, 0.0f), bm = bv;
        if (B) {
            bv = ((const float2 *)B)[m];
            bm = ((const float2 *)B)[mm];
        }
        const double *p0 = ph + 2 * i, *p1 = ph + 2 * (Ng + j);
        const double qr = p0[0] * p1[0] - p0[1] * p1[1], qi = p0[0] * p1[1] + p0[1] * p1[0];
        const double wij = winv[i] * winv[j];
        const int fi = kval(i, Ng), fj = kval(j, Ng), fn = kval(Mh, Ng);
        const KFac k0f = {ph[2 * (2 * Ng)], ph[2 * (2 * Ng) + 1], winv[0]};
        const KFac knf = {ph[2 * (2 * Ng + Mh)], ph[2 * (2 * Ng + Mh) + 1], winv[Mh]};
        if (i != 0 || j != 0) {   // (0, 0, 0) is the DC mode
            const double P = fused_power(0.5 * ((double)av.x + am.x), 0.5 * ((double)av.y - am.y),
                                         0.5 * ((double)bv.x + bm.x), 0.5 * ((double)bv.y - bm.y),
                                         B != nullptr, qr, qi, wij, k0f, na, nb, Vinv, Pshot);
            bin_mode<true>(a, o, 1, kamp(fi, fj, 0), los_dot(a, fi, fj, 0), P, 0);
        }
        const double P = fused_power(0.5 * ((double)av.y + am.y), -0.5 * ((double)av.x - am.x),
                                     0.5 * ((double)bv.y + bm.y), -0.5 * ((double)bv.x - bm.x),
                                     B != nullptr, qr, qi, wij, knf, na, nb, Vinv, Pshot);
        bin_mode<true>(a, o, 1, kamp(fi, fj, fn), los_dot(a, fi, fj, fn), P, 1);
    }
    window_flush(a, o);
}

static int bin_bk(const fgk_binspec *b) {
    return (b->bk == 32 || b->bk == 64 || b->bk == 128) ? b->bk : 16;
}

static BinArgs make_args(const fgk_binspec *b, int fx) {
    const int BK = bin_bk(b);
    BinArgs a;
    a.multi = 0;
    a.Ng = b->Ng;
    a.NK = b->NK;
    a.Nmu = b->Nmu;
    a.do_pl = b->do_pl;
    a.geometry = b->geometry;
    const int Wm = b->Nmu > 0 ? b->Nmu + 1 : 0;
    a.nvals = (b->do_pl ? 4 : 0) + Wm + (b->geometry ? 1 + Wm : 0);
    a.ncnt = b->geometry ? 1 + Wm : 0;
    a.fx = fx;
    size_t per = (size_t)a.nvals * 8 * lds_words(fx) + (size_t)a.ncnt * 4;
    // room for the widest window of a BI x BJ x BK block near the grid's
    // corner (about BK + 4 bins at dK = KF) even with many wedges
    a.wmax = per ? (int)(BIN_LDS_BYTES / per) : 1;
    if (a.wmax < BK + 32) a.wmax = BK + 32;
    a.fs_n1 = b->fs_n1;
    a.dKinv = b->dKinv;
    a.l0 = b->los[0];
    a.l1 = b->los[1];
    a.l2 = b->los[2];
    return a;
}

// Nk: k columns walked (Ng/2 + 1 padded layout, Ng/2 packed)
template <int BKT> static unsigned bin_blocks(int Ng, int Nk) {
    constexpr int BI = BinBlock<BKT>::BI, BJ = BinBlock<BKT>::BJ, BK = BinBlock<BKT>::BK;
    long n = (long)((Ng + BI - 1) / BI) * ((Ng + BJ - 1) / BJ) * ((Nk + BK - 1) / BK);
    return (unsigned)n;
}

// launch k_bin<MODE, bk> with the block count of that shape
#define BIN_LAUNCH(MODE, bk, Ng, Nk, grid_y, lds, s, ...)                                    \
    do {                                                                                      \
        switch (bk) {                                                                         \
        case 16:                                                                              \
            hipLaunchKernelGGL((k_bin<MODE, 16>), dim3(bin_blocks<16>(Ng, Nk), grid_y),        \
                               dim3(BIN_THREADS), lds, s, __VA_ARGS__);                       \
            break;                                                                            \
        case 32:                                                                              \
            hipLaunchKernelGGL((k_bin<MODE, 32>), dim3(bin_blocks<32>(Ng, Nk), grid_y),        \
                               dim3(BIN_THREADS), lds, s, __VA_ARGS__);                       \
            break;                                                                            \
        case 128:                                                                             \
            hipLaunchKernelGGL((k_bin<MODE, 128>), dim3(bin_blocks<128>(Ng, Nk), grid_y),      \
                               dim3(BIN_THREADS), lds, s, __VA_ARGS__);                       \
            break;                                                                            \
        default:                                                                              \
            hipLaunchKernelGGL((k_bin<MODE, 64>), dim3(bin_blocks<64>(Ng, Nk), grid_y),        \
                               dim3(BIN_THREADS), lds, s, __VA_ARGS__);                       \
        }                                                                                     \
    } while (0)

static size_t bin_lds(const BinArgs &a) {
    return (size_t)a.wmax * a.nvals * 8 * lds_words(a.fx) + (size_t)a.wmax * a.ncnt * 4 + 16;
}

extern "C" int fgk_bin_pgrid_f64(const double *f, const fgk_binspec *b, double *bins_dev,
                                 int64_t *counts_dev, fgk_stream_t s) {
    fgk_binspec b64 = *b;
    b64.bk = 64;             // MODE 0 and 2 run the default block
    BinArgs a = make_args(&b64, 0);
    hipLaunchKernelGGL((k_bin<0, 64>), dim3(bin_blocks<64>(b->Ng, b->Ng / 2 + 1)), dim3(BIN_THREADS),
                       bin_lds(a), (hipStream_t)s, f, (const float *)nullptr,
                       (const float *)nullptr, a, (const double *)nullptr,
                       (const double *)nullptr, 0.0, 0.0, 0.0, 0.0, bins_dev,
                       (unsigned long long *)counts_dev, (const fgk_binjob *)nullptr);
    FGK_LAUNCHED();
    return 0;
}

// Fixed bin scratch of the allocation-free Pl_bins_noalloc: the FP64 sums
// and int64 counts of fgk_bin_pgrid_f64 for up to FGK_BIN_SCRATCH_BYTES.
// Each device holds its own copy for the life of the process.  Host code
// never passes `scratch` itself to a kernel (on the host that name is a
// stand-in, not a device address): fgk_bin_scratch looks the device address
// up once per device with hipGetSymbolAddress and hands that out.
static __device__ __align__(256) unsigned char scratch[FGK_BIN_SCRATCH_BYTES];

#define FGK_SCRATCH_DEVICES 64
static void *g_scratch_ptr[FGK_SCRATCH_DEVICES];
static std::mutex g_scratch_mu;

extern "C" int fgk_bin_scratch(void **ptr, size_t *bytes) {
    *ptr = nullptr;
    *bytes = 0;
    int dev = -1;
    FGK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= FGK_SCRATCH_DEVICES) {
        fgk_set_error("fgk_bin_scratch: device %d beyond the %d this library tracks", dev,
                      FGK_SCRATCH_DEVICES);
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fgk_set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                          hipGetErrorString(e));
            return -1;
        }
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    *bytes = FGK_BIN_SCRATCH_BYTES;
    return 0;
}

extern "C" int fgk_bin_geometry(const fgk_binspec *b, int64_t *sums_dev, int64_t *counts_dev,
                                fgk_stream_t s) {
    if (!b->geometry) {
        fgk_set_error("fgk_bin_geometry: binspec without geometry");
        return -2;
    }
    fgk_binspec b64 = *b;
    b64.bk = 64;
    BinArgs a = make_args(&b64, 2);
    hipLaunchKernelGGL((k_bin<2, 64>), dim3(bin_blocks<64>(b->Ng, b->Ng / 2 + 1)), dim3(BIN_THREADS),
                       bin_lds(a), (hipStream_t)s, (const double *)nullptr,
                       (const float *)nullptr, (const float *)nullptr, a,
                       (const double *)nullptr, (const double *)nullptr, 0.0, 0.0, 0.0, 0.0,
                       (void *)sums_dev, (unsigned long long *)counts_dev,
                       (const fgk_binjob *)nullptr);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_fused_bin_f32(const float *A, const float *B, const fgk_binspec *b,
                                 const double *phase_dev, const double *winv_dev, double norm,
                                 double Vinv, double Pshot, double npart, int64_t *sums_dev,
                                 fgk_stream_t s) {
    if (b->geometry) {
        fgk_set_error("fgk_fused_bin_f32: geometry is computed by fgk_bin_geometry");
        return -2;
    }
    BinArgs a = make_args(b, 1);
    const int Ng = b->Ng;
    BIN_LAUNCH(1, bin_bk(b), Ng, Ng / 2, 1u, bin_lds(a), (hipStream_t)s, (const double *)nullptr,
               A, B, a, phase_dev, winv_dev, norm, Vinv, Pshot, npart, (void *)sums_dev,
               (unsigned long long *)nullptr, (const fgk_binjob *)nullptr);
    FGK_LAUNCHED();
    const unsigned nplane = (unsigned)(((Ng + PI - 1) / PI) * ((Ng + PJ - 1) / PJ));
    hipLaunchKernelGGL(k_bin_plane, dim3(nplane), dim3(PI * PJ), bin_lds(a), (hipStream_t)s, A, B,
                       a, phase_dev, winv_dev, norm, Vinv, Pshot, npart, (void *)sums_dev,
                       (const fgk_binjob *)nullptr);
    FGK_LAUNCHED();
    return 0;
}

extern "C" void fgk_bin_multi_blocks(const fgk_binspec *b, int Ng, unsigned *blocks,
                                     unsigned *pblocks) {
    unsigned nb = 0;
    switch (bin_bk(b)) {
    case 16: nb = bin_blocks<16>(Ng, Ng / 2); break;
    case 32: nb = bin_blocks<32>(Ng, Ng / 2); break;
    case 128: nb = bin_blocks<128>(Ng, Ng / 2); break;
    default: nb = bin_blocks<64>(Ng, Ng / 2); break;
    }
    if (blocks) *blocks = nb;
    if (pblocks) *pblocks = (unsigned)(((Ng + PI - 1) / PI) * ((Ng + PJ - 1) / PJ));
}

extern "C" int fgk_fused_bin_f32_multi(const fgk_binspec *b, const fgk_binjob *jobs_dev,
                                       int njobs, int NK_max, unsigned nblocks,
                                       unsigned npblocks, fgk_stream_t s) {
    if (b->geometry) {
        fgk_set_error("fgk_fused_bin_f32_multi: geometry is computed by fgk_bin_geometry");
        return -2;
    }
    if (njobs < 1) {
        fgk_set_error("fgk_fused_bin_f32_multi: no catalogs");
        return -2;
    }
    fgk_binspec bb = *b;
    bb.NK = NK_max;          // the LDS window size does not depend on NK
    BinArgs a = make_args(&bb, 1);
    a.multi = njobs;
    const int bk = bin_bk(b);
    switch (bk) {
    case 16:
        hipLaunchKernelGGL((k_bin<1, 16>), dim3(nblocks), dim3(BIN_THREADS), bin_lds(a),
                           (hipStream_t)s, (const double *)nullptr, (const float *)nullptr,
                           (const float *)nullptr, a, (const double *)nullptr,
                           (const double *)nullptr, 0.0, 0.0, 0.0, 0.0, (void *)nullptr,
                           (unsigned long long *)nullptr, jobs_dev);
        break;
    case 32:
        hipLaunchKernelGGL((k_bin<1, 32>), dim3(nblocks), dim3(BIN_THREADS), bin_lds(a),
                           (hipStream_t)s, (const double *)nullptr, (const float *)nullptr,
                           (const float *)nullptr, a, (const double *)nullptr,
                           (const double *)nullptr, 0.0, 0.0, 0.0, 0.0, (void *)nullptr,
                           (unsigned long long *)nullptr, jobs_dev);
        break;
    case 128:
        hipLaunchKernelGGL((k_bin<1, 128>), dim3(nblocks), dim3(BIN_THREADS), bin_lds(a),
                           (hipStream_t)s, (const double *)nullptr, (const float *)nullptr,
                           (const float *)nullptr, a, (const double *)nullptr,
                           (const double *)nullptr, 0.0, 0.0, 0.0, 0.0, (void *)nullptr,
                           (unsigned long long *)nullptr, jobs_dev);
        break;
    default:
        hipLaunchKernelGGL((k_bin<1, 64>), dim3(nblocks), dim3(BIN_THREADS), bin_lds(a),
                           (hipStream_t)s, (const double *)nullptr, (const float *)nullptr,
                           (const float *)nullptr, a, (const double *)nullptr,
                           (const double *)nullptr, 0.0, 0.0, 0.0, 0.0, (void *)nullptr,
                           (unsigned long long *)nullptr, jobs_dev);
    }
    FGK_LAUNCHED();
    hipLaunchKernelGGL(k_bin_plane, dim3(npblocks), dim3(PI * PJ), bin_lds(a), (hipStream_t)s,
                       (const float *)nullptr, (const float *)nullptr, a,
                       (const double *)nullptr, (const double *)nullptr, 0.0, 0.0, 0.0, 0.0,
                       (void *)nullptr, jobs_dev);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_fused_bin_f32_batch(const fgk_binspec *b, const fgk_binjob *jobs_dev,
                                       int njobs, int NK_max, const double *phase_dev,
                                       const double *winv_dev, fgk_stream_t s) {
    if (b->geometry) {
        fgk_set_error("fgk_fused_bin_f32_batch: geometry is computed by fgk_bin_geometry");
        return -2;
    }
    if (njobs < 1 || njobs > 65535) {
        fgk_set_error("fgk_fused_bin_f32_batch: need 1..65535 catalogs");
        return -2;
    }
    fgk_binspec bb = *b;
    bb.NK = NK_max;          // the LDS window size does not depend on NK
    BinArgs a = make_args(&bb, 1);
    const int Ng = b->Ng;
    BIN_LAUNCH(1, bin_bk(b), Ng, Ng / 2, (unsigned)njobs, bin_lds(a), (hipStream_t)s,
               (const double *)nullptr, (const float *)nullptr, (const float *)nullptr, a,
               phase_dev, winv_dev, 0.0, 0.0, 0.0, 0.0, (void *)nullptr,
               (unsigned long long *)nullptr, jobs_dev);
    FGK_LAUNCHED();
    const unsigned nplane = (unsigned)(((Ng + PI - 1) / PI) * ((Ng + PJ - 1) / PJ));
    hipLaunchKernelGGL(k_bin_plane, dim3(nplane, (unsigned)njobs), dim3(PI * PJ), bin_lds(a),
                       (hipStream_t)s, (const float *)nullptr, (const float *)nullptr, a,
                       phase_dev, winv_dev, 0.0, 0.0, 0.0, 0.0, (void *)nullptr, jobs_dev);
    FGK_LAUNCHED();
    return 0;
}
