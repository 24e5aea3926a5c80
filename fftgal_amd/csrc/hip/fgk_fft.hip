// Stockham mixed-radix FFT engine for cubic grids on gfx950.
//
// Replaces the reference's FFTW calls (src/fft.c:28-29 plan r2c/c2r 3d,
// src/fft.c:174,186 fftw_execute) with three LDS-staged passes over the
// padded in-place layout f[(i*Ng+j)*Ng_pad + k] (src/fft.h:10-13,20):
//   rows  (k, contiguous): real <-> half spectrum through a length Ng/2
//          complex transform plus the split post/pre-processing step;
//   cols  (j, stride Ng_pad/2 complex) and (i, stride Ng*(Ng/2+1)):
//          C consecutive k columns per workgroup, loaded as coalesced row
//          segments, transformed in LDS with the columns interleaved so
//          that butterfly reads are bank-conflict free.
// Each workgroup keeps its pencils in LDS for all stages (ping-pong
// buffers), so one pass reads and writes the grid exactly once.
// Twiddles exp(-2 pi i t/Ng) are generated in FP64 on the host (libm) and
// stored in the plan's precision.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "../fgk.h"
#include "fgk_common.h"

#define FGK_MAX_STAGES 24

struct Stages {
    int n;
    int ns;
    int radix[FGK_MAX_STAGES];
};

struct fgk_fftplan {
    int Ng;
    int prec;            // 4 or 8
    Stages sN;           // length Ng
    Stages sH;           // length Ng/2
    void *tw;            // Ng complex, exp(-2 pi i t / Ng)
    int cols_per_wg;     // 0 = automatic
    int rows_per_wg;     // 0 = automatic
    int spec;            // FP32 passes use the specialised kernels
    void *stw_cols;      // stage twiddles, length Ng (specialised FP32 only)
    void *stw_rows;      // stage twiddles, length Ng/2
    int fs_n1;           // i-axis pass as a four-step Ng = n1 x (Ng / n1), 0 = one pass
    void *stw_fs1;       // stage twiddles, lengths fs_n1 and Ng / fs_n1
    void *stw_fs2;
    // FP64 grids of any side (odd, or a prime factor above 13): rows are
    // transformed in pairs as one complex row, every pencil by a complex
    // length-Ng transform (Stockham, or Bluestein through a power of two M)
    int general;
    int blue_M;          // Bluestein length (0: Stockham of length Ng)
    Stages sM;
    void *twM;           // exp(-2 pi i t / M), t < M
    void *chirp;         // exp(-i pi (n^2 mod 2 Ng) / Ng), n < Ng
    void *bf[2];         // FFT_M of the convolution kernel: [0] forward, [1] inverse
};

static int factor_stages(int n, Stages *s) {
    s->n = n;
    s->ns = 0;
    int m = n;
    static const int order[] = {8, 4, 2, 3, 5, 7, 11, 13};
    for (int q = 0; q < 8; ++q) {
        int r = order[q];
        while (m % r == 0) {
            if (s->ns >= FGK_MAX_STAGES) return -1;
            s->radix[s->ns++] = r;
            m /= r;
        }
    }
    return m == 1 ? 0 : -1;
}

extern "C" int fgk_fft_supported(int Ng) {
    if (Ng < 2 || (Ng & 1)) return 0;
    Stages s;
    return factor_stages(Ng, &s) == 0 && factor_stages(Ng / 2, &s) == 0;
}

// FP64 plans take any side up to FGK_GENERAL_MAX (Bluestein lengths up to
// 2048: two LDS buffers of 32 KB per pencil)
#define FGK_GENERAL_MAX 1024
extern "C" int fgk_fft_supported_f64(int Ng) {
    return fgk_fft_supported(Ng) || (Ng >= 1 && Ng <= FGK_GENERAL_MAX);
}

// ------------------------------------------------------------ butterflies

template <typename T>
__device__ __forceinline__ cpx<T> root(const cpx<T> *tw, int e, int sign) {
    cpx<T> w = tw[e];
    if (sign > 0) w.y = -w.y;
    return w;
}

template <typename T, int R> struct Bfly {
    // generic prime radix: direct DFT with roots from the twiddle table
    __device__ static __forceinline__ void run(cpx<T> *v, int sign, const cpx<T> *tw,
                                               int n, int tws) {
        cpx<T> o[R];
        const int rs = (n / R) * tws;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            cpx<T> acc = v[0];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cpx<T> w = root(tw, ((q * r) % R) * rs, sign);
                acc = cadd(acc, cmul(v[r], w));
            }
            o[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = o[q];
    }
};

template <typename T> struct Bfly<T, 2> {
    __device__ static __forceinline__ void run(cpx<T> *v, int, const cpx<T> *, int, int) {
        cpx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename T> struct Bfly<T, 4> {
    __device__ static __forceinline__ void run(cpx<T> *v, int sign, const cpx<T> *, int, int) {
        cpx<T> s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
        cpx<T> s1 = cadd(v[1], v[3]), d1 = csub(v[1], v[3]);
        cpx<T> e = cmul_si(d1, sign);
        v[0] = cadd(s0, s1);
        v[1] = cadd(d0, e);
        v[2] = csub(s0, s1);
        v[3] = csub(d0, e);
    }
};

template <typename T> struct Bfly<T, 3> {
    __device__ static __forceinline__ void run(cpx<T> *v, int sign, const cpx<T> *, int, int) {
        const T h = (T)0.86602540378443864676;
        cpx<T> t1 = cadd(v[1], v[2]);
        cpx<T> t2 = csub(v[0], cscale(t1, (T)0.5));
        cpx<T> t3 = cscale(csub(v[1], v[2]), (T)sign * h);
        v[0] = cadd(v[0], t1);
        v[1] = cmk(t2.x - t3.y, t2.y + t3.x);
        v[2] = cmk(t2.x + t3.y, t2.y - t3.x);
    }
};

template <typename T> struct Bfly<T, 8> {
    __device__ static __forceinline__ void run(cpx<T> *v, int sign, const cpx<T> *, int, int) {
        cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
        Bfly<T, 4>::run(e, sign, nullptr, 0, 0);
        Bfly<T, 4>::run(o, sign, nullptr, 0, 0);
        const T r2 = (T)0.70710678118654752440;
        // w^1 = r2 (1 + sign i), w^2 = sign i, w^3 = r2 (-1 + sign i)
        cpx<T> w1o = cmk(r2 * (o[1].x - sign * o[1].y), r2 * (o[1].y + sign * o[1].x));
        cpx<T> w2o = cmul_si(o[2], sign);
        cpx<T> w3o = cmk(r2 * (-o[3].x - sign * o[3].y), r2 * (-o[3].y + sign * o[3].x));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], w1o);
        v[5] = csub(e[1], w1o);
        v[2] = cadd(e[2], w2o);
        v[6] = csub(e[2], w2o);
        v[3] = cadd(e[3], w3o);
        v[7] = csub(e[3], w3o);
    }
};

// One Stockham stage of radix R over all pencils held in LDS.  Element
// (t, p) of pencil p lives at buf[t*st + p*sp].  pfast selects whether
// consecutive threads take consecutive pencils (column layout) or
// consecutive butterflies (row layout).
template <typename T, int R>
__device__ __forceinline__ void stage_run(const cpx<T> *__restrict__ a, cpx<T> *__restrict__ b,
                                          int n, int Ns, int P, int lgP, int pfast, int st,
                                          int sp, const cpx<T> *__restrict__ tw, int tws,
                                          int sign) {
    const int m = n / R;
    const int tstep = (n / (Ns * R)) * tws;
    const int total = m * P;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        int p, j;
        if (pfast) {
            p = idx & (P - 1);
            j = idx >> lgP;
        } else {
            p = idx / m;
            j = idx - p * m;
        }
        const int k = j % Ns;
        cpx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = a[(j + r * m) * st + p * sp];
        if (k) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], root(tw, r * k * tstep, sign));
        }
        Bfly<T, R>::run(v, sign, tw, n, tws);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) b[(base + r * Ns) * st + p * sp] = v[r];
    }
}

template <typename T>
__device__ cpx<T> *lds_fft(cpx<T> *a, cpx<T> *b, const Stages &S, int P, int lgP, int pfast,
                           int st, int sp, const cpx<T> *tw, int tws, int sign) {
    int Ns = 1;
    for (int s = 0; s < S.ns; ++s) {
        const int R = S.radix[s];
        switch (R) {
        case 8: stage_run<T, 8>(a, b, S.n, Ns, P, lgP, pfast, st, sp, tw, tws, sign); break;
        case 4: stage_run<T, 4>(a, b, S.n, Ns, P, lgP, pfast, st, sp, tw, tws, sign); break;
        case 2: stage_run<T, 2>(a, b, S.n, Ns, P, lgP, pfast, st, sp, tw, tws, sign); break;
        case 3: stage_run<T, 3>(a, b, S.n, Ns, P, lgP, pfast, st, sp, tw, tws, sign); break;
        case 5: stage_run<T, 5>(a, b, S.n, Ns, P, lgP, pfast, st, sp, tw, tws, sign); break;
        case 7: stage_run<T, 7>(a, b, S.n, Ns, P, lgP, pfast, st, sp, tw, tws, sign); break;
        case 11: stage_run<T, 11>(a, b, S.n, Ns, P, lgP, pfast, st, sp, tw, tws, sign); break;
        default: stage_run<T, 13>(a, b, S.n, Ns, P, lgP, pfast, st, sp, tw, tws, sign); break;
        }
        __syncthreads();
        cpx<T> *t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

// ------------------------------------------------------------ row kernels

// MODE 0: real input in place in f; MODE 1: int64 fixed-point input src;
// MODE 2: int64 input, output in the packed half-spectrum layout (Ng/2
// complex per row, column 0 = (X[0], X[Ng/2]), both real)
template <typename T, int MODE>
__global__ void __launch_bounds__(256)
k_rows_r2c(T *f, int64_t *src, double in_scale, int clear_src, T out_scale, int N, int Ng_pad,
           long nrows, Stages S, const cpx<T> *tw, int P) {
    extern __shared__ char smem_raw[];
    cpx<T> *a = (cpx<T> *)smem_raw;
    const int M = N / 2;
    cpx<T> *b = a + (size_t)P * M;
    const long row0 = (long)xcd_block(blockIdx.x, gridDim.x) * P;
    for (int idx = threadIdx.x; idx < P * M; idx += blockDim.x) {
        const int p = idx / M;
        const int m = idx - p * M;
        const long row = row0 + p;
        cpx<T> z = cmk((T)0, (T)0);
        if (row < nrows) {
            if (MODE == 0) {
                const cpx<T> *r = (const cpx<T> *)(f + row * (long)Ng_pad);
                z = r[m];
            } else {
                int64_t *r = src + row * (long)N + 2 * m;
                longlong2 v = *(const longlong2 *)r;
                z = cmk((T)((double)v.x * in_scale), (T)((double)v.y * in_scale));
                if (clear_src) *(longlong2 *)r = make_longlong2(0, 0);
            }
        }
        a[idx] = z;
    }
    __syncthreads();
    cpx<T> *res = lds_fft<T>(a, b, S, P, 0, 0, 1, M, tw, 2, -1);
    const int K1 = MODE == 2 ? M : M + 1;
    for (int idx = threadIdx.x; idx < P * K1; idx += blockDim.x) {
        const int p = idx / K1;
        const int k = idx - p * K1;
        const long row = row0 + p;
        if (row >= nrows) continue;
        if (MODE == 2) {
            cpx<T> X;
            if (k == 0) {
                const cpx<T> Z = res[p * M];
                X = cmk((Z.x + Z.y) * out_scale, (Z.x - Z.y) * out_scale);
            } else {
                cpx<T> Z = res[p * M + k];
                cpx<T> Zc = res[p * M + M - k];
                Zc.y = -Zc.y;
                cpx<T> E = cscale(cadd(Z, Zc), (T)0.5);
                cpx<T> D = csub(Z, Zc);
                cpx<T> O = cmk((T)0.5 * D.y, (T)-0.5 * D.x);
                X = cscale(cadd(E, cmul(O, tw[k])), out_scale);
            }
            ((cpx<T> *)f)[row * (long)M + k] = X;
            continue;
        }
        const int k1 = (k == M) ? 0 : k;
        const int k2 = (k == 0) ? 0 : M - k;
        cpx<T> Z = res[p * M + k1];
        cpx<T> Zc = res[p * M + k2];
        Zc.y = -Zc.y;
        cpx<T> E = cscale(cadd(Z, Zc), (T)0.5);
        cpx<T> D = csub(Z, Zc);
        cpx<T> O = cmk((T)0.5 * D.y, (T)-0.5 * D.x);
        cpx<T> X = cadd(E, cmul(O, tw[k]));
        X = cscale(X, out_scale);
        ((cpx<T> *)(f + row * (long)Ng_pad))[k] = X;
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_rows_c2r(T *f, T out_scale, int N, int Ng_pad, long nrows, Stages S, const cpx<T> *tw, int P) {
    extern __shared__ char smem_raw[];
    const int M = N / 2;
    const int K1 = M + 1;
    cpx<T> *xs = (cpx<T> *)smem_raw;          // P*(M+1) staged half spectra
    cpx<T> *a = xs + (size_t)P * K1;
    cpx<T> *b = a + (size_t)P * M;
    const long row0 = (long)xcd_block(blockIdx.x, gridDim.x) * P;
    for (int idx = threadIdx.x; idx < P * K1; idx += blockDim.x) {
        const int p = idx / K1;
        const int k = idx - p * K1;
        const long row = row0 + p;
        xs[idx] = row < nrows ? ((const cpx<T> *)(f + row * (long)Ng_pad))[k] : cmk((T)0, (T)0);
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < P * M; idx += blockDim.x) {
        const int p = idx / M;
        const int k = idx - p * M;
        cpx<T> X = xs[p * K1 + k];
        cpx<T> Xc = xs[p * K1 + (M - k)];
        Xc.y = -Xc.y;
        cpx<T> E = cadd(X, Xc);
        cpx<T> O = cmulc(csub(X, Xc), tw[k]);   // times exp(+2 pi i k/N)
        a[idx] = cmk(E.x - O.y, E.y + O.x);      // E + i O
    }
    __syncthreads();
    cpx<T> *res = lds_fft<T>(a, b, S, P, 0, 0, 1, M, tw, 2, +1);
    for (int idx = threadIdx.x; idx < P * M; idx += blockDim.x) {
        const int p = idx / M;
        const int m = idx - p * M;
        const long row = row0 + p;
        if (row >= nrows) continue;
        ((cpx<T> *)(f + row * (long)Ng_pad))[m] = cscale(res[idx], out_scale);
    }
}

// --------------------------------------------------------- column kernel

template <typename T>
__global__ void __launch_bounds__(256)
k_cols(cpx<T> *g, int N, int Nk, long stride_t, long stride_o, int C, int lgC, Stages S,
       const cpx<T> *tw, int sign, T scale) {
    extern __shared__ char smem_raw[];
    cpx<T> *a = (cpx<T> *)smem_raw;
    cpx<T> *b = a + (size_t)N * C;
    const int nkb = (Nk + C - 1) / C;
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int o = lb / nkb;
    const int k0 = (lb - o * nkb) * C;
    cpx<T> *base = g + (long)o * stride_o + k0;
    const int kc = Nk - k0 < C ? Nk - k0 : C;
    for (int idx = threadIdx.x; idx < N * C; idx += blockDim.x) {
        const int c = idx & (C - 1);
        const int t = idx >> lgC;
        a[idx] = c < kc ? base[(long)t * stride_t + c] : cmk((T)0, (T)0);
    }
    __syncthreads();
    cpx<T> *res = lds_fft<T>(a, b, S, C, lgC, 1, C, 1, tw, 1, sign);
    for (int idx = threadIdx.x; idx < N * C; idx += blockDim.x) {
        const int c = idx & (C - 1);
        const int t = idx >> lgC;
        if (c < kc) base[(long)t * stride_t + c] = cscale(res[idx], scale);
    }
}

// ---------------------------------------------- FP64 grids of any side
//
// A complex transform of length N on P pencils in LDS (element t of pencil p
// at a[t*st + p*sp], buffers of L = N, or M for Bluestein, elements per
// pencil).  N whose prime factors are at most 13 runs the Stockham stages;
// otherwise Bluestein's chirp-z form (Bluestein 1970): with
// w[n] = exp(s i pi n^2 / N) (s = the transform's sign),
//   X[k] = w[k] sum_n (x[n] w[n]) conj(w[k - n]),
// a circular convolution of length M >= 2N - 1 (a power of two) done as
// FFT_M, a product with the precomputed FFT_M of the kernel conj(w[m])
// (m in (-N, N), wrapped), and an inverse FFT_M scaled by 1/M.
struct GenFFT {
    int N, M;                    // M: Bluestein length, 0 for Stockham
    Stages sN, sM;
    const cpx<double> *tw, *twM, *chirp;
    const cpx<double> *bf[2];    // [0] sign -1, [1] sign +1
};

__device__ __forceinline__ cpx<double> gen_chirp(const GenFFT &G, int t, int sign) {
    cpx<double> w = G.chirp[t];
    if (sign > 0) w.y = -w.y;
    return w;
}

__device__ cpx<double> *gen_cfft(cpx<double> *a, cpx<double> *b, const GenFFT &G, int P, int lgP,
                                 int pfast, int st, int sp, int sign) {
    if (!G.M) return lds_fft<double>(a, b, G.sN, P, lgP, pfast, st, sp, G.tw, 1, sign);
    const int N = G.N, M = G.M;
    for (int idx = threadIdx.x; idx < M * P; idx += blockDim.x) {
        int p, t;
        if (pfast) { p = idx & (P - 1); t = idx >> lgP; } else { p = idx / M; t = idx - p * M; }
        cpx<double> &v = a[t * st + p * sp];
        v = t < N ? cmul(v, gen_chirp(G, t, sign)) : cmk(0.0, 0.0);
    }
    __syncthreads();
    cpx<double> *r = lds_fft<double>(a, b, G.sM, P, lgP, pfast, st, sp, G.twM, 1, -1);
    cpx<double> *o = r == a ? b : a;
    const cpx<double> *bf = G.bf[sign > 0 ? 1 : 0];
    for (int idx = threadIdx.x; idx < M * P; idx += blockDim.x) {
        int p, t;
        if (pfast) { p = idx & (P - 1); t = idx >> lgP; } else { p = idx / M; t = idx - p * M; }
        cpx<double> &v = r[t * st + p * sp];
        v = cmul(v, bf[t]);
    }
    __syncthreads();
    r = lds_fft<double>(r, o, G.sM, P, lgP, pfast, st, sp, G.twM, 1, +1);
    const double inv = 1.0 / (double)M;
    for (int idx = threadIdx.x; idx < N * P; idx += blockDim.x) {
        int p, t;
        if (pfast) { p = idx & (P - 1); t = idx >> lgP; } else { p = idx / N; t = idx - p * N; }
        cpx<double> &v = r[t * st + p * sp];
        v = cscale(cmul(v, gen_chirp(G, t, sign)), inv);
    }
    __syncthreads();
    return r;
}

// rows of the padded layout (Ng_pad doubles each) in pairs (2q, 2q+1) as
// z = x_{2q} + i x_{2q+1}: r2c (dir 0) splits Z into the two half spectra,
// X1[k] = (Z[k] + conj Z[-k]) / 2, X2[k] = (Z[k] - conj Z[-k]) / 2i,
// k <= N/2; c2r (dir 1) forms Z[k] = X1[k] + i X2[k] from the Hermitian
// extensions, transforms with sign +1 and writes Re, Im to the two rows.
// P row pairs per workgroup, L = N or M complex of LDS per pair and buffer.
__global__ void __launch_bounds__(256)
k_rows_gen(double *f, int Ng_pad, long nrows, GenFFT G, int P, int dir, double out_scale) {
    extern __shared__ char smem_raw[];
    const int N = G.N, L = G.M ? G.M : N, K1 = N / 2 + 1;
    cpx<double> *a = (cpx<double> *)smem_raw;
    cpx<double> *b = a + (size_t)P * L;
    const long q0 = (long)xcd_block(blockIdx.x, gridDim.x) * P;
    auto rowp = [&](long r) -> double * { return r < nrows ? f + r * (long)Ng_pad : nullptr; };
    for (int idx = threadIdx.x; idx < P * N; idx += blockDim.x) {
        const int p = idx / N, t = idx - p * N;
        const double *r0 = rowp(2 * (q0 + p)), *r1 = rowp(2 * (q0 + p) + 1);
        cpx<double> z = cmk(0.0, 0.0);
        if (dir == 0) {
            z = cmk(r0 ? r0[t] : 0.0, r1 ? r1[t] : 0.0);
        } else {
            const int k = t <= N / 2 ? t : N - t;
            cpx<double> x1 = r0 ? ((const cpx<double> *)r0)[k] : cmk(0.0, 0.0);
            cpx<double> x2 = r1 ? ((const cpx<double> *)r1)[k] : cmk(0.0, 0.0);
            if (t > N / 2) { x1.y = -x1.y; x2.y = -x2.y; }
            z = cmk(x1.x - x2.y, x1.y + x2.x);   // X1 + i X2
        }
        a[p * L + t] = z;
    }
    __syncthreads();
    cpx<double> *res = gen_cfft(a, b, G, P, 0, 0, 1, L, dir == 0 ? -1 : +1);
    if (dir == 0) {
        for (int idx = threadIdx.x; idx < P * K1; idx += blockDim.x) {
            const int p = idx / K1, k = idx - p * K1;
            double *r0 = rowp(2 * (q0 + p)), *r1 = rowp(2 * (q0 + p) + 1);
            const cpx<double> Z = res[p * L + k];
            cpx<double> Zc = res[p * L + (k ? N - k : 0)];
            Zc.y = -Zc.y;
            const cpx<double> E = cscale(cadd(Z, Zc), 0.5);
            const cpx<double> D = csub(Z, Zc);
            if (r0) ((cpx<double> *)r0)[k] = cscale(E, out_scale);
            if (r1) ((cpx<double> *)r1)[k] = cscale(cmk(0.5 * D.y, -0.5 * D.x), out_scale);
        }
    } else {
        for (int idx = threadIdx.x; idx < P * N; idx += blockDim.x) {
            const int p = idx / N, t = idx - p * N;
            double *r0 = rowp(2 * (q0 + p)), *r1 = rowp(2 * (q0 + p) + 1);
            const cpx<double> z = res[p * L + t];
            if (r0) r0[t] = z.x * out_scale;
            if (r1) r1[t] = z.y * out_scale;
        }
    }
}

// columns: C consecutive columns of one plane per workgroup (as k_cols)
__global__ void __launch_bounds__(256)
k_cols_gen(cpx<double> *g, int Nk, long stride_t, long stride_o, int C, int lgC, GenFFT G,
           int sign, double scale) {
    extern __shared__ char smem_raw[];
    const int N = G.N, L = G.M ? G.M : N;
    cpx<double> *a = (cpx<double> *)smem_raw;
    cpx<double> *b = a + (size_t)L * C;
    const int nkb = (Nk + C - 1) / C;
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int o = lb / nkb;
    const int k0 = (lb - o * nkb) * C;
    cpx<double> *base = g + (long)o * stride_o + k0;
    const int kc = Nk - k0 < C ? Nk - k0 : C;
    for (int idx = threadIdx.x; idx < N * C; idx += blockDim.x) {
        const int c = idx & (C - 1);
        const int t = idx >> lgC;
        a[idx] = c < kc ? base[(long)t * stride_t + c] : cmk(0.0, 0.0);
    }
    __syncthreads();
    cpx<double> *res = gen_cfft(a, b, G, C, lgC, 1, C, 1, sign);
    for (int idx = threadIdx.x; idx < N * C; idx += blockDim.x) {
        const int c = idx & (C - 1);
        const int t = idx >> lgC;
        if (c < kc) base[(long)t * stride_t + c] = cscale(res[idx], scale);
    }
}

// ------------------------------------- specialised FP32 kernels (fast path)
//
// For the pencil lengths of the common grid sizes the whole transform is
// unrolled at compile time: the radix sequence, every stage's span and the
// butterflies per thread are constants, so index arithmetic folds away.
// Each thread owns whole butterflies; a stage loads its R inputs per
// butterfly into registers (from global memory in the first stage), applies
// the stage twiddles (a per-stage table [k][r-1], FP64-built), runs the
// radix-R butterfly and stores to LDS (to global memory in the last stage of
// a column pass).  The layout is the packed half spectrum: Ng/2 complex
// columns per row, so rows are 2^k-byte aligned and a 16-column workgroup
// moves whole 128-byte lines.

struct RSeq {
    int n;
    int r[16];
};

// radix-8 stages, one or two radix-4 stages for the remaining power of two,
// then the odd prime factors
constexpr RSeq radix_seq(int L) {
    RSeq s{0, {0}};
    int m = L, a = 0;
    while (m % 2 == 0) { m /= 2; ++a; }
    int n8 = a / 3;
    const int rem = a % 3;
    if (rem == 1 && n8 > 0) --n8;
    for (int i = 0; i < n8; ++i) s.r[s.n++] = 8;
    if (rem == 2) s.r[s.n++] = 4;
    if (rem == 1) {
        if (a >= 4) { s.r[s.n++] = 4; s.r[s.n++] = 4; }
        else s.r[s.n++] = 2;
    }
    const int odd[5] = {3, 5, 7, 11, 13};
    for (int q = 0; q < 5; ++q)
        while (m % odd[q] == 0) { s.r[s.n++] = odd[q]; m /= odd[q]; }
    return s;
}

constexpr int seq_rmax(int L) {
    RSeq s = radix_seq(L);
    int r = 1;
    for (int i = 0; i < s.n; ++i) r = s.r[i] > r ? s.r[i] : r;
    return r;
}

// offset of stage S's twiddles in the concatenated table: stage s holds
// Ns_s * (R_s - 1) entries
constexpr int stw_offset(int L, int S) {
    RSeq s = radix_seq(L);
    int off = 0, ns = 1;
    for (int i = 0; i < S; ++i) {
        off += ns * (s.r[i] - 1);
        ns *= s.r[i];
    }
    return off;
}

constexpr int stw_size(int L) { return stw_offset(L, radix_seq(L).n); }

__device__ __forceinline__ cpx<float> twmul(cpx<float> v, cpx<float> w, int sign) {
    return sign < 0 ? cmul(v, w) : cmulc(v, w);
}

// stage S (span NS) of a length-L transform; TPP threads per pencil, q = this
// thread's index within its pencil.  in(t): stage input (global in stage 0,
// LDS after), ld/st: LDS, gout: global output of the last stage when GOUT.
// act: this thread takes part (a multi-size launch runs more threads than
// the pencils of a small transform need; the others only meet the barriers).
template <int L, int TPP, int S, int NS, bool GOUT, class In, class Ld, class St, class GOut>
__device__ __forceinline__ void spencil(int q, In in, Ld ld, St st, GOut gout,
                                        const cpx<float> *__restrict__ stw,
                                        const cpx<float> *__restrict__ tw, int tws, int sign,
                                        bool act = true) {
    constexpr RSeq seq = radix_seq(L);
    constexpr int R = seq.r[S];
    constexpr bool last = S + 1 == seq.n;
    constexpr int NB = L / R;
    constexpr int B = (NB + TPP - 1) / TPP;
    cpx<float> v[B][R];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = q + b * TPP;
        if (act && (NB % TPP == 0 || j < NB)) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = S == 0 ? in(j + r * NB) : ld(j + r * NB);
        }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = q + b * TPP;
        if (act && (NB % TPP == 0 || j < NB)) {
            if (NS > 1) {
                const int k = j % NS;
                const cpx<float> *w = stw + stw_offset(L, S) + k * (R - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = twmul(v[b][r], w[r - 1], sign);
            }
            Bfly<float, R>::run(v[b], sign, tw, L, tws);
        }
    }
    if (last && GOUT) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = q + b * TPP;
            if (act && (NB % TPP == 0 || j < NB)) {
#pragma unroll
                for (int r = 0; r < R; ++r) gout(j + r * NB, v[b][r]);
            }
        }
        return;
    }
    if (S > 0) __syncthreads();          // every thread has read this stage's inputs
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = q + b * TPP;
        if (act && (NB % TPP == 0 || j < NB)) {
            const int k = j % NS;
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) st(base + r * NS, v[b][r]);
        }
    }
    __syncthreads();
    if constexpr (!last)
        spencil<L, TPP, S + 1, NS * R, GOUT>(q, in, ld, st, gout, stw, tw, tws, sign, act);
}

// column pass: C consecutive columns of one row/plane per workgroup, the
// pencil index t at stride stride_t; LDS row stride C+1 complex.  With tw4,
// output t of the pencils of outer group o is multiplied by tw4[o t] (the
// twiddle between the two passes of a four-step transform).
// (the body of k_cols32 and of the multi-size column pass: lb = the
// workgroup's logical block, buf = its L (C + 1) complex of LDS)
template <int L, int C, int TPP>
__device__ __forceinline__ void cols32_body(unsigned lb, cpx<float> *buf, cpx<float> *g,
                                            int ncols, long stride_t, long stride_o,
                                            const cpx<float> *stw, const cpx<float> *tw, int tws,
                                            int sign, const cpx<float> *tw4) {
    constexpr int LS = C + 1;
    const int c = threadIdx.x % C, q = threadIdx.x / C;
    const bool act = q < TPP;
    const int ncb = (ncols + C - 1) / C;
    const int o = (int)(lb / (unsigned)ncb);
    const int c0 = (int)(lb - (unsigned)o * ncb) * C;
    const bool live = act && c0 + c < ncols;
    cpx<float> *base = g + (long)o * stride_o + c0 + c;
    auto gin = [&](int t) -> cpx<float> {
        return live ? base[(long)t * stride_t] : cmk(0.0f, 0.0f);
    };
    auto gout = [&](int t, cpx<float> v) {
        if (tw4) v = twmul(v, tw4[o * t], sign);
        if (live) base[(long)t * stride_t] = v;
    };
    auto ld = [&](int t) -> cpx<float> { return buf[t * LS + c]; };
    auto st = [&](int t, cpx<float> v) { buf[t * LS + c] = v; };
    spencil<L, TPP, 0, 1, true>(q, gin, ld, st, gout, stw, tw, tws, sign, act);
}

template <int L, int C, int TPP>
__global__ void __launch_bounds__(C * TPP)
k_cols32(cpx<float> *g, int ncols, long stride_t, long stride_o, const cpx<float> *stw,
         const cpx<float> *tw, int tws, int sign, const cpx<float> *tw4) {
    __shared__ cpx<float> buf[L * (C + 1)];
    cols32_body<L, C, TPP>(xcd_block(blockIdx.x, gridDim.x), buf, g, ncols, stride_t, stride_o,
                           stw, tw, tws, sign, tw4);
}

// row pass: P rows of 2M int64 fixed-point values per workgroup, complex
// length-M transform of z[t] = x[2t] + i x[2t+1], then the real-input split
// X[k] = A[k] + W^k B[k], X[M-k] = conj(A[k] - W^k B[k]) with
// A = (Z[k] + conj Z[M-k]) / 2, B = (Z[k] - conj Z[M-k]) / 2i, W = e^{-2 pi i/2M},
// written to the packed layout (column 0 = (X[0], X[M])).  LDS holds each row
// with one pad slot per 8 elements.
// (the body of k_rows32 and of the multi-size row pass: lb = the
// workgroup's logical block, buf = its P (M + M/8 + 1) complex of LDS)
template <int M, int P, int TPP>
__device__ __forceinline__ void rows32_body(unsigned lb, cpx<float> *buf, cpx<float> *out,
                                            int64_t *src, double in_scale, int clear_src,
                                            long nrows, const cpx<float> *stw,
                                            const cpx<float> *tw) {
    constexpr int RS = M + M / 8 + 1;
    const int q = threadIdx.x % TPP, p = threadIdx.x / TPP;
    const bool act = p < P;
    const long row = (long)lb * P + p;
    const bool live = act && row < nrows;
    cpx<float> *mine = buf + (act ? p : 0) * RS;
    longlong2 *srow = (longlong2 *)(src + row * (2L * M));
    auto gin = [&](int t) -> cpx<float> {
        if (!live) return cmk(0.0f, 0.0f);
        const longlong2 v = srow[t];
        if (clear_src) srow[t] = make_longlong2(0, 0);
        return cmk((float)((double)v.x * in_scale), (float)((double)v.y * in_scale));
    };
    auto ld = [&](int t) -> cpx<float> { return mine[t + (t >> 3)]; };
    auto st = [&](int t, cpx<float> v) { mine[t + (t >> 3)] = v; };
    auto none = [&](int, cpx<float>) {};
    spencil<M, TPP, 0, 1, false>(q, gin, ld, st, none, stw, tw, 2, -1, act);
    if (!live) return;
    cpx<float> *orow = out + row * (long)M;
    for (int k = q; k <= M / 2; k += TPP) {
        const cpx<float> Z = ld(k);
        if (k == 0) {
            orow[0] = cmk(Z.x + Z.y, Z.x - Z.y);
            continue;
        }
        cpx<float> Zc = ld(M - k);
        Zc.y = -Zc.y;
        const cpx<float> A = cscale(cadd(Z, Zc), 0.5f);
        const cpx<float> D = csub(Z, Zc);
        const cpx<float> Bk = cmk(0.5f * D.y, -0.5f * D.x);
        const cpx<float> Tk = cmul(Bk, tw[k]);
        orow[k] = cadd(A, Tk);
        const cpx<float> Dm = csub(A, Tk);
        orow[M - k] = cmk(Dm.x, -Dm.y);
    }
}

template <int M, int P, int TPP>
__global__ void __launch_bounds__(P * TPP)
k_rows32(cpx<float> *out, int64_t *src, double in_scale, int clear_src, long nrows,
         const cpx<float> *stw, const cpx<float> *tw) {
    __shared__ cpx<float> buf[P * (M + M / 8 + 1)];
    rows32_body<M, P, TPP>(xcd_block(blockIdx.x, gridDim.x), buf, out, src, in_scale, clear_src,
                           nrows, stw, tw);
}

// launch shapes: columns per workgroup C (<= 16) and rows per workgroup P
// with at most 1024 threads and about 80 KB of LDS per workgroup
constexpr int cols_C(int L) {
    int C = 16;
    while (C > 1 && (C * (L / seq_rmax(L)) > 1024 || L * (C + 1) * 8 > 80 * 1024)) C /= 2;
    return C;
}
constexpr int cols_TPP(int L) {
    int t = L / seq_rmax(L);
    while (cols_C(L) * t > 1024) t /= 2;
    return t;
}
constexpr int rows_TPP(int M) {
    int t = M / seq_rmax(M);
    return t > 1024 ? 1024 : t;
}
constexpr int rows_P(int M) {
    int P = 1024 / rows_TPP(M);
    while (P > 1 && P * (M + M / 8 + 1) * 8 > 80 * 1024) P /= 2;
    return P;
}

// grid sides with specialised kernels (columns of length Ng, rows of Ng/2)
#define FGK_SPECIALISED_NG(X) \
    X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(2048)

extern "C" int fgk_fft_f32_specialised(int Ng) {
    switch (Ng) {
#define CASE_NG(n) case n:
        FGK_SPECIALISED_NG(CASE_NG)
#undef CASE_NG
        return 1;
    default:
        return 0;
    }
}

// nouter groups of ncols contiguous columns, group o at g + o * stride_o,
// pencil element t at stride stride_t
static int launch_cols32(int N, cpx<float> *g, int ncols, long stride_t, long stride_o,
                         long nouter, const cpx<float> *stw, const cpx<float> *tw, int sign,
                         hipStream_t s, const cpx<float> *tw4 = nullptr, int tws = 1) {
    switch (N) {
#define CASE_NG(n)                                                                            \
    case n: {                                                                                 \
        constexpr int C = cols_C(n), TPP = cols_TPP(n);                                       \
        const long nb = nouter * ((ncols + C - 1) / C);                                       \
        hipLaunchKernelGGL((k_cols32<n, C, TPP>), dim3((unsigned)nb), dim3(C * TPP), 0, s, g, \
                           ncols, stride_t, stride_o, stw, tw, tws, sign, tw4);               \
        break;                                                                                \
    }
        FGK_SPECIALISED_NG(CASE_NG)
#undef CASE_NG
    default:
        fgk_set_error("no specialised FFT kernel for Ng=%d", N);
        return -2;
    }
    FGK_LAUNCHED();
    return 0;
}

// the passes of a four-step transform are short (n <= 64): 64 columns per
// workgroup, so that each row piece is 512 bytes
#define FGK_FOURSTEP_L(X) X(16) X(24) X(32) X(48) X(64)
static int launch_cols32_wide(int N, cpx<float> *g, int ncols, long stride_t, long stride_o,
                              long nouter, const cpx<float> *stw, const cpx<float> *tw,
                              int sign, hipStream_t s, const cpx<float> *tw4, int tws) {
    switch (N) {
#define CASE_NG(n)                                                                            \
    case n: {                                                                                 \
        constexpr int C = 64, TPP = n / seq_rmax(n);                                          \
        static_assert(C * TPP <= 1024 && n * (C + 1) * 8 <= 80 * 1024, "launch shape");       \
        const long nb = nouter * ((ncols + C - 1) / C);                                       \
        hipLaunchKernelGGL((k_cols32<n, C, TPP>), dim3((unsigned)nb), dim3(C * TPP), 0, s, g, \
                           ncols, stride_t, stride_o, stw, tw, tws, sign, tw4);               \
        break;                                                                                \
    }
        FGK_FOURSTEP_L(CASE_NG)
#undef CASE_NG
    default:
        return launch_cols32(N, g, ncols, stride_t, stride_o, nouter, stw, tw, sign, s, tw4, tws);
    }
    FGK_LAUNCHED();
    return 0;
}

static int launch_rows32(int N, long nrows, cpx<float> *out, int64_t *src, double in_scale,
                         int clear_src, const cpx<float> *stw, const cpx<float> *tw,
                         hipStream_t s) {
    switch (N) {
#define CASE_NG(n)                                                                            \
    case n: {                                                                                 \
        constexpr int M = n / 2, P = rows_P(M), TPP = rows_TPP(M);                            \
        hipLaunchKernelGGL((k_rows32<M, P, TPP>), dim3((unsigned)((nrows + P - 1) / P)),      \
                           dim3(P * TPP), 0, s, out, src, in_scale, clear_src, nrows, stw, tw); \
        break;                                                                                \
    }
        FGK_SPECIALISED_NG(CASE_NG)
#undef CASE_NG
    default:
        fgk_set_error("no specialised FFT kernel for Ng=%d", N);
        return -2;
    }
    FGK_LAUNCHED();
    return 0;
}

// ---------------------------------------------- multi-size passes (batches)
//
// One launch per FFT pass over groups of grids of different sides: group g
// holds nbatch grids of side N one after the other.  Each workgroup finds
// its group by the groups' first logical blocks (binary search, scalar
// loads) and runs the compile-time specialised body of that side; the
// launch has 1024 threads and the largest LDS of the sides it serves, and
// threads a side does not need only meet the barriers.
#define FGK_MULTI_NG(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

// columns per workgroup of a side in the multi-size column pass: the most
// (a power of two) that 1024 threads and 80 KB of LDS hold, so that small
// sides fill the launch's workgroups too
constexpr int mcols_C(int L) {
    int C = 1;
    while (2 * C * cols_TPP(L) <= 1024 && L * (2 * C + 1) * 8 <= 80 * 1024 && 2 * C <= 64) C *= 2;
    return C;
}
constexpr int multi_rows_lds(int N) {
    return rows_P(N / 2) * (N / 2 + N / 16 + 1) * 8;
}
constexpr int multi_cols_lds(int N) { return N * (mcols_C(N) + 1) * 8; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }
#define MAX_ROWS_LDS(n) cmax(multi_rows_lds(n),
#define MAX_COLS_LDS(n) cmax(multi_cols_lds(n),
constexpr int MULTI_ROWS_LDS =
    FGK_MULTI_NG(MAX_ROWS_LDS) 0))))))))) ;
constexpr int MULTI_COLS_LDS =
    FGK_MULTI_NG(MAX_COLS_LDS) 0))))))))) ;
#undef MAX_ROWS_LDS
#undef MAX_COLS_LDS
static_assert(MULTI_ROWS_LDS <= 80 * 1024 && MULTI_COLS_LDS <= 80 * 1024, "multi LDS");

struct MultiGroup {
    cpx<float> *f;               // spectra of the group's grids
    int64_t *src[2];             // int64 grids, grid set 0 and 1 (row pass)
    int N, nbatch;
    unsigned rblk0, cblk0[2];    // first logical block: row pass, column pass axis 0 / 1
    const cpx<float> *stw_rows, *stw_cols, *tw;
    int ncols[2];
    long stride_t[2], stride_o[2];
};

__device__ __forceinline__ int multi_group(const MultiGroup *__restrict__ G, int ng, unsigned lb,
                                           int which) {
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const unsigned b0 = which < 0 ? G[mid].rblk0 : G[mid].cblk0[which];
        if (b0 <= lb) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(1024)
k_rows32_multi(const MultiGroup *__restrict__ G, int ng, int set, int clear_src) {
    __shared__ cpx<float> buf[MULTI_ROWS_LDS / 8];
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int gi = multi_group(G, ng, lb, -1);
    const MultiGroup &g = G[gi];
    const long nrows = (long)g.N * g.N * g.nbatch;
    switch (g.N) {
#define CASE_NG(n)                                                                            \
    case n:                                                                                   \
        rows32_body<n / 2, rows_P(n / 2), rows_TPP(n / 2)>(lb - g.rblk0, buf, g.f, g.src[set],  \
                                                           1.0, clear_src, nrows, g.stw_rows, \
                                                           g.tw);                             \
        break;
        FGK_MULTI_NG(CASE_NG)
#undef CASE_NG
    default:
        break;
    }
}

__global__ void __launch_bounds__(1024)
k_cols32_multi(const MultiGroup *__restrict__ G, int ng, int axis, int sign) {
    __shared__ cpx<float> buf[MULTI_COLS_LDS / 8];
    const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
    const int gi = multi_group(G, ng, lb, axis);
    const MultiGroup &g = G[gi];
    switch (g.N) {
#define CASE_NG(n)                                                                            \
    case n:                                                                                   \
        cols32_body<n, mcols_C(n), cols_TPP(n)>(lb - g.cblk0[axis], buf, g.f, g.ncols[axis],    \
                                               g.stride_t[axis], g.stride_o[axis], g.stw_cols,\
                                               g.tw, 1, sign, nullptr);                       \
        break;
        FGK_MULTI_NG(CASE_NG)
#undef CASE_NG
    default:
        break;
    }
}

struct fgk_fftmulti {
    int ng;
    MultiGroup *d;               // device table
    unsigned rblocks, cblocks[2];
};

// concatenated per-stage twiddles of a length-L transform (FP64-built):
// stage s, k < Ns_s, r = 1..R_s-1 -> exp(-2 pi i r k / (Ns_s R_s))
static int stage_twiddles(int L, void **dev) {
    *dev = NULL;
    const RSeq seq = radix_seq(L);
    const int n = stw_size(L);
    float *h = (float *)malloc((size_t)(n > 0 ? n : 1) * 2 * sizeof(float));
    if (!h) { fgk_set_error("out of host memory"); return -1; }
    int off = 0, ns = 1;
    for (int st = 0; st < seq.n; ++st) {
        const int R = seq.r[st];
        for (int k = 0; k < ns; ++k)
            for (int r = 1; r < R; ++r) {
                const double ang = -2.0 * M_PI * (double)r * k / ((double)ns * R);
                h[2 * off] = (float)cos(ang);
                h[2 * off + 1] = (float)sin(ang);
                ++off;
            }
        ns *= R;
    }
    int rc = fgk_malloc(dev, (size_t)(n > 0 ? n : 1) * 2 * sizeof(float));
    if (!rc) {
        hipError_t e = hipMemcpy(*dev, h, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fgk_hip_fail(e, "hipMemcpy stage twiddles", __FILE__, __LINE__);
    }
    free(h);
    return rc;
}

// ------------------------------------------------------------------ host

// FP64 tables of a general plan: Bluestein length, twiddles of M, the chirp
// and the kernel's FFT for both signs (a direct FP64 DFT on the host, once)
static int general_setup(fgk_fftplan *p) {
    const int N = p->Ng;
    p->general = 1;
    if (factor_stages(N, &p->sN) == 0) return 0;   // Stockham of length N
    int M = 1;
    while (M < 2 * N - 1) M <<= 1;
    p->blue_M = M;
    factor_stages(M, &p->sM);
    double *twM = (double *)malloc((size_t)M * 2 * sizeof(double));
    double *chirp = (double *)malloc((size_t)N * 2 * sizeof(double));
    double *ker = (double *)malloc((size_t)M * 2 * sizeof(double));
    double *bf = (double *)malloc((size_t)M * 2 * sizeof(double));
    int rc = (!twM || !chirp || !ker || !bf) ? -1 : 0;
    if (rc) fgk_set_error("out of host memory");
    if (!rc) {
        for (int t = 0; t < M; ++t) {
            const double ang = 2.0 * M_PI * (double)t / (double)M;
            twM[2 * t] = cos(ang);
            twM[2 * t + 1] = -sin(ang);
        }
        for (int n = 0; n < N; ++n) {   // exp(-i pi (n^2 mod 2N) / N), exact argument
            const long e = ((long)n * n) % (2L * N);
            const double ang = M_PI * (double)e / (double)N;
            chirp[2 * n] = cos(ang);
            chirp[2 * n + 1] = -sin(ang);
        }
    }
    for (int sg = 0; sg < 2 && !rc; ++sg) {
        // kernel conj(w_s[m]) for m in (-N, N) wrapped into M, w_s = chirp
        // (sign -1) or its conjugate (sign +1)
        for (int m = 0; m < M; ++m) ker[2 * m] = ker[2 * m + 1] = 0.0;
        for (int m = 0; m < N; ++m) {
            const double re = chirp[2 * m], im = sg == 0 ? -chirp[2 * m + 1] : chirp[2 * m + 1];
            ker[2 * m] = re;
            ker[2 * m + 1] = im;
            if (m) {
                ker[2 * (M - m)] = re;
                ker[2 * (M - m) + 1] = im;
            }
        }
        for (int k = 0; k < M; ++k) {   // FFT_M, sign -1
            double sr = 0.0, si = 0.0;
            for (int m = 0; m < M; ++m) {
                const long e = ((long)k * m) % M;
                const double c = twM[2 * e], sn = twM[2 * e + 1];
                sr += ker[2 * m] * c - ker[2 * m + 1] * sn;
                si += ker[2 * m] * sn + ker[2 * m + 1] * c;
            }
            bf[2 * k] = sr;
            bf[2 * k + 1] = si;
        }
        if (!rc) rc = fgk_malloc(&p->bf[sg], (size_t)M * 2 * sizeof(double));
        if (!rc) rc = fgk_h2d(p->bf[sg], bf, (size_t)M * 2 * sizeof(double), NULL);
    }
    if (!rc) rc = fgk_malloc(&p->twM, (size_t)M * 2 * sizeof(double));
    if (!rc) rc = fgk_h2d(p->twM, twM, (size_t)M * 2 * sizeof(double), NULL);
    if (!rc) rc = fgk_malloc(&p->chirp, (size_t)N * 2 * sizeof(double));
    if (!rc) rc = fgk_h2d(p->chirp, chirp, (size_t)N * 2 * sizeof(double), NULL);
    if (!rc) rc = fgk_stream_sync(NULL);
    free(twM);
    free(chirp);
    free(ker);
    free(bf);
    return rc;
}

extern "C" int fgk_fft_plan_create(fgk_fftplan **out, int Ng, int precision) {
    *out = NULL;
    if (precision != 4 && precision != 8) {
        fgk_set_error("precision must be 4 or 8");
        return -2;
    }
    const int general = !fgk_fft_supported(Ng);
    if (general && (precision != 8 || !fgk_fft_supported_f64(Ng))) {
        fgk_set_error("grid size Ng=%d unsupported: need even Ng = 2^a 3^b 5^c 7^d 11^e 13^f "
                      "(FP64 plans: any Ng in [1, %d])", Ng, FGK_GENERAL_MAX);
        return -2;
    }
    fgk_fftplan *p = (fgk_fftplan *)calloc(1, sizeof(fgk_fftplan));
    if (!p) { fgk_set_error("out of host memory"); return -1; }
    p->Ng = Ng;
    p->prec = precision;
    if (!general) {
        factor_stages(Ng, &p->sN);
        factor_stages(Ng / 2, &p->sH);
    }
    size_t bytes = (size_t)Ng * 2 * precision;
    void *host = malloc(bytes);
    if (!host) { free(p); fgk_set_error("out of host memory"); return -1; }
    for (int t = 0; t < Ng; ++t) {
        double ang = 2.0 * M_PI * (double)t / (double)Ng;
        double c = cos(ang), s = -sin(ang);
        if (precision == 8) {
            ((double *)host)[2 * t] = c;
            ((double *)host)[2 * t + 1] = s;
        } else {
            ((float *)host)[2 * t] = (float)c;
            ((float *)host)[2 * t + 1] = (float)s;
        }
    }
    if (fgk_malloc(&p->tw, bytes)) { free(host); free(p); return -1; }
    hipError_t e = hipMemcpy(p->tw, host, bytes, hipMemcpyHostToDevice);
    free(host);
    if (e != hipSuccess) {
        fgk_free(p->tw);
        free(p);
        return fgk_hip_fail(e, "hipMemcpy twiddles", __FILE__, __LINE__);
    }
    if (general && general_setup(p)) {
        fgk_fft_plan_destroy(p);
        return -1;
    }
    if (precision == 4 && fgk_fft_f32_specialised(Ng)) {
        p->spec = 1;
        if (stage_twiddles(Ng, &p->stw_cols) || stage_twiddles(Ng / 2, &p->stw_rows)) {
            fgk_fft_plan_destroy(p);
            return -1;
        }
        // the i-axis column pass of large grids strides over Ng planes of
        // Ng^2/2 complex (16 MB apart at Ng = 2048), one row per plane per
        // workgroup: split it into passes of n1 and Ng / n1 rows
        // (FFTGAL_FFT_FOURSTEP=n1 chooses the split for any grid, 0 turns
        // it off)
        int n1 = (Ng >= 2048 && Ng % 64 == 0) ? 64 : 0;
        const char *fs = getenv("FFTGAL_FFT_FOURSTEP");
        if (fs) n1 = atoi(fs);
        if (n1 > 1 && Ng % n1 == 0 && fgk_fft_f32_specialised(n1) &&
            fgk_fft_f32_specialised(Ng / n1)) {
            p->fs_n1 = n1;
            if (stage_twiddles(n1, &p->stw_fs1) || stage_twiddles(Ng / n1, &p->stw_fs2)) {
                fgk_fft_plan_destroy(p);
                return -1;
            }
        }
    }
    *out = p;
    return 0;
}

extern "C" int fgk_fft_fourstep(const fgk_fftplan *p) { return p->spec ? p->fs_n1 : 0; }

extern "C" int fgk_fft_force_generic(fgk_fftplan *p, int on) {
    p->spec = (!on && p->stw_cols && p->stw_rows) ? 1 : 0;
    return 0;
}

extern "C" void fgk_fft_plan_destroy(fgk_fftplan *p) {
    if (!p) return;
    fgk_free(p->tw);
    fgk_free(p->twM);
    fgk_free(p->chirp);
    fgk_free(p->bf[0]);
    fgk_free(p->bf[1]);
    fgk_free(p->stw_cols);
    fgk_free(p->stw_rows);
    fgk_free(p->stw_fs1);
    fgk_free(p->stw_fs2);
    free(p);
}

static int cols_per_wg(const fgk_fftplan *p, size_t elem);
static int rows_per_wg(const fgk_fftplan *p, size_t elem, int extra_row);

extern "C" int fgk_fft_launch_shape(const fgk_fftplan *p, int *cols, int *rows) {
    const size_t elem = p->prec == 8 ? sizeof(cpx<double>) : sizeof(cpx<float>);
    if (cols) *cols = cols_per_wg(p, elem);
    if (rows) *rows = rows_per_wg(p, elem, 1);
    return 0;
}

extern "C" int fgk_fft_set_tuning(fgk_fftplan *p, int cols_per_wg, int rows_per_wg) {
    if (cols_per_wg && (cols_per_wg & (cols_per_wg - 1))) {
        fgk_set_error("cols_per_wg must be a power of two");
        return -2;
    }
    p->cols_per_wg = cols_per_wg;
    p->rows_per_wg = rows_per_wg;
    return 0;
}

static const size_t kLdsBudget = 64 * 1024;

// LDS per row: two ping-pong pencils of Ng/2 points, plus the staged half
// spectrum (Ng/2+1 points) for the c2r pre-processing
static size_t rows_lds(int P, int M, size_t elem, int extra_row) {
    return (size_t)P * (2 * M + extra_row * (M + 1)) * elem;
}

static int rows_per_wg(const fgk_fftplan *p, size_t elem, int extra_row) {
    const int M = p->Ng / 2;
    if (p->rows_per_wg > 0) return p->rows_per_wg;
    int P = 1;
    while (P * M < 1024 && rows_lds(2 * P, M, elem, extra_row) <= kLdsBudget) P *= 2;
    return P;
}

static int cols_per_wg(const fgk_fftplan *p, size_t elem) {
    if (p->cols_per_wg > 0) return p->cols_per_wg;
    int C = 16;
    while (C > 1 && (size_t)2 * p->Ng * C * elem > kLdsBudget) C >>= 1;
    return C;
}

static int ilog2(int c) {
    int l = 0;
    while ((1 << l) < c) ++l;
    return l;
}

// packed: the FP32 packed half-spectrum layout (Ng/2 columns), else the
// padded layout (Ng/2+1 columns in rows of Ng/2+1 complex)
// Column passes of nbatch contiguous grids.  axis 1 (along j): N * nbatch
// planes of Nk columns at stride rowc; axis 0 (along i): nbatch grids of
// N * Nk columns (the whole (j, k) plane is contiguous) at stride planec.
struct ColShape {
    int ncols;
    long stride_t, stride_o, nouter;
};

static ColShape col_shape(int N, int Nk, int axis, int nbatch) {
    const long rowc = Nk;                  // complex elements between j rows
    const long planec = (long)N * Nk;      // complex elements between i planes
    if (axis == 1) return ColShape{Nk, rowc, planec, (long)N * nbatch};
    return ColShape{N * Nk, planec, (long)N * planec, (long)nbatch};
}

template <typename T>
static int launch_cols(fgk_fftplan *p, T *f, int axis, int sign, T scale, hipStream_t s,
                       int packed = 0, int nbatch = 1) {
    const int N = p->Ng, Nk = packed ? N / 2 : N / 2 + 1;
    const int C = cols_per_wg(p, sizeof(cpx<T>));
    const ColShape cs = col_shape(N, Nk, axis, nbatch);
    const long nb = cs.nouter * ((cs.ncols + C - 1) / C);
    size_t lds = (size_t)2 * N * C * sizeof(cpx<T>);
    hipLaunchKernelGGL(k_cols<T>, dim3((unsigned)nb), dim3(256), lds, s, (cpx<T> *)f, N,
                       cs.ncols, cs.stride_t, cs.stride_o, C, ilog2(C), p->sN,
                       (const cpx<T> *)p->tw, sign, scale);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_fft_cols_f32(fgk_fftplan *p, float *f, int axis, int sign, int nbatch,
                                fgk_stream_t s) {
    if (p->prec != 4) { fgk_set_error("plan precision is not FP32"); return -2; }
    if (nbatch < 1) { fgk_set_error("FFT: nbatch must be >= 1"); return -2; }
    const int N = p->Ng, M = N / 2;
    if (!p->spec) return launch_cols<float>(p, f, axis, sign, 1.0f, (hipStream_t)s, 1, nbatch);
    if (axis == 0 && p->fs_n1) {
        // four-step over i = n2 + N2 n1 (N1 = fs_n1, N2 = N / N1), per grid:
        // pass A, length N1 over n1 (planes N2 apart) for each n2, output k1
        // times W_N^(n2 k1); pass B, length N2 over n2 (consecutive planes)
        // for each k1.  Frequency i = k1 + N1 k2 lands on plane k2 + N2 k1.
        const int N1 = p->fs_n1, N2 = N / N1;
        const long P = (long)N * M;                   // complex per i-plane
        for (int g = 0; g < nbatch; ++g) {
            cpx<float> *b = (cpx<float> *)f + (long)g * N * P;
            if (launch_cols32_wide(N1, b, (int)P, N2 * P, P, N2, (const cpx<float> *)p->stw_fs1,
                                   (const cpx<float> *)p->tw, sign, (hipStream_t)s,
                                   (const cpx<float> *)p->tw, N / N1))
                return -1;
            if (launch_cols32_wide(N2, b, (int)P, P, N2 * P, N1, (const cpx<float> *)p->stw_fs2,
                                   (const cpx<float> *)p->tw, sign, (hipStream_t)s, nullptr,
                                   N / N2))
                return -1;
        }
        return 0;
    }
    const ColShape cs = col_shape(N, M, axis, nbatch);
    return launch_cols32(N, (cpx<float> *)f, cs.ncols, cs.stride_t, cs.stride_o, cs.nouter,
                         (const cpx<float> *)p->stw_cols, (const cpx<float> *)p->tw, sign,
                         (hipStream_t)s);
}

static bool multi_side(int N) {
    switch (N) {
#define CASE_NG(n) case n:
        FGK_MULTI_NG(CASE_NG)
#undef CASE_NG
        return true;
    default:
        return false;
    }
}

static int multi_rows_P(int N) {
    switch (N) {
#define CASE_NG(n) case n: return rows_P(n / 2);
        FGK_MULTI_NG(CASE_NG)
#undef CASE_NG
    default:
        return 1;
    }
}

static int multi_cols_C(int N) {
    switch (N) {
#define CASE_NG(n) case n: return mcols_C(n);
        FGK_MULTI_NG(CASE_NG)
#undef CASE_NG
    default:
        return 1;
    }
}

extern "C" int fgk_fft_multi_create(fgk_fftmulti **out, int ngroups, fgk_fftplan *const *plans,
                                    float *const *f, int64_t *const *src0,
                                    int64_t *const *src1, const int *nbatch) {
    *out = nullptr;
    if (ngroups < 1) return 0;
    for (int g = 0; g < ngroups; ++g) {
        const fgk_fftplan *p = plans[g];
        if (p->prec != 4 || !p->spec || p->fs_n1 || !multi_side(p->Ng) || nbatch[g] < 1)
            return 0;   // a side without a multi-size body: the caller launches per group
    }
    MultiGroup *h = (MultiGroup *)calloc((size_t)ngroups, sizeof(MultiGroup));
    fgk_fftmulti *m = (fgk_fftmulti *)calloc(1, sizeof(fgk_fftmulti));
    if (!h || !m) {
        free(h);
        free(m);
        fgk_set_error("out of host memory");
        return -1;
    }
    unsigned long rb = 0, cb[2] = {0, 0};
    for (int g = 0; g < ngroups; ++g) {
        const fgk_fftplan *p = plans[g];
        const int N = p->Ng, M = N / 2;
        MultiGroup &G = h[g];
        G.f = (cpx<float> *)f[g];
        G.src[0] = src0[g];
        G.src[1] = src1 ? src1[g] : nullptr;
        G.N = N;
        G.nbatch = nbatch[g];
        G.stw_rows = (const cpx<float> *)p->stw_rows;
        G.stw_cols = (const cpx<float> *)p->stw_cols;
        G.tw = (const cpx<float> *)p->tw;
        G.rblk0 = (unsigned)rb;
        const int P = multi_rows_P(N), C = multi_cols_C(N);
        rb += ((unsigned long)N * N * nbatch[g] + P - 1) / P;
        for (int axis = 0; axis < 2; ++axis) {
            const ColShape cs = col_shape(N, M, axis, nbatch[g]);
            G.ncols[axis] = cs.ncols;
            G.stride_t[axis] = cs.stride_t;
            G.stride_o[axis] = cs.stride_o;
            G.cblk0[axis] = (unsigned)cb[axis];
            cb[axis] += (unsigned long)cs.nouter * ((cs.ncols + C - 1) / C);
        }
    }
    if (rb >= (1ul << 31) || cb[0] >= (1ul << 31) || cb[1] >= (1ul << 31)) {
        free(h);
        free(m);
        fgk_set_error("FFT multi: too many workgroups for one launch");
        return -2;
    }
    m->ng = ngroups;
    m->rblocks = (unsigned)rb;
    m->cblocks[0] = (unsigned)cb[0];
    m->cblocks[1] = (unsigned)cb[1];
    int rc = fgk_malloc((void **)&m->d, (size_t)ngroups * sizeof(MultiGroup));
    if (!rc) rc = fgk_h2d(m->d, h, (size_t)ngroups * sizeof(MultiGroup), nullptr);
    if (!rc) rc = fgk_stream_sync(nullptr);
    free(h);
    if (rc) {
        fgk_free(m->d);
        free(m);
        return rc;
    }
    *out = m;
    return 0;
}

extern "C" void fgk_fft_multi_destroy(fgk_fftmulti *m) {
    if (!m) return;
    fgk_free(m->d);
    free(m);
}

extern "C" int fgk_fft_multi_rows(fgk_fftmulti *m, int set, int clear_src, fgk_stream_t s) {
    hipLaunchKernelGGL(k_rows32_multi, dim3(m->rblocks), dim3(1024), 0, (hipStream_t)s, m->d,
                       m->ng, set, clear_src);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_fft_multi_cols(fgk_fftmulti *m, int axis, int sign, fgk_stream_t s) {
    hipLaunchKernelGGL(k_cols32_multi, dim3(m->cblocks[axis]), dim3(1024), 0, (hipStream_t)s,
                       m->d, m->ng, axis, sign);
    FGK_LAUNCHED();
    return 0;
}

static GenFFT gen_args(const fgk_fftplan *p) {
    GenFFT G;
    G.N = p->Ng;
    G.M = p->blue_M;
    G.sN = p->sN;
    G.sM = p->sM;
    G.tw = (const cpx<double> *)p->tw;
    G.twM = (const cpx<double> *)p->twM;
    G.chirp = (const cpx<double> *)p->chirp;
    G.bf[0] = (const cpx<double> *)p->bf[0];
    G.bf[1] = (const cpx<double> *)p->bf[1];
    return G;
}

// LDS of general pencils: two buffers of L complex doubles per pencil, up
// to 64 KB per workgroup
static const size_t kGenLds = 64 * 1024;

static int general_rows(fgk_fftplan *p, double *f, int dir, double scale, hipStream_t s) {
    const GenFFT G = gen_args(p);
    const int N = p->Ng, L = G.M ? G.M : N, Ng_pad = 2 * (N / 2 + 1);
    const long nrows = (long)N * N, npairs = (nrows + 1) / 2;
    int P = 1;
    while (2 * P * L < 1024 && (size_t)2 * 2 * P * L * sizeof(cpx<double>) <= kGenLds) P *= 2;
    const size_t lds = (size_t)2 * P * L * sizeof(cpx<double>);
    hipLaunchKernelGGL(k_rows_gen, dim3(fgk_ceil_div(npairs, P)), dim3(256), lds, s, f, Ng_pad,
                       nrows, G, P, dir, scale);
    FGK_LAUNCHED();
    return 0;
}

static int general_cols(fgk_fftplan *p, double *f, int axis, int sign, double scale,
                        hipStream_t s) {
    const GenFFT G = gen_args(p);
    const int N = p->Ng, L = G.M ? G.M : N, Nk = N / 2 + 1;
    int C = 16;
    while (C > 1 && (size_t)2 * L * C * sizeof(cpx<double>) > kGenLds) C >>= 1;
    const ColShape cs = col_shape(N, Nk, axis, 1);
    const long nb = cs.nouter * ((cs.ncols + C - 1) / C);
    const size_t lds = (size_t)2 * L * C * sizeof(cpx<double>);
    hipLaunchKernelGGL(k_cols_gen, dim3((unsigned)nb), dim3(256), lds, s, (cpx<double> *)f,
                       cs.ncols, cs.stride_t, cs.stride_o, C, ilog2(C), G, sign, scale);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_fft_r2c_f64(fgk_fftplan *p, double *f, double scale, fgk_stream_t st) {
    if (p->prec != 8) { fgk_set_error("plan precision is not FP64"); return -2; }
    hipStream_t s = (hipStream_t)st;
    if (p->general) {
        if (general_rows(p, f, 0, 1.0, s)) return -1;
        if (general_cols(p, f, 1, -1, 1.0, s)) return -1;
        return general_cols(p, f, 0, -1, scale, s);
    }
    const int N = p->Ng, M = N / 2;
    const long nrows = (long)N * N;
    const int P = rows_per_wg(p, sizeof(cpx<double>), 0);
    size_t lds = (size_t)2 * P * M * sizeof(cpx<double>);
    hipLaunchKernelGGL((k_rows_r2c<double, 0>), dim3(fgk_ceil_div(nrows, P)), dim3(256), lds, s,
                       f, (int64_t *)nullptr, 1.0, 0, 1.0, N, 2 * (M + 1), nrows, p->sH,
                       (const cpx<double> *)p->tw, P);
    FGK_LAUNCHED();
    if (launch_cols<double>(p, f, 1, -1, 1.0, s)) return -1;
    return launch_cols<double>(p, f, 0, -1, scale, s);
}

extern "C" int fgk_fft_c2r_f64(fgk_fftplan *p, double *f, double scale, fgk_stream_t st) {
    if (p->prec != 8) { fgk_set_error("plan precision is not FP64"); return -2; }
    hipStream_t s = (hipStream_t)st;
    if (p->general) {
        if (general_cols(p, f, 0, +1, 1.0, s)) return -1;
        if (general_cols(p, f, 1, +1, 1.0, s)) return -1;
        return general_rows(p, f, 1, scale, s);
    }
    const int N = p->Ng, M = N / 2;
    const long nrows = (long)N * N;
    if (launch_cols<double>(p, f, 0, +1, 1.0, s)) return -1;
    if (launch_cols<double>(p, f, 1, +1, 1.0, s)) return -1;
    const int P = rows_per_wg(p, sizeof(cpx<double>), 1);
    size_t lds = (size_t)P * (2 * M + M + 1) * sizeof(cpx<double>);
    hipLaunchKernelGGL(k_rows_c2r<double>, dim3(fgk_ceil_div(nrows, P)), dim3(256), lds, s, f,
                       scale, N, 2 * (M + 1), nrows, p->sH, (const cpx<double> *)p->tw, P);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_fft_rows_f32_from_fixed(fgk_fftplan *p, float *f, int64_t *src,
                                           double in_scale, int clear_src, int nbatch,
                                           fgk_stream_t st) {
    if (p->prec != 4) { fgk_set_error("plan precision is not FP32"); return -2; }
    if (nbatch < 1) { fgk_set_error("FFT: nbatch must be >= 1"); return -2; }
    hipStream_t s = (hipStream_t)st;
    const int N = p->Ng, M = N / 2;
    const long nrows = (long)N * N * nbatch;   // rows of contiguous grids
    if (p->spec)
        return launch_rows32(N, nrows, (cpx<float> *)f, src, in_scale, clear_src,
                             (const cpx<float> *)p->stw_rows, (const cpx<float> *)p->tw, s);
    const int P = rows_per_wg(p, sizeof(cpx<float>), 0);
    size_t lds = (size_t)2 * P * M * sizeof(cpx<float>);
    hipLaunchKernelGGL((k_rows_r2c<float, 2>), dim3(fgk_ceil_div(nrows, P)), dim3(256), lds, s,
                       f, src, in_scale, clear_src, 1.0f, N, 2 * (M + 1), nrows, p->sH,
                       (const cpx<float> *)p->tw, P);
    FGK_LAUNCHED();
    return 0;
}

extern "C" int fgk_fft_r2c_f32_from_fixed(fgk_fftplan *p, float *f, int64_t *src,
                                          double in_scale, int clear_src, fgk_stream_t s) {
    if (fgk_fft_rows_f32_from_fixed(p, f, src, in_scale, clear_src, 1, s)) return -1;
    if (fgk_fft_cols_f32(p, f, 1, -1, 1, s)) return -1;
    return fgk_fft_cols_f32(p, f, 0, -1, 1, s);
}
