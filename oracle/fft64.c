/*
 * Own FP64 FFT for the CPU oracle (test infrastructure, see oracle.h).
 *
 * The reference calls FFTW3 (third-party, not vendored, version unpinned:
 * Makefile:3 "-lfftw3"); its call sites are src/fft.c:28-29 (plan r2c/c2r 3d)
 * and src/fft.c:174,186 (fftw_execute).  FFTW is absent here, so this file
 * restates the published transform FFTW computes:
 *   r2c: F[i,j,k] = sum_{x,y,z} f[x,y,z] exp(-2 pi i (ix+jy+kz)/Ng), k = 0..Ng/2,
 *   c2r: f[x,y,z] = sum over the Hermitian-extended half spectrum of
 *        F exp(+2 pi i (...)/Ng),
 * both unnormalised, in place on the padded layout f[(i*Ng+j)*Ng_pad+k]
 * (src/fft.h:10-13,20).
 *
 * Algorithm: Stockham autosort mixed-radix FFT (Stockham 1966; formulation of
 * Govindaraju et al., SC'08) with radix 4, 2, 3 butterflies written out and a
 * direct DFT for other prime radices; twiddles exp(-2 pi i t/n) computed once
 * in FP64 with libm.  Real rows of even length use the standard half-length
 * complex transform plus the split post-processing step; odd lengths use a
 * full complex transform of the real row.
 *
 * Pinned by numpy.fft.rfftn/irfftn golden vectors (tests/golden), brute-force
 * DFT, Parseval and the octet known answer (src/tests/octet.c:33-56).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <assert.h>
#include <pthread.h>
#include "oracle.h"

static int orc_nthreads = 1;

void oracle_set_threads(int n) { orc_nthreads = n < 1 ? 1 : (n > 256 ? 256 : n); }
int oracle_threads(void) { return orc_nthreads; }

typedef struct orc_range {
    void (*fn)(void *, long, long);
    void *arg;
    long lo, hi;
} orc_range;

static void *orc_range_run(void *p) {
    orc_range *r = (orc_range *)p;
    r->fn(r->arg, r->lo, r->hi);
    return NULL;
}

void orc_parallel(long n, void (*fn)(void *arg, long lo, long hi), void *arg) {
    int T = orc_nthreads;
    if (T > n) T = n > 0 ? (int)n : 1;
    if (T <= 1) {
        fn(arg, 0, n);
        return;
    }
    orc_range r[256];
    pthread_t th[256];
    for (int t = 0; t < T; ++t) {
        r[t].fn = fn;
        r[t].arg = arg;
        r[t].lo = n * t / T;
        r[t].hi = n * (t + 1) / T;
    }
    for (int t = 1; t < T; ++t)
        assert(pthread_create(&th[t], NULL, orc_range_run, &r[t]) == 0);
    orc_range_run(&r[0]);
    for (int t = 1; t < T; ++t)
        pthread_join(th[t], NULL);
}

int offt64_plan_init(offt64_plan *p, int n) {
    assert(n >= 1);
    p->n = n;
    p->nstage = 0;
    int m = n;
    while (m % 4 == 0) { p->radix[p->nstage++] = 4; m /= 4; }
    while (m % 2 == 0) { p->radix[p->nstage++] = 2; m /= 2; }
    for (int q = 3; m > 1; q += 2)
        while (m % q == 0) { p->radix[p->nstage++] = q; m /= q; }
    p->tw = (double *)malloc(2 * (size_t)n * sizeof(double));
    if (!p->tw) return -1;
    for (int t = 0; t < n; ++t) {
        double a = 2.0 * M_PI * (double)t / (double)n;
        p->tw[2 * t] = cos(a);
        p->tw[2 * t + 1] = -sin(a);
    }
    return 0;
}

void offt64_plan_free(offt64_plan *p) {
    if (p) { free(p->tw); p->tw = NULL; }
}

/* one Stockham stage of radix R over n points: a -> b */
static void stage(const offt64_plan *p, const double *a, double *b,
                  int R, int Ns, int sign) {
    const int n = p->n, m = n / R;
    const int tstep = n / (Ns * R);
    const double *tw = p->tw;
    double vr[64], vi[64];
    double *gr = NULL, *gi = NULL;
    if (R > 64) {
        gr = (double *)malloc((size_t)R * 2 * sizeof(double));
        assert(gr != NULL);
        gi = gr + R;
    }
    double *ur = gr ? gr : vr, *ui = gr ? gi : vi;
    for (int j = 0; j < m; ++j) {
        const int k = j % Ns;
        for (int r = 0; r < R; ++r) {
            double xr = a[2 * (j + r * m)], xi = a[2 * (j + r * m) + 1];
            if (r > 0 && k > 0) {
                int t = r * k * tstep;
                double wr = tw[2 * t], wi = sign < 0 ? tw[2 * t + 1] : -tw[2 * t + 1];
                double yr = xr * wr - xi * wi;
                double yi = xr * wi + xi * wr;
                xr = yr; xi = yi;
            }
            ur[r] = xr; ui[r] = xi;
        }
        const int base = (j / Ns) * Ns * R + k;
        if (R == 2) {
            double ar = ur[0], ai = ui[0], br = ur[1], bi = ui[1];
            b[2 * base] = ar + br;          b[2 * base + 1] = ai + bi;
            b[2 * (base + Ns)] = ar - br;   b[2 * (base + Ns) + 1] = ai - bi;
        } else if (R == 4) {
            double s0r = ur[0] + ur[2], s0i = ui[0] + ui[2];
            double d0r = ur[0] - ur[2], d0i = ui[0] - ui[2];
            double s1r = ur[1] + ur[3], s1i = ui[1] + ui[3];
            double d1r = ur[1] - ur[3], d1i = ui[1] - ui[3];
            /* sign*i*d1 = (-sign*d1i, sign*d1r) */
            double er = -sign * d1i, ei = sign * d1r;
            b[2 * base] = s0r + s1r;               b[2 * base + 1] = s0i + s1i;
            b[2 * (base + Ns)] = d0r + er;         b[2 * (base + Ns) + 1] = d0i + ei;
            b[2 * (base + 2 * Ns)] = s0r - s1r;    b[2 * (base + 2 * Ns) + 1] = s0i - s1i;
            b[2 * (base + 3 * Ns)] = d0r - er;     b[2 * (base + 3 * Ns) + 1] = d0i - ei;
        } else if (R == 3) {
            const double h = 0.86602540378443864676;  /* sqrt(3)/2 */
            double t1r = ur[1] + ur[2], t1i = ui[1] + ui[2];
            double t2r = ur[0] - 0.5 * t1r, t2i = ui[0] - 0.5 * t1i;
            double t3r = sign * h * (ur[1] - ur[2]), t3i = sign * h * (ui[1] - ui[2]);
            /* X1 = t2 + i t3, X2 = t2 - i t3 */
            b[2 * base] = ur[0] + t1r;             b[2 * base + 1] = ui[0] + t1i;
            b[2 * (base + Ns)] = t2r - t3i;        b[2 * (base + Ns) + 1] = t2i + t3r;
            b[2 * (base + 2 * Ns)] = t2r + t3i;    b[2 * (base + 2 * Ns) + 1] = t2i - t3r;
        } else {
            const int rstep = n / R;
            for (int q = 0; q < R; ++q) {
                double sr = 0, si = 0;
                for (int r = 0; r < R; ++r) {
                    int t = ((q * r) % R) * rstep;
                    double wr = tw[2 * t], wi = sign < 0 ? tw[2 * t + 1] : -tw[2 * t + 1];
                    sr += ur[r] * wr - ui[r] * wi;
                    si += ur[r] * wi + ui[r] * wr;
                }
                b[2 * (base + q * Ns)] = sr;
                b[2 * (base + q * Ns) + 1] = si;
            }
        }
    }
    free(gr);
}

void offt64_exec(const offt64_plan *p, double *x, double *work, int sign) {
    double *a = x, *b = work;
    int Ns = 1;
    for (int s = 0; s < p->nstage; ++s) {
        int R = p->radix[s];
        stage(p, a, b, R, Ns, sign);
        double *t = a; a = b; b = t;
        Ns *= R;
    }
    if (a != x)
        memcpy(x, a, 2 * (size_t)p->n * sizeof(double));
}

/* real row (N doubles at row[0..N)) -> half spectrum row[0..2(N/2+1)) */
static void row_r2c(const offt_t *g, double *row, double *buf, double *work) {
    const int N = g->Ng;
    if (N % 2 == 0) {
        const int M = N / 2;
        memcpy(buf, row, (size_t)N * sizeof(double));   /* z[m] = x[2m] + i x[2m+1] */
        offt64_exec(g->pH, buf, work, -1);
        const double *tw = g->pN->tw;                   /* exp(-2 pi i k / N) */
        for (int k = 0; k <= M; ++k) {
            int k1 = k % M, k2 = (M - k) % M;
            double zr = buf[2 * k1], zi = buf[2 * k1 + 1];
            double cr = buf[2 * k2], ci = -buf[2 * k2 + 1];  /* conj Z[M-k] */
            double er = 0.5 * (zr + cr), ei = 0.5 * (zi + ci);
            /* O = -i (Z - conj) / 2 */
            double dr = zr - cr, di = zi - ci;
            double or_ = 0.5 * di, oi = -0.5 * dr;
            double wr, wi;
            if (k < N) { wr = tw[2 * k]; wi = tw[2 * k + 1]; } else { wr = 1; wi = 0; }
            row[2 * k] = er + (wr * or_ - wi * oi);
            row[2 * k + 1] = ei + (wr * oi + wi * or_);
        }
    } else {
        for (int n = 0; n < N; ++n) { buf[2 * n] = row[n]; buf[2 * n + 1] = 0; }
        offt64_exec(g->pN, buf, work, -1);
        memcpy(row, buf, (size_t)2 * (N / 2 + 1) * sizeof(double));
    }
}

/* half spectrum row -> real row (unnormalised inverse) */
static void row_c2r(const offt_t *g, double *row, double *buf, double *work) {
    const int N = g->Ng;
    if (N % 2 == 0) {
        const int M = N / 2;
        const double *tw = g->pN->tw;
        for (int k = 0; k < M; ++k) {
            double xr = row[2 * k], xi = row[2 * k + 1];
            double cr = row[2 * (M - k)], ci = -row[2 * (M - k) + 1];  /* conj X[M-k] */
            double er = xr + cr, ei = xi + ci;
            double dr = xr - cr, di = xi - ci;
            double wr = tw[2 * k], wi = -tw[2 * k + 1];                /* exp(+2 pi i k/N) */
            double or_ = dr * wr - di * wi, oi = dr * wi + di * wr;
            buf[2 * k] = er - oi;       /* Z = E + i O */
            buf[2 * k + 1] = ei + or_;
        }
        offt64_exec(g->pH, buf, work, +1);
        memcpy(row, buf, (size_t)N * sizeof(double));
    } else {
        const int H = N / 2;
        for (int k = 0; k <= H; ++k) { buf[2 * k] = row[2 * k]; buf[2 * k + 1] = row[2 * k + 1]; }
        for (int k = H + 1; k < N; ++k) {
            buf[2 * k] = row[2 * (N - k)];
            buf[2 * k + 1] = -row[2 * (N - k) + 1];
        }
        offt64_exec(g->pN, buf, work, +1);
        for (int n = 0; n < N; ++n) row[n] = buf[2 * n];
    }
}

#define OBLK 8

typedef struct axis_args {
    offt_t *g;
    int axis, sign;
} axis_args;

/* complex transforms along j (axis=1) or i (axis=0) for all k <= Ng/2, for
 * the other spatial axis o in [o0, o1) */
static void axis_c2c_range(void *arg, long o0, long o1) {
    const axis_args *a = (const axis_args *)arg;
    offt_t *g = a->g;
    const int axis = a->axis, sign = a->sign;
    const int N = g->Ng, Nk = N / 2 + 1;
    const long rowstride = g->Ng_pad;            /* doubles between j rows */
    const long planestride = (long)N * g->Ng_pad; /* doubles between i planes */
    double *buf = (double *)malloc((size_t)OBLK * 2 * N * sizeof(double));
    double *work = (double *)malloc((size_t)2 * N * sizeof(double));
    assert(buf && work);
    for (long o = o0; o < o1; ++o) {
        for (int k0 = 0; k0 < Nk; k0 += OBLK) {
            int nb = Nk - k0 < OBLK ? Nk - k0 : OBLK;
            for (int t = 0; t < N; ++t) {
                long off = axis == 1 ? o * planestride + (long)t * rowstride
                                     : (long)t * planestride + o * rowstride;
                const double *src = g->f + off + 2 * k0;
                for (int b = 0; b < nb; ++b) {
                    buf[(size_t)b * 2 * N + 2 * t] = src[2 * b];
                    buf[(size_t)b * 2 * N + 2 * t + 1] = src[2 * b + 1];
                }
            }
            for (int b = 0; b < nb; ++b)
                offt64_exec(g->pN, buf + (size_t)b * 2 * N, work, sign);
            for (int t = 0; t < N; ++t) {
                long off = axis == 1 ? o * planestride + (long)t * rowstride
                                     : (long)t * planestride + o * rowstride;
                double *dst = g->f + off + 2 * k0;
                for (int b = 0; b < nb; ++b) {
                    dst[2 * b] = buf[(size_t)b * 2 * N + 2 * t];
                    dst[2 * b + 1] = buf[(size_t)b * 2 * N + 2 * t + 1];
                }
            }
        }
    }
    free(buf);
    free(work);
}

static void axis_c2c(offt_t *g, int axis, int sign) {
    axis_args a = {g, axis, sign};
    orc_parallel(g->Ng, axis_c2c_range, &a);
}

/* real rows [r0, r1) of the grid, forward (r2c) or inverse (c2r) */
static void rows_r2c(void *arg, long r0, long r1) {
    offt_t *g = (offt_t *)arg;
    double *buf = (double *)malloc((size_t)2 * g->Ng * sizeof(double));
    double *work = (double *)malloc((size_t)2 * g->Ng * sizeof(double));
    assert(buf && work);
    for (long r = r0; r < r1; ++r)
        row_r2c(g, g->f + r * g->Ng_pad, buf, work);
    free(buf);
    free(work);
}

static void rows_c2r(void *arg, long r0, long r1) {
    offt_t *g = (offt_t *)arg;
    double *buf = (double *)malloc((size_t)2 * g->Ng * sizeof(double));
    double *work = (double *)malloc((size_t)2 * g->Ng * sizeof(double));
    assert(buf && work);
    for (long r = r0; r < r1; ++r)
        row_c2r(g, g->f + r * g->Ng_pad, buf, work);
    free(buf);
    free(work);
}

void offt64_r2c_3d(offt_t *g) {
    orc_parallel((long)g->Ng * g->Ng, rows_r2c, g);
    axis_c2c(g, 1, -1);
    axis_c2c(g, 0, -1);
}

void offt64_c2r_3d(offt_t *g) {
    axis_c2c(g, 0, +1);
    axis_c2c(g, 1, +1);
    orc_parallel((long)g->Ng * g->Ng, rows_c2r, g);
}
