/* QPM mock catalogs in ASCII (src/io/qpm.c:9-34 of eelregit/fftgal): one
 * galaxy per line, nine whitespace-separated columns
 *     x y z vx vy vz  halo_mass satellite_flag halo_id
 * of which the first six are kept.
 *
 * The reference reads the file twice with stdio (count '\n', then fscanf
 * line by line).  Here the file is mapped once and cut into line-aligned
 * chunks of about 8 MiB; threads count the lines of the chunks, and then
 * parse each chunk into its own slice of the arrays.  Numbers are converted
 * by strtod, which rounds correctly as fscanf's %lf does, so the values are
 * the reference's bit for bit.
 *
 * gal_loadqpm_cubic      -- the drop-in (double arrays with velocities).
 * fftgal_qpm_load_device -- positions straight to device float32 arrays for
 *   the fast path: consecutive chunks are parsed, a piece at a time, into
 *   one of two pinned host staging buffers, and each piece's copy to the
 *   device runs while the next piece is parsed. */
#define _GNU_SOURCE
#include <errno.h>
#include <fcntl.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include "../../../include/gal.h"
#include "../../../include/io.h"
#include "../../../include/fftgal_batch.h"
#include "../fgk.h"
#include "internal.h"

#define QPM_MAX_THREADS 64
#define QPM_LINE_MAX 1024

/* chunk size (8 MiB) and staged lines per piece (2^22: more than a chunk can
 * hold, since a line with six numbers has at least 12 characters); both can
 * be lowered through the environment (FFTGAL_QPM_CHUNK_BYTES,
 * FFTGAL_QPM_PIECE_LINES) to exercise many chunks and pieces on small files */
static long env_long(const char *name, long dflt, long lo, long hi) {
    const char *e = getenv(name);
    long v = e ? atol(e) : dflt;
    return v < lo ? lo : (v > hi ? hi : v);
}
#define QPM_CHUNK_BYTES env_long("FFTGAL_QPM_CHUNK_BYTES", 8l << 20, 4096, 1l << 30)
#define QPM_PIECE_LINES env_long("FFTGAL_QPM_PIECE_LINES", 1l << 22, 1024, 1l << 26)

typedef struct qpm_chunk {
    const char *a, *b;     /* [a, b): whole lines ('\n'-terminated) but the last chunk's tail */
    long lines, first;     /* '\n' count (the reference's line count), first line's index */
    long bad;              /* first line (0-based, file-wide) without six numbers, or -1 */
} qpm_chunk;

typedef struct qpm_file {
    const char *p;
    size_t n;
    int fd;
    qpm_chunk *c;
    long nc, lines;
} qpm_file;

static void qpm_close(qpm_file *f) {
    if (f->p) munmap((void *)f->p, f->n);
    if (f->fd >= 0) close(f->fd);
    free(f->c);
}

static int qpm_threads(long nchunks) {
    long c = sysconf(_SC_NPROCESSORS_ONLN);
    const char *e = getenv("FFTGAL_IO_THREADS");
    if (!(e && atoi(e) > 0)) e = getenv("OMP_NUM_THREADS");
    if (e && atoi(e) > 0 && atoi(e) < c) c = atoi(e);
    if (c < 1) c = 1;
    if (c > QPM_MAX_THREADS) c = QPM_MAX_THREADS;
    return (int)(c < nchunks ? c : (nchunks > 0 ? nchunks : 1));
}

/* run fn(ctx, chunk index) over chunks [c0, c1) on up to nt threads, each
 * thread taking the chunks t, t + nt, ... */
typedef struct qpm_job {
    void (*fn)(void *, long);
    void *ctx;
    long c0, c1;
    int t, nt;
} qpm_job;

static void *qpm_worker(void *arg) {
    qpm_job *j = (qpm_job *)arg;
    for (long c = j->c0 + j->t; c < j->c1; c += j->nt) j->fn(j->ctx, c);
    return NULL;
}

static void qpm_parallel(int nt, long c0, long c1, void (*fn)(void *, long), void *ctx) {
    if (nt > c1 - c0) nt = (int)(c1 - c0);
    if (nt < 1) return;
    pthread_t th[QPM_MAX_THREADS];
    qpm_job job[QPM_MAX_THREADS];
    int ok[QPM_MAX_THREADS] = {0};
    for (int t = 0; t < nt; ++t) job[t] = (qpm_job){fn, ctx, c0, c1, t, nt};
    for (int t = 1; t < nt; ++t) ok[t] = pthread_create(&th[t], NULL, qpm_worker, &job[t]) == 0;
    qpm_worker(&job[0]);
    for (int t = 1; t < nt; ++t) {
        if (ok[t]) pthread_join(th[t], NULL);
        else qpm_worker(&job[t]);
    }
}

static void count_chunk(void *ctx, long i) {
    qpm_chunk *c = (qpm_chunk *)ctx + i;
    long n = 0;
    for (const char *p = c->a; p < c->b;) {
        const char *nl = memchr(p, '\n', (size_t)(c->b - p));
        if (!nl) break;
        ++n;
        p = nl + 1;
    }
    c->lines = n;
}

/* map the file, cut it into line-aligned chunks and count their lines */
static int qpm_open(const char *file, qpm_file *f) {
    memset(f, 0, sizeof *f);
    f->fd = open(file, O_RDONLY);
    if (f->fd < 0) {
        fgk_set_error("qpm: cannot open %s: %s", file, strerror(errno));
        return -1;
    }
    struct stat st;
    if (fstat(f->fd, &st) != 0) {
        fgk_set_error("qpm: cannot stat %s", file);
        return -1;
    }
    f->n = (size_t)st.st_size;
    if (f->n == 0) return 0;
    void *p = mmap(NULL, f->n, PROT_READ, MAP_PRIVATE, f->fd, 0);
    if (p == MAP_FAILED) {
        fgk_set_error("qpm: cannot map %s", file);
        return -1;
    }
    madvise(p, f->n, MADV_SEQUENTIAL);
    f->p = (const char *)p;
    const long cb = QPM_CHUNK_BYTES;
    const long maxc = (long)(f->n / cb) + 1;
    f->c = (qpm_chunk *)calloc((size_t)maxc, sizeof(qpm_chunk));
    if (!f->c) {
        fgk_set_error("out of host memory");
        return -1;
    }
    const char *end = f->p + f->n, *s = f->p;
    while (s < end) {
        const char *e = s + cb < end ? s + cb : end;
        if (e < end) {
            const char *nl = memchr(e, '\n', (size_t)(end - e));
            e = nl ? nl + 1 : end;
        }
        f->c[f->nc].a = s;
        f->c[f->nc].b = e;
        ++f->nc;
        s = e;
    }
    qpm_parallel(qpm_threads(f->nc), 0, f->nc, count_chunk, f->c);
    for (long i = 0; i < f->nc; ++i) {
        f->c[i].first = f->lines;
        f->lines += f->c[i].lines;
    }
    return 0;
}

/* six numbers of the line [p, nl); strtod converts as fscanf's %lf does.
 * The line is copied and terminated first, so that a short line cannot make
 * strtod run on past its end. */
static int parse_line(const char *p, const char *nl, double v[6]) {
    char buf[QPM_LINE_MAX];
    size_t len = (size_t)(nl - p);
    if (len >= sizeof buf) return -1;
    memcpy(buf, p, len);
    buf[len] = '\0';
    char *q = buf;
    for (int i = 0; i < 6; ++i) {
        char *t;
        v[i] = strtod(q, &t);
        if (t == q) return -1;
        q = t;
    }
    return 0;
}

/* parse targets: doubles with velocities, or float32 positions at an
 * offset into staging arrays */
typedef struct qpm_out {
    qpm_chunk *c;
    double *x, *y, *z, *vx, *vy, *vz;
    float *fx, *fy, *fz;
    long base;             /* line index of element 0 of the float arrays */
} qpm_out;

static void parse_chunk(void *ctx, long i) {
    qpm_out *o = (qpm_out *)ctx;
    qpm_chunk *c = o->c + i;
    const char *p = c->a;
    double v[6];
    c->bad = -1;
    for (long l = 0; l < c->lines; ++l) {
        const char *nl = memchr(p, '\n', (size_t)(c->b - p));
        const long k = c->first + l;
        if (parse_line(p, nl, v)) {
            c->bad = k;
            return;
        }
        if (o->x) {
            o->x[k] = v[0];
            o->y[k] = v[1];
            o->z[k] = v[2];
            o->vx[k] = v[3];
            o->vy[k] = v[4];
            o->vz[k] = v[5];
        } else {
            o->fx[k - o->base] = (float)v[0];
            o->fy[k - o->base] = (float)v[1];
            o->fz[k - o->base] = (float)v[2];
        }
        p = nl + 1;
    }
}

/* src/io/qpm.c:9-34 */
gal_t *gal_loadqpm_cubic(char *file, double L) {
    double t0 = fftgal_now();
    qpm_file f;
    if (qpm_open(file, &f)) {
        fprintf(stderr, "fftgal_amd: %s\n", fgk_last_error());
        abort();
    }
    const long Np = f.lines;
    gal_t *part = gal_init(Np, L * L * L);
    size_t n = Np > 0 ? (size_t)Np : 1;
    part->vx = (double *)malloc(n * sizeof(double));
    part->vy = (double *)malloc(n * sizeof(double));
    part->vz = (double *)malloc(n * sizeof(double));
    DI_ASSERT(part->vx && part->vy && part->vz, "gal_loadqpm_cubic: out of memory");
    qpm_out o = {f.c, part->x, part->y, part->z, part->vx, part->vy, part->vz, NULL, NULL, NULL, 0};
    const int nt = qpm_threads(f.nc);
    qpm_parallel(nt, 0, f.nc, parse_chunk, &o);
    for (long i = 0; i < f.nc; ++i)
        if (f.c[i].bad >= 0) {
            fprintf(stderr, "gal_loadqpm_cubic: line %ld of %s does not start with six numbers\n",
                    f.c[i].bad + 1, file);
            abort();
        }
    qpm_close(&f);
    if (fftgal_verbose())
        fprintf(stderr, "gal_loadqpm_cubic() %.2fs, loaded %ld galaxies (%d threads)\n",
                fftgal_now() - t0, Np, nt);
    return part;
}

int fftgal_qpm_load_device(const char *file, long Np_max, float *x, float *y, float *z,
                           long *Np_out) {
    if (Np_out) *Np_out = 0;
    if (fgk_init()) return -1;
    qpm_file f;
    if (qpm_open(file, &f)) {
        qpm_close(&f);
        return -1;
    }
    if (f.lines > Np_max) {
        fgk_set_error("qpm: %s holds %ld lines, more than Np_max=%ld", file, f.lines, Np_max);
        qpm_close(&f);
        return -2;
    }
    const int nt = qpm_threads(f.nc);
    const long PL = QPM_PIECE_LINES;
    fgk_stream_t s = NULL;
    float *h[2] = {NULL, NULL};
    fgk_event_t ev[2] = {NULL, NULL};
    int used[2] = {0, 0};
    int rc = fgk_stream_create(&s);
    for (int b = 0; b < 2 && !rc; ++b) {
        rc = fgk_host_alloc((void **)&h[b], (size_t)3 * PL * sizeof(float));
        if (!rc) rc = fgk_event_create(&ev[b]);
    }
    long c0 = 0;
    for (int piece = 0; !rc && c0 < f.nc; ++piece) {
        /* chunks [c0, c1) of at most PL lines */
        long c1 = c0 + 1, n = f.c[c0].lines;
        while (c1 < f.nc && n + f.c[c1].lines <= PL) n += f.c[c1++].lines;
        if (n > PL) {   /* a chunk with more lines than a piece holds */
            fgk_set_error("qpm: a chunk of %s holds %ld lines, more than a staged piece (%ld)",
                          file, n, PL);
            rc = -2;
            break;
        }
        const int b = piece & 1;
        if (used[b]) rc = fgk_event_sync(ev[b]);   /* this buffer's previous copy is done */
        if (rc) break;
        qpm_out o = {f.c, NULL, NULL, NULL, NULL, NULL, NULL, h[b], h[b] + PL, h[b] + 2 * PL,
                     f.c[c0].first};
        qpm_parallel(nt, c0, c1, parse_chunk, &o);
        for (long i = c0; i < c1 && !rc; ++i)
            if (f.c[i].bad >= 0) {
                fgk_set_error("qpm: line %ld of %s does not start with six numbers",
                              f.c[i].bad + 1, file);
                rc = -2;
            }
        if (rc) break;
        const long at = f.c[c0].first;
        const size_t bytes = (size_t)n * sizeof(float);
        rc = fgk_h2d(x + at, h[b], bytes, s);
        if (!rc) rc = fgk_h2d(y + at, h[b] + PL, bytes, s);
        if (!rc) rc = fgk_h2d(z + at, h[b] + 2 * PL, bytes, s);
        if (!rc) rc = fgk_event_record(ev[b], s);
        used[b] = 1;
        c0 = c1;
    }
    if (s) {
        int rs = fgk_stream_sync(s);
        if (!rc) rc = rs;
        fgk_stream_destroy(s);
    }
    for (int b = 0; b < 2; ++b) {
        fgk_host_free(h[b]);
        if (ev[b]) fgk_event_destroy(ev[b]);
    }
    if (!rc && Np_out) *Np_out = f.lines;
    qpm_close(&f);
    return rc;
}
