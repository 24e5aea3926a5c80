/* Tuning cache file (see tune.h). */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "tune.h"
#include "../../../include/fftgal_batch.h"

static void defaults(int Ng, fftgal_tuning *t) {
    t->paint_tile = Ng >= 16 ? 16 : Ng;
    t->fft_cols = 0;
    t->fft_rows = 0;
    t->fft_generic = 0;
    t->bin_bk = 0;
}

/* the interlace field of a line: 0, 1, or the FP64 drop-in grid's marker */
static int mode_of(int interlace) {
    return interlace == FFTGAL_TUNE_FP64_GRID ? FFTGAL_TUNE_FP64_GRID : (interlace ? 1 : 0);
}

static int usable(const char *path) {
    return path && *path && strcmp(path, "FFTW_ESTIMATE") != 0;
}

int fftgal_tune_lookup(const char *path, int Ng, int interlace, fftgal_tuning *t) {
    defaults(Ng, t);
    if (!usable(path)) return 0;
    FILE *fp = fopen(path, "r");
    if (!fp) return 0;
    char line[256];
    int found = 0;
    while (fgets(line, sizeof line, fp)) {
        int n, il, a, b, c, g = 0, k = 0;
        if (line[0] == '#') continue;
        if (sscanf(line, "%d %d %d %d %d %d %d", &n, &il, &a, &b, &c, &g, &k) >= 5 && n == Ng &&
            il == mode_of(interlace)) {
            t->paint_tile = a;
            t->fft_cols = b;
            t->fft_rows = c;
            t->fft_generic = g;
            t->bin_bk = k;
            found = 1;
        }
    }
    fclose(fp);
    return found;
}

int fftgal_tune_store(const char *path, int Ng, int interlace, const fftgal_tuning *t) {
    if (!usable(path)) return 0;
    char **keep = NULL;
    size_t nkeep = 0;
    FILE *fp = fopen(path, "r");
    if (fp) {
        char line[256];
        while (fgets(line, sizeof line, fp)) {
            int n, il;
            if (line[0] != '#' && sscanf(line, "%d %d", &n, &il) == 2 && n == Ng &&
                il == mode_of(interlace))
                continue;
            char **k2 = (char **)realloc(keep, (nkeep + 1) * sizeof(char *));
            if (!k2) break;
            keep = k2;
            keep[nkeep++] = strdup(line);
        }
        fclose(fp);
    }
    fp = fopen(path, "w");
    int rc = fp ? 0 : -1;
    if (fp) {
        if (nkeep == 0)
            fprintf(fp, "# Ng interlace paint_tile fft_cols fft_rows fft_generic bin_bk\n");
        for (size_t i = 0; i < nkeep; ++i) fputs(keep[i], fp);
        fprintf(fp, "%d %d %d %d %d %d %d\n", Ng, mode_of(interlace), t->paint_tile, t->fft_cols,
                t->fft_rows, t->fft_generic, t->bin_bk);
        fclose(fp);
    }
    for (size_t i = 0; i < nkeep; ++i) free(keep[i]);
    free(keep);
    return rc;
}
