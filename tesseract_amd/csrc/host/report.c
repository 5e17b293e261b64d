/* Run summaries in the reference's raw, machine-readable forms, and their merge: tester 1 per-size
   results (reference src/tester.c:228-240,355-431) and enumeration shares (reference
   src/work.c:341-371,435-441). A GPU run and CPU runs of other shares write the same lines, and one
   merge adds them up. */
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "tesseract_host.h"

size_t ts_t1_program_len(const ts_t1_config *cfg, uint32_t cells)
{
    size_t len = (size_t)(cells * cfg->ratio + 0.5);
    return len ? len : 1;
}

void ts_t1_add_size(ts_t1_size *to, const ts_t1_size *from)
{
    to->programs += from->programs;
    to->starts += from->starts;
    to->passes += from->passes;
    to->unfinished += from->unfinished;
    for (int b = 0; b < TS_T1_BUCKETS; b++)
        to->hist[b] += from->hist[b];
    if (from->longest > to->longest) {
        to->longest = from->longest;
        to->longest_seed = from->longest_seed;
    }
}

/* a growing text buffer; ok turns 0 once an allocation fails */
typedef struct {
    char  *p;
    size_t n, cap;
    int    ok;
} text;

static void put(text *t, const char *fmt, ...)
{
    if (!t->ok)
        return;
    for (;;) {
        va_list ap;
        va_start(ap, fmt);
        int k = vsnprintf(t->p + t->n, t->cap - t->n, fmt, ap);
        va_end(ap);
        if (k < 0) {
            t->ok = 0;
            return;
        }
        if ((size_t)k < t->cap - t->n) {
            t->n += (size_t)k;
            return;
        }
        size_t cap = 2 * t->cap + (size_t)k + 1;
        char *q = realloc(t->p, cap);
        if (!q) {
            t->ok = 0;
            return;
        }
        t->p = q;
        t->cap = cap;
    }
}

static char *finish(text *t)
{
    if (t->ok)
        return t->p;
    free(t->p);
    return NULL;
}

static text start(void)
{
    text t = {malloc(256), 0, 256, 1};
    if (!t.p)
        t.ok = 0;
    else
        t.p[0] = 0;
    return t;
}

char *ts_t1_raw(const ts_t1_size *s, int n, const char *host, double seconds, int threads,
                uint64_t max_passes)
{
    text t = start();
    put(&t, "T1RUN %s %d %.6f %llu\n", host, threads, seconds, (unsigned long long)max_passes);
    for (int i = 0; i < n; i++) {
        const ts_t1_size *x = &s[i];
        put(&t, "T1 %u %zu %llu %llu %llu %llu %llu %llx", x->cells, x->len,
            (unsigned long long)x->programs, (unsigned long long)x->starts,
            (unsigned long long)x->passes, (unsigned long long)x->unfinished,
            (unsigned long long)x->longest, (unsigned long long)x->longest_seed);
        for (int b = 0; b < TS_T1_BUCKETS; b++)
            put(&t, " %llu", (unsigned long long)x->hist[b]);
        put(&t, "\n");
    }
    return finish(&t);
}

int ts_t1_merge(char **files, int nfiles, ts_t1_size *s, double *seconds, int *threads,
                uint64_t *max_passes)
{
    int n = 0;
    *seconds = 0;
    *threads = 0;
    *max_passes = 0;
    char line[4096];
    for (int f = 0; f < nfiles; f++) {
        FILE *in = fopen(files[f], "r");
        if (!in)
            return -1;
        while (fgets(line, sizeof line, in)) {
            char host[128];
            int th;
            double sec;
            unsigned long long mp;
            if (sscanf(line, "T1RUN %127s %d %lf %llu", host, &th, &sec, &mp) == 4) {
                *threads += th;
                if (sec > *seconds)
                    *seconds = sec;
                *max_passes = mp;
                continue;
            }
            ts_t1_size x;
            memset(&x, 0, sizeof x);
            unsigned long long v[7];
            int used = 0;
            if (sscanf(line, "T1 %u %zu %llu %llu %llu %llu %llu %llx%n", &x.cells, &x.len, &v[0],
                       &v[1], &v[2], &v[3], &v[4], &v[5], &used) != 8)
                continue;
            x.programs = v[0];
            x.starts = v[1];
            x.passes = v[2];
            x.unfinished = v[3];
            x.longest = v[4];
            x.longest_seed = v[5];
            const char *p = line + used;
            for (int b = 0; b < TS_T1_BUCKETS; b++) {
                char *end;
                x.hist[b] = strtoull(p, &end, 10);
                p = end;
            }
            int i = 0;
            while (i < n && !(s[i].cells == x.cells && s[i].len == x.len))
                i++;
            if (i == n) {
                if (n == TS_T1_SIZES)
                    continue;  /* more sizes than a summary holds: dropped, as the reference does */
                memset(&s[n], 0, sizeof s[n]);
                s[n].cells = x.cells;
                s[n].len = x.len;
                n++;
            }
            ts_t1_add_size(&s[i], &x);
        }
        fclose(in);
    }
    return n;
}

char *ts_enum_raw(int mode, const ts_enum_result *r)
{
    text t = start();
    put(&t, "EN %d %llu %llu %llu %.6f %d %llu\n", mode, (unsigned long long)r->space,
        (unsigned long long)r->tried, (unsigned long long)r->hits, r->seconds, r->threads,
        (unsigned long long)r->ops);
    for (int h = 0; h < r->nhit; h++)
        put(&t, "ENHIT %llx\n", (unsigned long long)r->hit[h]);
    return finish(&t);
}

int ts_enum_merge(char **files, int nfiles, ts_enum_result *r, char *err, size_t errlen)
{
    memset(r, 0, sizeof *r);
    char line[512];
    for (int f = 0; f < nfiles; f++) {
        FILE *in = fopen(files[f], "r");
        if (!in) {
            snprintf(err, errlen, "cannot open %s", files[f]);
            return 0;
        }
        while (fgets(line, sizeof line, in)) {
            unsigned long long space, tried, hits, ops, x;
            double sec;
            int mode, threads;
            if (sscanf(line, "EN %d %llu %llu %llu %lf %d %llu", &mode, &space, &tried, &hits, &sec,
                       &threads, &ops) == 7) {
                r->space = space;
                r->tried += tried;
                r->hits += hits;
                r->ops += ops;
                r->threads += threads;
                if (sec > r->seconds)
                    r->seconds = sec;
            } else if (sscanf(line, "ENHIT %llx", &x) == 1) {
                ts_enum_add_hit(r, x);
            }
        }
        fclose(in);
    }
    return 1;
}
