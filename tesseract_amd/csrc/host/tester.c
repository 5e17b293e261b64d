/* Random programs, their seeds and start states for the cycle tester (reference src/tester.c:21-91),
   and the small enumeration helpers (reference src/enum.c:31-77). */
#include <ctype.h>
#include <stdlib.h>
#include <string.h>

#include "tesseract_host.h"

/* Vigna's splitmix64: add the golden-ratio increment, then the variant-13 finaliser. */
uint64_t ts_splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* reference src/tester.c:51-55 */
uint64_t ts_t1_program_seed(uint64_t seed, uint32_t cells, uint64_t k)
{
    uint64_t s = seed ^ ((uint64_t)cells << 40) ^ (k * 0xD1B54A32D192ED03ull);
    return ts_splitmix64(&s);
}

/* One splitmix64 draw per instruction: bit 0 is the opcode, the rest picks the cell; a NAND may
   read any cell, a STO writes a writable one (reference src/tester.c:57-81). */
void ts_random_program(ts_program *p, uint32_t cells, size_t len, uint64_t seed)
{
    memset(p, 0, sizeof *p);
    p->ncells = cells + TS_CELL_FIRST_FREE;
    p->len = len;
    p->code = malloc((len ? len : 1) * sizeof *p->code);
    p->cell_names = malloc(p->ncells * sizeof *p->cell_names);
    p->inputs = malloc(sizeof *p->inputs);
    p->outputs = malloc(sizeof *p->outputs);
    for (uint32_t a = 0; a < p->ncells; a++) {
        p->cell_names[a] = malloc(16);
        if (a == TS_CELL_ZERO)
            strcpy(p->cell_names[a], "ZERO");
        else if (a == TS_CELL_ONE)
            strcpy(p->cell_names[a], "ONE");
        else
            snprintf(p->cell_names[a], 16, "c%u", a - TS_CELL_FIRST_FREE);
    }
    uint64_t s = seed;
    for (size_t i = 0; i < len; i++) {
        uint64_t r = ts_splitmix64(&s);
        p->code[i].op = (uint8_t)(r & 1);
        p->code[i].addr = p->code[i].op == TS_OP_NAND
                              ? (uint32_t)((r >> 1) % p->ncells)
                              : TS_CELL_FIRST_FREE + (uint32_t)((r >> 1) % cells);
    }
}

/* Group 0 is exactly the reference's t1_start (src/tester.c:83-91): stream seed ^ 0x5DEECE66D,
   8 words per writable cell in cell order, then 8 words of RR. Further groups continue the same
   stream in the same order. */
void ts_t1_start(const ts_program *p, uint64_t seed, unsigned groups, ts_line *mem, ts_line *rr)
{
    uint64_t s = seed ^ 0x5DEECE66Dull;
    for (unsigned g = 0; g < groups; g++) {
        ts_line *m = mem + (size_t)g * p->ncells;
        for (uint32_t a = TS_CELL_FIRST_FREE; a < p->ncells; a++)
            for (int w = 0; w < 8; w++)
                m[a].w[w] = ts_splitmix64(&s);
        for (int w = 0; w < 8; w++)
            rr[g].w[w] = ts_splitmix64(&s);
    }
}

/* bucket b holds cycle lengths in (2^(b-1), 2^b] (reference src/tester.c:42-48) */
int ts_t1_bucket(uint64_t lam)
{
    int b = 0;
    while (b < TS_T1_BUCKETS - 2 && (1ull << b) < lam)
        b++;
    return b;
}

/* keeps the TS_EN_SHOW smallest, ascending (reference src/enum.c:31-46) */
void ts_enum_add_hit(ts_enum_result *r, uint64_t x)
{
    int i;
    if (r->nhit < TS_EN_SHOW) {
        i = r->nhit++;
    } else {
        if (x >= r->hit[TS_EN_SHOW - 1])
            return;
        i = TS_EN_SHOW - 1;
    }
    for (; i > 0 && r->hit[i - 1] > x; i--)
        r->hit[i] = r->hit[i - 1];
    r->hit[i] = x;
}

static int eq_nocase(const char *a, const char *b)
{
    while (*a && *b && tolower((unsigned char)*a) == tolower((unsigned char)*b)) {
        a++;
        b++;
    }
    return *a == 0 && *b == 0;
}

static int match(const ts_program *a, const uint32_t *la, size_t na, const ts_program *b,
                 const uint32_t *lb, size_t nb, int *map, int inputs, char *err, size_t errlen)
{
    if (na != nb) {
        snprintf(err, errlen, "the programs have %zu and %zu %s", na, nb,
                 inputs ? "inputs" : "outputs");
        return 0;
    }
    for (size_t i = 0; i < na; i++) {
        map[i] = -1;
        for (size_t j = 0; j < nb; j++)  /* the last match wins, as in the reference */
            if (eq_nocase(a->cell_names[la[i]], b->cell_names[lb[j]]))
                map[i] = (int)j;
        if (map[i] < 0) {
            snprintf(err, errlen, "the second program has no %s called %s",
                     inputs ? "input" : "output", a->cell_names[la[i]]);
            return 0;
        }
    }
    return 1;
}

int ts_enum_match(const ts_program *a, const ts_program *b, int *in_map, int *out_map, char *err,
                  size_t errlen)
{
    return match(a, a->inputs, a->ninputs, b, b->inputs, b->ninputs, in_map, 1, err, errlen) &&
           match(a, a->outputs, a->noutputs, b, b->outputs, b->noutputs, out_map, 0, err, errlen);
}
