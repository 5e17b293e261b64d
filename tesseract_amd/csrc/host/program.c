/* Program helpers: the read-only rule, link resolution, cube size (reference src/asm.c:767-820). */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tesseract_host.h"

static int same_name(const char *a, const char *b)
{
    while (*a && *b && tolower((unsigned char)*a) == tolower((unsigned char)*b)) {
        a++;
        b++;
    }
    return *a == 0 && *b == 0;
}

void ts_program_free(ts_program *p)
{
    if (p->cell_names)
        for (uint32_t i = 0; i < p->ncells; i++)
            free(p->cell_names[i]);
    free(p->cell_names);
    free(p->code);
    free(p->inputs);
    free(p->outputs);
    free(p->special);
    memset(p, 0, sizeof *p);
}

int ts_program_find_cell(const ts_program *p, const char *name, uint32_t *addr)
{
    for (uint32_t i = 0; i < p->ncells; i++)
        if (same_name(p->cell_names[i], name)) {
            *addr = i;
            return 1;
        }
    return 0;
}

/* ZERO, ONE, lane-id and link cells ignore writes (reference src/asm.c:779-782). */
int ts_program_readonly(const ts_program *p, uint32_t cell)
{
    if (cell < TS_CELL_FIRST_FREE)
        return 1;
    return p->special != NULL && p->special[cell].kind != TS_CK_PLAIN;
}

/* Follow a chain of links to the cell really read; every link crossed flips its dimension in the
   machine number (reference src/asm.c:793-801). */
uint32_t ts_program_resolve(const ts_program *p, uint32_t cell, unsigned *mask)
{
    unsigned m = 0;
    if (p->special)
        while (p->special[cell].kind == TS_CK_LINK) {
            m ^= 1u << p->special[cell].dim;
            cell = p->special[cell].src;
        }
    *mask = m;
    return cell;
}

/* reference src/asm.c:784-791 */
unsigned ts_program_dims(const ts_program *p)
{
    unsigned dims = TS_CUBE_DIMS;
    if (!p->special)
        return dims;
    for (uint32_t c = 0; c < p->ncells; c++) {
        const ts_cellinfo *ci = &p->special[c];
        if (ci->kind != TS_CK_PLAIN && ci->dim + 1u > dims)
            dims = ci->dim + 1u;
    }
    return dims;
}

/* "ID0".."ID15", case-insensitive, no leading zeros (reference src/asm.c:251-265). */
int ts_program_id_name(const char *name)
{
    if (tolower((unsigned char)name[0]) != 'i' || tolower((unsigned char)name[1]) != 'd')
        return -1;
    const char *d = name + 2;
    if (!isdigit((unsigned char)d[0]))
        return -1;
    if (d[1] == 0)
        return d[0] - '0';
    if (d[0] == '0' || !isdigit((unsigned char)d[1]) || d[2] != 0)
        return -1;
    int v = (d[0] - '0') * 10 + (d[1] - '0');
    return v < TS_CUBE_MAX_DIMS ? v : -1;
}

unsigned ts_addr_bits(uint32_t ncells)
{
    unsigned b = 1;
    while (b < 32 && (1ull << b) < ncells)
        b++;
    return b;
}

static void *dup_bytes(const void *src, size_t n)
{
    void *d = malloc(n ? n : 1);
    if (d && n)
        memcpy(d, src, n);
    return d;
}

int ts_program_copy(ts_program *dst, const ts_program *src)
{
    memset(dst, 0, sizeof *dst);
    dst->len = src->len;
    dst->ncells = src->ncells;
    dst->ninputs = src->ninputs;
    dst->noutputs = src->noutputs;
    dst->code = dup_bytes(src->code, src->len * sizeof *src->code);
    dst->inputs = dup_bytes(src->inputs, src->ninputs * sizeof *src->inputs);
    dst->outputs = dup_bytes(src->outputs, src->noutputs * sizeof *src->outputs);
    if (src->special)
        dst->special = dup_bytes(src->special, src->ncells * sizeof *src->special);
    dst->cell_names = calloc(src->ncells ? src->ncells : 1, sizeof *dst->cell_names);
    if (!dst->code || !dst->inputs || !dst->outputs || !dst->cell_names ||
        (src->special && !dst->special)) {
        ts_program_free(dst);
        return 0;
    }
    for (uint32_t i = 0; i < src->ncells; i++) {
        const char *n = src->cell_names ? src->cell_names[i] : NULL;
        dst->cell_names[i] = n ? dup_bytes(n, strlen(n) + 1) : dup_bytes("?", 2);
        if (!dst->cell_names[i]) {
            ts_program_free(dst);
            return 0;
        }
    }
    return 1;
}

int ts_program_from_arrays(ts_program *out, const uint32_t *code, size_t len, uint32_t ncells,
                           const char *names, size_t names_bytes, const uint32_t *inputs,
                           size_t ninputs, const uint32_t *outputs, size_t noutputs,
                           const uint8_t *kind, const uint8_t *dim, const uint32_t *src, char *err,
                           size_t errlen)
{
    memset(out, 0, sizeof *out);
    if (ncells < TS_CELL_FIRST_FREE) {
        snprintf(err, errlen, "a program has at least the cells ZERO and ONE");
        return 0;
    }
    for (size_t i = 0; i < len; i++)
        if ((code[i] & 0x7FFFFFFFu) >= ncells) {
            snprintf(err, errlen, "instruction %zu names cell %u of %u", i, code[i] & 0x7FFFFFFFu,
                     ncells);
            return 0;
        }
    for (size_t i = 0; i < ninputs + noutputs; i++) {
        uint32_t c = i < ninputs ? inputs[i] : outputs[i - ninputs];
        if (c >= ncells) {
            snprintf(err, errlen, "input or output %zu names cell %u of %u", i, c, ncells);
            return 0;
        }
    }
    out->code = malloc((len ? len : 1) * sizeof *out->code);
    out->cell_names = calloc(ncells, sizeof *out->cell_names);
    out->inputs = dup_bytes(inputs, ninputs * sizeof *inputs);
    out->outputs = dup_bytes(outputs, noutputs * sizeof *outputs);
    if (kind)
        out->special = calloc(ncells, sizeof *out->special);
    if (!out->code || !out->cell_names || !out->inputs || !out->outputs || (kind && !out->special)) {
        ts_program_free(out);
        snprintf(err, errlen, "out of memory");
        return 0;
    }
    out->ncells = ncells;
    out->len = len;
    out->ninputs = ninputs;
    out->noutputs = noutputs;
    for (size_t i = 0; i < len; i++) {
        out->code[i].op = (uint8_t)(code[i] >> 31);
        out->code[i].addr = code[i] & 0x7FFFFFFFu;
    }
    size_t at = 0;
    for (uint32_t c = 0; c < ncells; c++) {
        size_t n = at < names_bytes ? strnlen(names + at, names_bytes - at) : 0;
        if (at + n >= names_bytes) {
            ts_program_free(out);
            snprintf(err, errlen, "the names end before cell %u", c);
            return 0;
        }
        out->cell_names[c] = dup_bytes(names + at, n + 1);
        if (!out->cell_names[c]) {
            ts_program_free(out);
            snprintf(err, errlen, "out of memory");
            return 0;
        }
        at += n + 1;
    }
    for (uint32_t c = 0; kind && c < ncells; c++) {
        out->special[c].kind = kind[c];
        out->special[c].dim = dim[c];
        out->special[c].src = src[c];
        int bad = kind[c] > TS_CK_LINK || (kind[c] != TS_CK_PLAIN && dim[c] >= TS_CUBE_MAX_DIMS) ||
                  (kind[c] == TS_CK_LINK && src[c] >= c);  /* a link reads an older cell */
        if (bad) {
            ts_program_free(out);
            snprintf(err, errlen, "cell %u is a malformed id or link cell", c);
            return 0;
        }
    }
    return 1;
}
