/* Gate builder: logic gates -> NAND/STO runs (reference src/build.c).

   Every gate leaves its result in a cell. The instruction sequences are the shortest known ones
   the reference publishes (AND 6, OR 6, XOR 8, half adder 11, full adder 19, mux 8); they are
   reproduced here instruction for instruction, and cells are numbered in order of first use, so a
   generator produces the same program as the reference's generator of the same name. A few
   scratch cells (__t0..__t5) are shared by all gates. Registers (b_reg) latch at the end of the
   pass; a register fed by another register is staged first so a chain shifts by one place. */
#include <ctype.h>
#include <stdlib.h>
#include <string.h>

#include "tesseract_host.h"

struct ts_builder {
    ts_instr    *code;
    size_t       len, cap;
    char       **names;
    ts_cellinfo *info;
    uint32_t     ncells, capcells;
    uint32_t    *in, *out, *reg_q, *reg_d;
    size_t       nin, capin, nout, capout, nreg, capreg;
    uint32_t    *table;  /* open addressing: cell + 1, 0 = free */
    size_t       tsize;
    uint32_t     scratch[6];
    int          special;
};

static void *grow(void *p, size_t *cap, size_t need, size_t elem)
{
    if (need <= *cap)
        return p;
    size_t c = *cap ? *cap : 32;
    while (c < need)
        c *= 2;
    void *q = realloc(p, c * elem);
    if (!q) {
        fputs("out of memory\n", stderr);
        abort();
    }
    *cap = c;
    return q;
}

static uint64_t name_hash(const char *s)
{
    uint64_t h = 0xcbf29ce484222325ull;
    for (; *s; s++)
        h = (h ^ (uint64_t)tolower((unsigned char)*s)) * 0x100000001b3ull;
    return h;
}

static int name_eq(const char *a, const char *b)
{
    while (*a && *b && tolower((unsigned char)*a) == tolower((unsigned char)*b)) {
        a++;
        b++;
    }
    return *a == 0 && *b == 0;
}

static size_t slot_of(const ts_builder *b, const char *name)
{
    size_t mask = b->tsize - 1, i = (size_t)name_hash(name) & mask;
    while (b->table[i] && !name_eq(b->names[b->table[i] - 1], name))
        i = (i + 1) & mask;
    return i;
}

static void rehash(ts_builder *b)
{
    free(b->table);
    b->tsize = b->tsize ? b->tsize * 2 : 1024;
    b->table = calloc(b->tsize, sizeof *b->table);
    if (!b->table) {
        fputs("out of memory\n", stderr);
        abort();
    }
    for (uint32_t c = 0; c < b->ncells; c++)
        b->table[slot_of(b, b->names[c])] = c + 1;
}

uint32_t ts_b_cell(ts_builder *b, const char *name)
{
    size_t s = slot_of(b, name);
    if (b->table[s])
        return b->table[s] - 1;
    size_t cap = b->capcells;
    b->names = grow(b->names, &cap, b->ncells + 1, sizeof *b->names);
    cap = b->capcells;
    b->info = grow(b->info, &cap, b->ncells + 1, sizeof *b->info);
    b->capcells = (uint32_t)cap;
    size_t n = strlen(name) + 1;
    b->names[b->ncells] = malloc(n);
    memcpy(b->names[b->ncells], name, n);
    ts_cellinfo *ci = &b->info[b->ncells];
    memset(ci, 0, sizeof *ci);
    int d = ts_program_id_name(name);
    if (d >= 0) {
        ci->kind = TS_CK_ID;
        ci->dim = (uint8_t)d;
        b->special = 1;
    }
    b->table[s] = ++b->ncells;
    if (2 * (size_t)b->ncells > b->tsize)
        rehash(b);
    return b->ncells - 1;
}

int ts_b_has(const ts_builder *b, const char *name)
{
    return b->table[slot_of(b, name)] != 0;
}

uint32_t ts_b_link(ts_builder *b, const char *name, uint32_t src, int dim)
{
    uint32_t c = ts_b_cell(b, name);
    b->info[c].kind = TS_CK_LINK;
    b->info[c].dim = (uint8_t)dim;
    b->info[c].src = src;
    b->special = 1;
    return c;
}

ts_builder *ts_b_new(void)
{
    ts_builder *b = calloc(1, sizeof *b);
    if (!b)
        return NULL;
    rehash(b);
    ts_b_cell(b, "ZERO");
    ts_b_cell(b, "ONE");
    return b;
}

static uint32_t scratch(ts_builder *b, int i)
{
    static const char *const names[6] = {"__t0", "__t1", "__t2", "__t3", "__t4", "__t5"};
    if (!b->scratch[i])
        b->scratch[i] = ts_b_cell(b, names[i]);
    return b->scratch[i];
}

void ts_b_input(ts_builder *b, uint32_t cell)
{
    b->in = grow(b->in, &b->capin, b->nin + 1, sizeof *b->in);
    b->in[b->nin++] = cell;
}

void ts_b_output(ts_builder *b, uint32_t cell)
{
    b->out = grow(b->out, &b->capout, b->nout + 1, sizeof *b->out);
    b->out[b->nout++] = cell;
}

void ts_b_reg(ts_builder *b, uint32_t q, uint32_t d)
{
    size_t cap = b->capreg;
    b->reg_q = grow(b->reg_q, &cap, b->nreg + 1, sizeof *b->reg_q);
    cap = b->capreg;
    b->reg_d = grow(b->reg_d, &cap, b->nreg + 1, sizeof *b->reg_d);
    b->capreg = cap;
    b->reg_q[b->nreg] = q;
    b->reg_d[b->nreg++] = d;
}

static void emit(ts_builder *b, int op, uint32_t cell)
{
    b->code = grow(b->code, &b->cap, b->len + 1, sizeof *b->code);
    b->code[b->len].op = (uint8_t)op;
    b->code[b->len].addr = cell;
    b->len++;
}

void ts_b_nand(ts_builder *b, uint32_t c) { emit(b, TS_OP_NAND, c); }
void ts_b_sto(ts_builder *b, uint32_t c) { emit(b, TS_OP_STO, c); }

#define NAND(c) emit(b, TS_OP_NAND, (c))
#define STO(c)  emit(b, TS_OP_STO, (c))

/* RR = c in 3, RR = !c in 2 */
static void load(ts_builder *b, uint32_t c) { NAND(TS_CELL_ZERO); NAND(c); NAND(TS_CELL_ONE); }
static void load_not(ts_builder *b, uint32_t c) { NAND(TS_CELL_ZERO); NAND(c); }

/* RR = x ^ y in 7 (reference src/build.c:168-179), scratch 0 */
static void xor_into_rr(ts_builder *b, uint32_t x, uint32_t y)
{
    uint32_t t = scratch(b, 0);
    NAND(TS_CELL_ZERO);
    NAND(x);
    NAND(y);
    STO(t);
    NAND(y);
    NAND(x);
    NAND(t);
}

/* reference src/build.c:181-233 */
void ts_b_gate(ts_builder *b, int type, uint32_t out, const uint32_t *in, int n)
{
    switch (type) {
    case TS_G_BUF:
        load(b, in[0]);
        break;
    case TS_G_NOT:
        load_not(b, in[0]);
        break;
    case TS_G_AND:
    case TS_G_NAND:
        load(b, in[0]);
        for (int i = 1; i < n; i++) {
            NAND(in[i]);
            if (i + 1 < n)
                NAND(TS_CELL_ONE);
        }
        /* after the chain RR is the NAND of all inputs (one input: the input itself) */
        if ((n == 1) == (type == TS_G_NAND))
            NAND(TS_CELL_ONE);
        break;
    case TS_G_OR:
    case TS_G_NOR: {
        uint32_t t = scratch(b, 0);
        load_not(b, in[0]);
        for (int i = 1; i < n; i++) {
            if (i > 1)
                NAND(TS_CELL_ONE);
            STO(t);
            NAND(in[i]);
            NAND(t);
        }
        if ((n == 1) == (type == TS_G_OR))
            NAND(TS_CELL_ONE);
        break;
    }
    case TS_G_XOR:
    case TS_G_XNOR:
        if (n == 1) {
            load(b, in[0]);
        } else {
            uint32_t acc = scratch(b, 2);
            xor_into_rr(b, in[0], in[1]);
            for (int i = 2; i < n; i++) {
                STO(acc);
                xor_into_rr(b, acc, in[i]);
            }
        }
        if (type == TS_G_XNOR)
            NAND(TS_CELL_ONE);
        break;
    default:
        abort();
    }
    STO(out);
}

void ts_b_gate2(ts_builder *b, int type, uint32_t out, uint32_t x, uint32_t y)
{
    uint32_t in[2] = {x, y};
    ts_b_gate(b, type, out, in, 2);
}

/* out = s ? a1 : a0 given ns = !s, 8 instructions (reference src/build.c:236-248) */
void ts_b_mux(ts_builder *b, uint32_t out, uint32_t s, uint32_t ns, uint32_t a1, uint32_t a0)
{
    uint32_t t = scratch(b, 0);
    NAND(TS_CELL_ZERO);
    NAND(s);
    NAND(a0);
    STO(t);
    NAND(ns);
    NAND(a1);
    NAND(t);
    STO(out);
}

/* s = x ^ y, k = x & y, 11 instructions (reference src/build.c:251-266) */
void ts_b_half_adder(ts_builder *b, uint32_t s, uint32_t k, uint32_t x, uint32_t y)
{
    uint32_t t = scratch(b, 0);
    NAND(TS_CELL_ZERO);
    NAND(x);
    NAND(y);
    STO(t);
    NAND(y);
    NAND(TS_CELL_ONE);
    STO(k);
    NAND(TS_CELL_ONE);
    NAND(x);
    NAND(t);
    STO(s);
}

/* s = x ^ y ^ c, k = carry, 19 instructions (reference src/build.c:269-291) */
void ts_b_full_adder(ts_builder *b, uint32_t s, uint32_t k, uint32_t x, uint32_t y, uint32_t c)
{
    uint32_t nx = scratch(b, 0), nxy = scratch(b, 1), o = scratch(b, 2), q = scratch(b, 3);
    NAND(TS_CELL_ZERO);
    NAND(x);
    STO(nx);           /* !x */
    NAND(TS_CELL_ONE);
    NAND(y);
    STO(nxy);          /* !(x & y) */
    NAND(y);
    NAND(nx);
    STO(o);            /* x | y */
    NAND(nxy);
    NAND(c);
    STO(q);            /* (x ^ y) | !c */
    NAND(c);
    NAND(nxy);
    STO(k);            /* (x ^ y) & c | x & y */
    NAND(TS_CELL_ONE);
    NAND(o);
    NAND(q);
    STO(s);            /* x ^ y ^ c */
}

/* out = OR of sel[i] & val[i] (reference src/build.c:293-318) */
void ts_b_select(ts_builder *b, uint32_t out, const uint32_t *sel, const uint32_t *val, int n)
{
    uint32_t acc = scratch(b, 0);
    int terms = 0;
    for (int i = 0; i < n; i++) {
        if (val[i] == TS_CELL_ZERO)
            continue;
        if (val[i] == TS_CELL_ONE) {
            load_not(b, sel[i]);
        } else {
            load(b, sel[i]);
            NAND(val[i]);
        }
        if (terms++) {
            NAND(acc);
            NAND(TS_CELL_ONE);
        }
        STO(acc);
    }
    if (terms)
        load_not(b, acc);
    else
        load(b, TS_CELL_ZERO);
    STO(out);
}

size_t ts_b_len(const ts_builder *b)
{
    return b->len;
}

/* reference src/build.c:349-399 */
void ts_b_finish(ts_builder *b, ts_program *p)
{
    unsigned char *is_q = calloc((size_t)b->ncells + 1, 1);
    uint32_t *stage = calloc(b->nreg + 1, sizeof *stage);
    for (size_t r = 0; r < b->nreg; r++)
        is_q[b->reg_q[r]] = 1;
    for (size_t r = 0; r < b->nreg; r++)
        if (is_q[b->reg_d[r]]) {
            char name[32];
            snprintf(name, sizeof name, "__r%zu", r);
            stage[r] = ts_b_cell(b, name);
            load(b, b->reg_d[r]);
            STO(stage[r]);
        }
    for (size_t r = 0; r < b->nreg; r++)
        if (!stage[r]) {
            load(b, b->reg_d[r]);
            STO(b->reg_q[r]);
        }
    for (size_t r = 0; r < b->nreg; r++)
        if (stage[r]) {
            load(b, stage[r]);
            STO(b->reg_q[r]);
        }
    free(stage);
    free(is_q);

    memset(p, 0, sizeof *p);
    p->code = b->code;
    p->len = b->len;
    p->ncells = b->ncells;
    p->cell_names = b->names;
    p->inputs = b->in;
    p->ninputs = b->nin;
    p->outputs = b->out;
    p->noutputs = b->nout;
    if (b->special)
        p->special = b->info;
    else
        free(b->info);
    free(b->reg_q);
    free(b->reg_d);
    free(b->table);
    free(b);
}

static const char *const gate_names[TS_G_COUNT] = {"BUF", "NOT", "AND", "NAND",
                                                   "OR",  "NOR", "XOR", "XNOR"};

const char *ts_gate_name(int type)
{
    return type >= 0 && type < TS_G_COUNT ? gate_names[type] : "?";
}

int ts_gate_type(const char *name)
{
    if (name_eq(name, "BUFF"))
        return TS_G_BUF;
    for (int t = 0; t < TS_G_COUNT; t++)
        if (name_eq(name, gate_names[t]))
            return t;
    return -1;
}
