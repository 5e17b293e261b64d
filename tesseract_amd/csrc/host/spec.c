/* Verify/search specifications (reference src/spec.h, src/spec.c).

   "c:s = a + b + c": left of '=' the output vectors, joined, the first giving the highest bits
   (at most 64 in all); right of '=' an expression over input vectors and numbers with
   ~ - (unary), * , + -, << >>, &, ^, | (in rising precedence order: | ^ & shifts +- *), evaluated
   on 64-bit words and cut to the width of the left side. Shifts by 64 or more give 0. A vector is
   the cell with exactly that name or the cells name0, name1, ... (bit i in name<i>).

   The expression is kept as an RPN program so engines can lower it (the GPU engine turns it into
   per-lane straight-line code). */
#include <ctype.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "tesseract_host.h"

enum { MAX_VARS = 32, MAX_RPN = 256, MAX_STACK = 64 };

struct ts_spec {
    int          width;
    uint32_t     out[64];
    int          nvars;
    char         var_name[MAX_VARS][64];
    int          var_bits[MAX_VARS];
    int          var_pos[MAX_VARS][64];
    ts_spec_item rpn[MAX_RPN];
    int          nrpn, depth, max_depth;
};

typedef struct {
    const char       *at;
    const ts_program *p;
    ts_spec          *s;
    int               inputs_allowed;
    char             *err;
    size_t            errlen;
    int               failed;
} parser;

static int eq_n(const char *a, const char *b, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (tolower((unsigned char)a[i]) != tolower((unsigned char)b[i]))
            return 0;
    return 1;
}

/* reference src/spec.c:49-82: exact name first, else name<digits> without leading zeros */
int ts_vec_resolve(const ts_program *p, const uint32_t *list, size_t n, const char *name, int *idx)
{
    size_t len = strlen(name);
    for (size_t i = 0; i < n; i++) {
        const char *c = p->cell_names[list[i]];
        if (strlen(c) == len && eq_n(c, name, len)) {
            idx[0] = (int)i;
            return 1;
        }
    }
    int width = 0;
    uint64_t seen = 0;
    for (size_t i = 0; i < n; i++) {
        const char *c = p->cell_names[list[i]];
        if (strlen(c) <= len || !eq_n(c, name, len))
            continue;
        const char *d = c + len;
        if (!isdigit((unsigned char)d[0]) || (d[0] == '0' && d[1]))
            continue;
        char *end;
        unsigned long bit = strtoul(d, &end, 10);
        if (*end)
            continue;
        if (bit >= 64)
            return -1;
        idx[bit] = (int)i;
        seen |= 1ull << bit;
        if ((int)bit + 1 > width)
            width = (int)bit + 1;
    }
    if (width && seen != (width == 64 ? ~0ull : (1ull << width) - 1))
        return -1;
    return width;
}

static void fail(parser *ps, const char *fmt, ...)
{
    if (ps->failed)
        return;
    ps->failed = 1;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ps->err, ps->errlen, fmt, ap);
    va_end(ap);
}

static void blanks(parser *ps)
{
    while (*ps->at == ' ' || *ps->at == '\t')
        ps->at++;
}

static void emit(parser *ps, int op, uint64_t v)
{
    ts_spec *s = ps->s;
    if (s->nrpn == MAX_RPN) {
        fail(ps, "the expression is too long");
        return;
    }
    s->rpn[s->nrpn].op = op;
    s->rpn[s->nrpn++].v = v;
    if (op == TS_T_NUM || op == TS_T_VAR)
        s->depth++;
    else if (op != TS_T_NOT && op != TS_T_NEG)
        s->depth--;
    if (s->depth > s->max_depth)
        s->max_depth = s->depth;
    if (s->max_depth > MAX_STACK)
        fail(ps, "the expression is nested too deeply");
}

static int identifier(parser *ps, char *buf, size_t len)
{
    if (!isalpha((unsigned char)*ps->at) && *ps->at != '_')
        return 0;
    size_t n = 0;
    for (; isalnum((unsigned char)*ps->at) || *ps->at == '_'; ps->at++)
        if (n + 1 < len)
            buf[n++] = *ps->at;
    buf[n] = 0;
    return 1;
}

static int variable(parser *ps, const char *name)
{
    ts_spec *s = ps->s;
    size_t len = strlen(name);
    for (int k = 0; k < s->nvars; k++)
        if (strlen(s->var_name[k]) == len && eq_n(s->var_name[k], name, len))
            return k;
    if (!ps->inputs_allowed) {
        fail(ps, "'%s': the right side must be a number here", name);
        return -1;
    }
    if (s->nvars == MAX_VARS) {
        fail(ps, "too many input vectors");
        return -1;
    }
    int idx[64];
    int w = ts_vec_resolve(ps->p, ps->p->inputs, ps->p->ninputs, name, idx);
    if (w <= 0) {
        if (w < 0)
            fail(ps, "input vector %s has missing bits or more than 64", name);
        else
            fail(ps, "no input called %s", name);
        return -1;
    }
    int k = s->nvars++;
    snprintf(s->var_name[k], sizeof s->var_name[k], "%s", name);
    s->var_bits[k] = w;
    for (int b = 0; b < w; b++)  /* input i is bit n-1-i of the assignment */
        s->var_pos[k][b] = (int)ps->p->ninputs - 1 - idx[b];
    return k;
}

static void expression(parser *ps, int min_prec);

static void operand(parser *ps)
{
    char name[64];
    blanks(ps);
    if (*ps->at == '(') {
        ps->at++;
        expression(ps, 0);
        blanks(ps);
        if (*ps->at == ')')
            ps->at++;
        else
            fail(ps, "expected )");
    } else if (isdigit((unsigned char)*ps->at)) {
        char *end;
        uint64_t v = strtoull(ps->at, &end, 0);
        ps->at = end;
        emit(ps, TS_T_NUM, v);
    } else if (identifier(ps, name, sizeof name)) {
        int k = variable(ps, name);
        if (k >= 0)
            emit(ps, TS_T_VAR, (uint64_t)k);
    } else if (*ps->at) {
        fail(ps, "unexpected '%c'", *ps->at);
    } else {
        fail(ps, "the expression ends too early");
    }
}

static void prefixed(parser *ps)
{
    blanks(ps);
    if (*ps->at == '~' || *ps->at == '-') {
        int op = *ps->at == '~' ? TS_T_NOT : TS_T_NEG;
        ps->at++;
        prefixed(ps);
        emit(ps, op, 0);
    } else {
        operand(ps);
    }
}

/* the binary operator at s, its precedence and length; 0 if there is none */
static int binary(const char *s, int *prec, int *len)
{
    *len = 1;
    switch (s[0]) {
    case '|': *prec = 1; return TS_T_OR;
    case '^': *prec = 2; return TS_T_XOR;
    case '&': *prec = 3; return TS_T_AND;
    case '<': *len = 2; *prec = 4; return s[1] == '<' ? TS_T_SHL : 0;
    case '>': *len = 2; *prec = 4; return s[1] == '>' ? TS_T_SHR : 0;
    case '+': *prec = 5; return TS_T_ADD;
    case '-': *prec = 5; return TS_T_SUB;
    case '*': *prec = 6; return TS_T_MUL;
    default:  return 0;
    }
}

/* precedence climbing, left associative */
static void expression(parser *ps, int min_prec)
{
    prefixed(ps);
    while (!ps->failed) {
        blanks(ps);
        int prec, len, op = binary(ps->at, &prec, &len);
        if (!op || prec < min_prec)
            return;
        ps->at += len;
        expression(ps, prec + 1);
        emit(ps, op, 0);
    }
}

ts_spec *ts_spec_parse(const char *text, const ts_program *p, int inputs_allowed, char *err,
                       size_t errlen)
{
    ts_spec *s = calloc(1, sizeof *s);
    if (!s) {
        snprintf(err, errlen, "out of memory");
        return NULL;
    }
    parser ps = {text, p, s, inputs_allowed, err, errlen, 0};
    int parts[64][64], widths[64], nparts = 0, total = 0;
    for (;;) {
        char name[64];
        blanks(&ps);
        if (!identifier(&ps, name, sizeof name)) {
            fail(&ps, "expected an output name before =");
            break;
        }
        int w = ts_vec_resolve(p, p->outputs, p->noutputs, name, parts[nparts]);
        if (w <= 0) {
            if (w < 0)
                fail(&ps, "output vector %s has missing bits or more than 64", name);
            else
                fail(&ps, "no output called %s", name);
            break;
        }
        widths[nparts++] = w;
        total += w;
        blanks(&ps);
        if (*ps.at != ':' || nparts == 64)
            break;
        ps.at++;
    }
    if (!ps.failed && total > 64)
        fail(&ps, "the left side has %d bits, at most 64 work", total);
    if (!ps.failed) {
        for (int k = nparts - 1; k >= 0; k--)  /* the last vector gives the lowest bits */
            for (int b = 0; b < widths[k]; b++)
                s->out[s->width++] = p->outputs[parts[k][b]];
        blanks(&ps);
        if (*ps.at == '=')
            ps.at++;
        else
            fail(&ps, "expected = after the outputs");
    }
    if (!ps.failed)
        expression(&ps, 0);
    if (!ps.failed) {
        blanks(&ps);
        if (*ps.at)
            fail(&ps, "unexpected '%c'", *ps.at);
    }
    if (ps.failed) {
        free(s);
        return NULL;
    }
    return s;
}

void ts_spec_free(ts_spec *s) { free(s); }
int ts_spec_width(const ts_spec *s) { return s->width; }
uint32_t ts_spec_out_cell(const ts_spec *s, int bit) { return s->out[bit]; }
int ts_spec_nrpn(const ts_spec *s) { return s->nrpn; }
const ts_spec_item *ts_spec_rpn(const ts_spec *s) { return s->rpn; }
int ts_spec_nvars(const ts_spec *s) { return s->nvars; }
int ts_spec_var_bits(const ts_spec *s, int k) { return s->var_bits[k]; }
int ts_spec_var_pos(const ts_spec *s, int k, int bit) { return s->var_pos[k][bit]; }

uint64_t ts_spec_value(const ts_spec *s, uint64_t x)
{
    uint64_t var[MAX_VARS], st[MAX_STACK + 1];
    for (int k = 0; k < s->nvars; k++) {
        uint64_t v = 0;
        for (int b = 0; b < s->var_bits[k]; b++)
            v |= (x >> s->var_pos[k][b] & 1) << b;
        var[k] = v;
    }
    int sp = 0;
    for (int i = 0; i < s->nrpn; i++) {
        const ts_spec_item *it = &s->rpn[i];
        if (it->op == TS_T_NUM) {
            st[sp++] = it->v;
        } else if (it->op == TS_T_VAR) {
            st[sp++] = var[it->v];
        } else if (it->op == TS_T_NOT) {
            st[sp - 1] = ~st[sp - 1];
        } else if (it->op == TS_T_NEG) {
            st[sp - 1] = 0 - st[sp - 1];
        } else {
            uint64_t r = st[--sp], *l = &st[sp - 1];
            switch (it->op) {
            case TS_T_ADD: *l += r; break;
            case TS_T_SUB: *l -= r; break;
            case TS_T_MUL: *l *= r; break;
            case TS_T_AND: *l &= r; break;
            case TS_T_OR:  *l |= r; break;
            case TS_T_XOR: *l ^= r; break;
            case TS_T_SHL: *l = r < 64 ? *l << r : 0; break;
            case TS_T_SHR: *l = r < 64 ? *l >> r : 0; break;
            }
        }
    }
    return s->width == 64 ? st[0] : st[0] & ((1ull << s->width) - 1);
}
