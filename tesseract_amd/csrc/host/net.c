/* ISCAS .bench netlists -> programs (reference src/net.c).

     INPUT(a)  OUTPUT(z)  z = NAND(a, b)  q = DFF(d)   # comment

   Signals are numbered in order of first appearance. Combinational gates are placed in Kahn
   order (inputs and flip-flop outputs are ready from the start; a placed gate releases its
   readers, most recently seen reader first) and built with the gate builder; DFFs become
   registers. Signal names become cell names: characters outside [A-Za-z0-9_] turn into '_', and
   a name starting with a digit, equal to ZERO/ONE or starting with "__" gets an 'n' in front. */
#include <ctype.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "tesseract_host.h"

enum { FLIPFLOP = -2, UNDRIVEN = -1, PRIMARY = -2 };

typedef struct {
    int  type;
    int  out;
    int *in;
    int  nin;
    int  line;
} gate;

typedef struct {
    char      **name;
    int        *driver;  /* gate index, UNDRIVEN or PRIMARY */
    int         n, cap;
    gate       *g;
    int         ng, capg;
    int        *inputs, *outputs;
    int         nin, nout, capin, capout;
    int        *slot;    /* open-addressing table of signal numbers by name, -1 = empty */
    int         nslot;   /* a power of two, at least twice n */
    char       *err;
    size_t      errlen;
    const char *file;
} netlist;

static void *xgrow(void *p, int *cap, int need, size_t elem)
{
    if (need <= *cap)
        return p;
    int c = *cap ? *cap : 16;
    while (c < need)
        c *= 2;
    void *q = realloc(p, (size_t)c * elem);
    if (!q)
        abort();
    *cap = c;
    return q;
}

static int push(int **arr, int *n, int *cap, int v)
{
    *arr = xgrow(*arr, cap, *n + 1, sizeof **arr);
    (*arr)[(*n)++] = v;
    return v;
}

static int failed(netlist *net, int line, const char *fmt, ...)
{
    int k = snprintf(net->err, net->errlen, "%s:%d: ", net->file, line);
    if (k > 0 && (size_t)k < net->errlen) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(net->err + k, net->errlen - (size_t)k, fmt, ap);
        va_end(ap);
    }
    return 0;
}

static unsigned name_hash(const char *s)
{
    unsigned h = 2166136261u;  /* FNV-1a */
    for (; *s; s++)
        h = (h ^ (unsigned char)*s) * 16777619u;
    return h;
}

static void rehash(netlist *net, int nslot)
{
    free(net->slot);
    net->slot = malloc((size_t)nslot * sizeof *net->slot);
    if (!net->slot)
        abort();
    net->nslot = nslot;
    for (int k = 0; k < nslot; k++)
        net->slot[k] = -1;
    for (int i = 0; i < net->n; i++) {
        unsigned k = name_hash(net->name[i]) & (unsigned)(nslot - 1);
        while (net->slot[k] >= 0)
            k = (k + 1) & (unsigned)(nslot - 1);
        net->slot[k] = i;
    }
}

static int signal_of(netlist *net, const char *name)
{
    if (2 * (net->n + 1) > net->nslot)
        rehash(net, net->nslot ? 2 * net->nslot : 1024);
    unsigned k = name_hash(name) & (unsigned)(net->nslot - 1);
    for (; net->slot[k] >= 0; k = (k + 1) & (unsigned)(net->nslot - 1))
        if (!strcmp(net->name[net->slot[k]], name))
            return net->slot[k];
    net->slot[k] = net->n;
    int c1 = net->cap;
    net->name = xgrow(net->name, &c1, net->n + 1, sizeof *net->name);
    int c2 = net->cap;
    net->driver = xgrow(net->driver, &c2, net->n + 1, sizeof *net->driver);
    net->cap = c1;
    net->name[net->n] = strdup(name);
    net->driver[net->n] = UNDRIVEN;
    return net->n++;
}

static int word(const char **p, char *buf, size_t len)
{
    const char *s = *p;
    while (*s == ' ' || *s == '\t')
        s++;
    size_t n = 0;
    for (; isalnum((unsigned char)*s) || (*s && strchr("_.[]$", *s)); s++)
        if (n + 1 < len)
            buf[n++] = *s;
    buf[n] = 0;
    *p = s;
    return n > 0;
}

static int punct(const char **p, char c)
{
    while (**p == ' ' || **p == '\t')
        (*p)++;
    if (**p != c)
        return 0;
    (*p)++;
    return 1;
}

static int one_line(netlist *net, char *line, int ln)
{
    char *hash = strchr(line, '#');
    if (hash)
        *hash = 0;
    const char *p = line;
    char head[256], arg[256], type[64];
    if (!word(&p, head, sizeof head)) {
        while (*p == ' ' || *p == '\t' || *p == '\r')
            p++;
        return *p ? failed(net, ln, "cannot read this line") : 1;
    }
    int is_in = !strcmp(head, "INPUT") || !strcmp(head, "input");
    int is_out = !strcmp(head, "OUTPUT") || !strcmp(head, "output");
    if (is_in || is_out) {
        if (!punct(&p, '(') || !word(&p, arg, sizeof arg) || !punct(&p, ')'))
            return failed(net, ln, "expected %s(name)", head);
        int s = signal_of(net, arg);
        if (is_in) {
            if (net->driver[s] != UNDRIVEN)
                return failed(net, ln, "%s is driven twice", arg);
            net->driver[s] = PRIMARY;
            push(&net->inputs, &net->nin, &net->capin, s);
        } else {
            push(&net->outputs, &net->nout, &net->capout, s);
        }
        return 1;
    }
    if (!punct(&p, '=') || !word(&p, type, sizeof type) || !punct(&p, '('))
        return failed(net, ln, "expected name = GATE(inputs)");
    int t = ts_gate_type(type);
    if (t < 0 && (!strcmp(type, "DFF") || !strcmp(type, "dff")))
        t = FLIPFLOP;
    if (t == -1)
        return failed(net, ln, "unknown gate %s", type);
    int out = signal_of(net, head);
    if (net->driver[out] != UNDRIVEN)
        return failed(net, ln, "%s is driven twice", head);
    net->g = xgrow(net->g, &net->capg, net->ng + 1, sizeof *net->g);
    gate *g = &net->g[net->ng];
    memset(g, 0, sizeof *g);
    g->type = t;
    g->out = out;
    g->line = ln;
    int cap = 0;
    do {
        if (!word(&p, arg, sizeof arg))
            return failed(net, ln, "expected an input name");
        push(&g->in, &g->nin, &cap, signal_of(net, arg));
    } while (punct(&p, ','));
    if (!punct(&p, ')'))
        return failed(net, ln, "expected )");
    if ((t == TS_G_NOT || t == TS_G_BUF || t == FLIPFLOP) && g->nin != 1)
        return failed(net, ln, "%s takes one input", type);
    net->driver[out] = net->ng++;
    return 1;
}

static int eq_nocase(const char *a, const char *b)
{
    while (*a && *b && tolower((unsigned char)*a) == tolower((unsigned char)*b)) {
        a++;
        b++;
    }
    return *a == 0 && *b == 0;
}

static void cell_name(const char *s, char *buf, size_t len)
{
    size_t n = 0;
    if (isdigit((unsigned char)*s) || eq_nocase(s, "ZERO") || eq_nocase(s, "ONE") ||
        !strncmp(s, "__", 2))
        buf[n++] = 'n';
    for (; *s && n + 1 < len; s++)
        buf[n++] = isalnum((unsigned char)*s) || *s == '_' ? *s : '_';
    buf[n] = 0;
}

static int build(netlist *net, ts_program *out)
{
    for (int s = 0; s < net->n; s++)
        if (net->driver[s] == UNDRIVEN)
            return failed(net, 0, "signal %s is used but never driven", net->name[s]);
    int ng = net->ng;
    int *pending = calloc((size_t)ng + 1, sizeof *pending), *order = malloc(((size_t)ng + 1) * sizeof *order);
    int *head = malloc(((size_t)net->n + 1) * sizeof *head);
    int *reader = NULL, *link = NULL, nedge = 0, cap_r = 0, cap_l = 0;
    for (int s = 0; s < net->n; s++)
        head[s] = -1;
    for (int i = 0; i < ng; i++) {
        if (net->g[i].type == FLIPFLOP)
            continue;
        for (int k = 0; k < net->g[i].nin; k++) {
            int d = net->driver[net->g[i].in[k]];
            if (d >= 0 && net->g[d].type != FLIPFLOP) {
                pending[i]++;
                int sig = net->g[d].out;
                int e = nedge;
                push(&reader, &nedge, &cap_r, i);
                int nl = e;
                push(&link, &nl, &cap_l, head[sig]);
                head[sig] = e;
            }
        }
    }
    int placed = 0, done = 0;
    for (int i = 0; i < ng; i++)
        if (net->g[i].type != FLIPFLOP && pending[i] == 0)
            order[placed++] = i;
    while (done < placed) {
        int g = order[done++];
        for (int e = head[net->g[g].out]; e >= 0; e = link[e])
            if (--pending[reader[e]] == 0)
                order[placed++] = reader[e];
    }
    int combinational = 0;
    for (int i = 0; i < ng; i++)
        combinational += net->g[i].type != FLIPFLOP;
    int ok = placed == combinational;
    if (!ok)
        for (int i = 0; i < ng; i++)
            if (net->g[i].type != FLIPFLOP && pending[i] > 0) {
                failed(net, net->g[i].line, "%s is part of a loop without a DFF",
                       net->name[net->g[i].out]);
                break;
            }
    if (ok) {
        ts_builder *b = ts_b_new();
        uint32_t *cell = malloc(((size_t)net->n + 1) * sizeof *cell);
        for (int s = 0; ok && s < net->n; s++) {
            char nm[300];
            cell_name(net->name[s], nm, sizeof nm);
            if (ts_b_has(b, nm))
                ok = failed(net, 0, "two signals both become cell %s", nm);
            else
                cell[s] = ts_b_cell(b, nm);
        }
        if (ok) {
            for (int i = 0; i < net->nin; i++)
                ts_b_input(b, cell[net->inputs[i]]);
            uint32_t *in = NULL;
            int capin = 0;
            for (int k = 0; k < placed; k++) {
                const gate *g = &net->g[order[k]];
                in = xgrow(in, &capin, g->nin, sizeof *in);
                for (int j = 0; j < g->nin; j++)
                    in[j] = cell[g->in[j]];
                ts_b_gate(b, g->type, cell[g->out], in, g->nin);
            }
            free(in);
            for (int i = 0; i < ng; i++)
                if (net->g[i].type == FLIPFLOP)
                    ts_b_reg(b, cell[net->g[i].out], cell[net->g[i].in[0]]);
            for (int i = 0; i < net->nout; i++)
                ts_b_output(b, cell[net->outputs[i]]);
        }
        ts_b_finish(b, out);
        if (!ok)
            ts_program_free(out);
        free(cell);
    }
    free(pending);
    free(order);
    free(head);
    free(reader);
    free(link);
    return ok;
}

int ts_net_parse(const char *text, const char *name, ts_program *out, char *err, size_t errlen)
{
    netlist net;
    memset(&net, 0, sizeof net);
    memset(out, 0, sizeof *out);
    net.err = err;
    net.errlen = errlen;
    net.file = name;
    char *copy = strdup(text);
    int ok = 1, ln = 1;
    for (char *line = copy; ok && line; ln++) {
        char *nl = strchr(line, '\n');
        if (nl)
            *nl = 0;
        ok = one_line(&net, line, ln);
        line = nl ? nl + 1 : NULL;
    }
    free(copy);
    if (ok)
        ok = build(&net, out);
    for (int i = 0; i < net.n; i++)
        free(net.name[i]);
    for (int i = 0; i < net.ng; i++)
        free(net.g[i].in);
    free(net.name);
    free(net.driver);
    free(net.g);
    free(net.inputs);
    free(net.outputs);
    free(net.slot);
    return ok;
}
