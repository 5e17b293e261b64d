/* ORACLE — test infrastructure only (see oracle.h). A plain restatement of the reference machine
   and the closed forms; no part of the product calls it. */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---- the reference machine (reference src/ref.c) ------------------------------------------ */

typedef struct {
    uint8_t  kind;    /* what a NAND of this instruction reads: 0 plain, 1 id, 2 constant */
    uint8_t  op;      /* 0 NAND, 1 STO, 2 STO to a read-only cell (ignored) */
    uint32_t cell;    /* resolved cell (or id bit for kind 1) */
    uint32_t mask;    /* machine XOR for a read across links */
} decoded;

static int readonly(const ts_program *p, uint32_t c)
{
    return c < TS_CELL_FIRST_FREE || (p->special && p->special[c].kind != TS_CK_PLAIN);
}

/* reference src/asm.c:793-801 */
static uint32_t resolve(const ts_program *p, uint32_t c, uint32_t *mask)
{
    *mask = 0;
    while (p->special && p->special[c].kind == TS_CK_LINK) {
        *mask ^= 1u << p->special[c].dim;
        c = p->special[c].src;
    }
    return c;
}

static unsigned cube_dims(const ts_program *p)
{
    unsigned d = TS_CUBE_DIMS;
    for (uint32_t c = 0; p->special && c < p->ncells; c++)
        if (p->special[c].kind != TS_CK_PLAIN && p->special[c].dim + 1u > d)
            d = p->special[c].dim + 1u;
    return d;
}

static decoded *decode(const ts_program *p)
{
    decoded *d = calloc(p->len ? p->len : 1, sizeof *d);
    for (size_t i = 0; i < p->len; i++) {
        uint32_t a = p->code[i].addr;
        if (p->code[i].op == TS_OP_STO) {
            d[i].op = readonly(p, a) ? 2 : 1;
            d[i].cell = a;
            continue;
        }
        uint32_t mask, c = resolve(p, a, &mask);
        d[i].op = 0;
        d[i].mask = mask;
        if (p->special && p->special[c].kind == TS_CK_ID) {
            d[i].kind = 1;
            d[i].cell = p->special[c].dim;
        } else {
            d[i].kind = c < TS_CELL_FIRST_FREE ? 2 : 0;
            d[i].cell = c;
        }
    }
    return d;
}

static int get_bit(const ts_line *l, unsigned k)
{
    return (int)(l->w[k >> 6] >> (k & 63) & 1);
}

static void put_bit(ts_line *l, unsigned k, int v)
{
    uint64_t b = 1ull << (k & 63);
    l->w[k >> 6] = v ? l->w[k >> 6] | b : l->w[k >> 6] & ~b;
}

/* One cube of M machines, one byte per cell: mem[m * ncells + c], rr[m].
   Machine by machine within an instruction equals lockstep: a NAND changes only RR and nobody
   reads during a STO (reference src/ref.c:117-118). */
static void cube_steps(const ts_program *p, const decoded *d, uint32_t M, uint8_t *mem, uint8_t *rr,
                       size_t from, size_t to)
{
    const uint32_t n = p->ncells;
    for (size_t i = from; i < to; i++) {
        const decoded *in = &d[i];
        if (in->op == 2)
            continue;
        if (in->op == 1) {
            for (uint32_t m = 0; m < M; m++)
                mem[(size_t)m * n + in->cell] = rr[m];
            continue;
        }
        for (uint32_t m = 0; m < M; m++) {
            uint32_t src = m ^ in->mask;
            uint8_t v;
            if (in->kind == 1)
                v = (uint8_t)(src >> in->cell & 1);
            else if (in->kind == 2)
                v = (uint8_t)in->cell;  /* ZERO reads 0, ONE reads 1 */
            else
                v = mem[(size_t)src * n + in->cell];
            rr[m] = (uint8_t)!(rr[m] & v);
        }
    }
}

static void cube_pass(const ts_program *p, const decoded *d, uint32_t M, uint8_t *mem, uint8_t *rr)
{
    cube_steps(p, d, M, mem, rr, 0, p->len);
}

/* `steps` instructions in every machine from instruction pc (wrapping at the end of the
   program), reference src/ref.c:117-130 (ref_cube_step); returns the new pc, or -1 if groups is
   not a whole number of cubes. */
int64_t oracle_steps_lines(const ts_program *p, unsigned groups, ts_line *mem, ts_line *rr,
                           uint64_t pc, uint64_t steps)
{
    unsigned dims = cube_dims(p), cube_groups = 1u << (dims - TS_CUBE_DIMS);
    if (groups % cube_groups || (p->len && pc >= p->len))
        return -1;
    const uint32_t n = p->ncells, M = cube_groups * TS_CUBE_SIZE;
    decoded *d = decode(p);
    uint8_t *bm = malloc((size_t)M * n), *br = malloc(M);
    uint64_t end_pc = pc;
    for (unsigned c0 = 0; c0 < groups; c0 += cube_groups) {
        for (uint32_t m = 0; m < M; m++) {
            const ts_line *g = mem + (size_t)(c0 + m / TS_CUBE_SIZE) * n;
            for (uint32_t c = 0; c < n; c++)
                bm[(size_t)m * n + c] = (uint8_t)get_bit(&g[c], m % TS_CUBE_SIZE);
            bm[(size_t)m * n + TS_CELL_ZERO] = 0;
            bm[(size_t)m * n + TS_CELL_ONE] = 1;
            br[m] = (uint8_t)get_bit(&rr[c0 + m / TS_CUBE_SIZE], m % TS_CUBE_SIZE);
        }
        uint64_t at = pc, left = p->len ? steps : 0;
        while (left) {
            size_t to = p->len - at < left ? p->len : (size_t)(at + left);
            cube_steps(p, d, M, bm, br, (size_t)at, to);
            left -= to - at;
            at = to == p->len ? 0 : to;
        }
        end_pc = at;
        for (uint32_t m = 0; m < M; m++) {
            ts_line *g = mem + (size_t)(c0 + m / TS_CUBE_SIZE) * n;
            for (uint32_t c = TS_CELL_FIRST_FREE; c < n; c++)
                if (!readonly(p, c))
                    put_bit(&g[c], m % TS_CUBE_SIZE, bm[(size_t)m * n + c]);
            put_bit(&rr[c0 + m / TS_CUBE_SIZE], m % TS_CUBE_SIZE, br[m]);
        }
    }
    free(bm);
    free(br);
    free(d);
    return (int64_t)end_pc;
}

int oracle_run_lines(const ts_program *p, unsigned groups, ts_line *mem, ts_line *rr,
                     uint64_t passes)
{
    unsigned dims = cube_dims(p), cube_groups = 1u << (dims - TS_CUBE_DIMS);
    if (groups % cube_groups)
        return 0;
    const uint32_t n = p->ncells, M = cube_groups * TS_CUBE_SIZE;
    decoded *d = decode(p);
    uint8_t *bm = malloc((size_t)M * n), *br = malloc(M);
    for (unsigned c0 = 0; c0 < groups; c0 += cube_groups) {
        for (uint32_t m = 0; m < M; m++) {
            const ts_line *g = mem + (size_t)(c0 + m / TS_CUBE_SIZE) * n;
            for (uint32_t c = 0; c < n; c++)
                bm[(size_t)m * n + c] = (uint8_t)get_bit(&g[c], m % TS_CUBE_SIZE);
            bm[(size_t)m * n + TS_CELL_ZERO] = 0;
            bm[(size_t)m * n + TS_CELL_ONE] = 1;
            br[m] = (uint8_t)get_bit(&rr[c0 + m / TS_CUBE_SIZE], m % TS_CUBE_SIZE);
        }
        for (uint64_t t = 0; t < passes; t++)
            cube_pass(p, d, M, bm, br);
        for (uint32_t m = 0; m < M; m++) {
            ts_line *g = mem + (size_t)(c0 + m / TS_CUBE_SIZE) * n;
            for (uint32_t c = TS_CELL_FIRST_FREE; c < n; c++)
                if (!readonly(p, c))
                    put_bit(&g[c], m % TS_CUBE_SIZE, bm[(size_t)m * n + c]);
            put_bit(&rr[c0 + m / TS_CUBE_SIZE], m % TS_CUBE_SIZE, br[m]);
        }
    }
    free(bm);
    free(br);
    free(d);
    return 1;
}

/* reference src/ref.c:24-32 and src/lanes.c:173-188 (id cells hold bit d of the machine number
   within its cube) */
void oracle_reset_lines(const ts_program *p, unsigned groups, ts_line *mem, ts_line *rr)
{
    unsigned cube_groups = 1u << (cube_dims(p) - TS_CUBE_DIMS);
    memset(mem, 0, (size_t)groups * p->ncells * sizeof *mem);
    memset(rr, 0, (size_t)groups * sizeof *rr);
    for (unsigned g = 0; g < groups; g++) {
        ts_line *m = mem + (size_t)g * p->ncells;
        memset(&m[TS_CELL_ONE], 0xFF, sizeof(ts_line));
        for (uint32_t c = 0; p->special && c < p->ncells; c++)
            if (p->special[c].kind == TS_CK_ID)
                for (unsigned k = 0; k < TS_CUBE_SIZE; k++) {
                    unsigned machine = (g % cube_groups) * TS_CUBE_SIZE + k;
                    put_bit(&m[c], k, (int)(machine >> p->special[c].dim & 1));
                }
    }
}

static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* reference src/check.c:66-77 */
void oracle_random_lines(const ts_program *p, unsigned groups, uint64_t seed, ts_line *mem,
                         ts_line *rr)
{
    oracle_reset_lines(p, groups, mem, rr);
    uint64_t s = seed;
    for (unsigned g = 0; g < groups; g++) {
        ts_line *m = mem + (size_t)g * p->ncells;
        for (uint32_t a = TS_CELL_FIRST_FREE; a < p->ncells; a++)
            for (int w = 0; w < 8; w++) {
                uint64_t v = splitmix64(&s);
                if (!readonly(p, a))
                    m[a].w[w] = v;
            }
        for (int w = 0; w < 8; w++)
            rr[g].w[w] = splitmix64(&s);
    }
}

/* ---- enumeration (reference src/enum.c, src/enum.h:3-5) ----------------------------------- */

static void add_hit(oracle_enum_result *r, uint64_t x)
{
    int i;
    if (r->nhit < 16) {
        i = r->nhit++;
    } else {
        if (x >= r->hit[15])
            return;
        i = 15;
    }
    for (; i > 0 && r->hit[i - 1] > x; i--)
        r->hit[i] = r->hit[i - 1];
    r->hit[i] = x;
}

/* one machine from power-on with input k = bit n-1-k of x, `ticks` passes */
static void run_one(const ts_program *p, const decoded *d, const uint32_t *in_cells, int n,
                    uint64_t x, unsigned ticks, uint8_t *mem)
{
    uint8_t rr = 0;
    memset(mem, 0, p->ncells);
    mem[TS_CELL_ONE] = 1;
    for (int k = 0; k < n; k++)
        mem[in_cells[k]] = (uint8_t)(x >> (n - 1 - k) & 1);
    for (unsigned t = 0; t < ticks; t++)
        cube_pass(p, d, 1, mem, &rr);
}

int oracle_enum(int mode, const ts_program *a, const ts_program *b, const uint32_t *b_in,
                const uint32_t *b_out, const uint32_t *out_cells, int width,
                const uint64_t *expect, uint64_t target, unsigned ticks, int part, int part_end,
                int parts, oracle_enum_result *r)
{
    memset(r, 0, sizeof *r);
    int n = (int)a->ninputs;
    if (n > 30 || a->special || (b && b->special))
        return 0;
    decoded *da = decode(a), *db = b ? decode(b) : NULL;
    uint8_t *ma = malloc(a->ncells), *mb = b ? malloc(b->ncells) : NULL;
    r->space = 1ull << n;
    for (uint64_t x = 0; x < r->space; x++) {
        int share = (int)((x >> 15) % (uint64_t)parts);  /* chunk = 64 passes of 512 */
        if (share < part || share > part_end)
            continue;
        r->tried++;
        run_one(a, da, a->inputs, n, x, ticks, ma);
        int hit;
        if (mode == 2) {
            uint64_t got = 0;
            for (int bit = 0; bit < width; bit++)
                got |= (uint64_t)ma[out_cells[bit]] << bit;
            hit = got == target;
        } else if (mode == 1) {
            run_one(b, db, b_in, n, x, ticks, mb);
            hit = 0;
            for (size_t o = 0; o < a->noutputs; o++)
                hit |= ma[a->outputs[o]] != mb[b_out[o]];
        } else {
            uint64_t got = 0;
            for (int bit = 0; bit < width; bit++)
                got |= (uint64_t)ma[out_cells[bit]] << bit;
            hit = got != expect[x];
        }
        if (hit) {
            r->hits++;
            add_hit(r, x);
        }
    }
    free(da);
    free(db);
    free(ma);
    free(mb);
    return 1;
}

/* ---- cycle lengths (reference src/tester.c:96-164) ---------------------------------------- */

static void lane_state(const ts_program *p, const ts_line *mem, const ts_line *rr, unsigned k,
                       uint8_t *m, uint8_t *r)
{
    for (uint32_t c = 0; c < p->ncells; c++)
        m[c] = (uint8_t)get_bit(&mem[c], k);
    m[TS_CELL_ZERO] = 0;
    m[TS_CELL_ONE] = 1;
    *r = (uint8_t)get_bit(rr, k);
}

uint64_t oracle_brent_lines(const ts_program *p, const ts_line *mem, const ts_line *rr,
                            uint64_t max_passes, uint64_t *lam_out)
{
    decoded *d = decode(p);
    const uint32_t n = p->ncells;
    uint8_t *hare = malloc(n), *tort = malloc(n), hr, tr;
    uint64_t most = 0;
    for (unsigned k = 0; k < TS_CUBE_SIZE; k++) {
        lane_state(p, mem, rr, k, hare, &hr);
        memcpy(tort, hare, n);
        tr = hr;
        cube_pass(p, d, 1, hare, &hr);
        uint64_t lam = 1, power = 1, steps = 1, found = 0;
        for (;;) {
            if (hr == tr && !memcmp(hare + TS_CELL_FIRST_FREE, tort + TS_CELL_FIRST_FREE,
                                    n - TS_CELL_FIRST_FREE)) {
                found = lam;
                break;
            }
            if (steps >= max_passes)
                break;
            if (power == lam) {
                memcpy(tort, hare, n);
                tr = hr;
                power *= 2;
                lam = 0;
            }
            cube_pass(p, d, 1, hare, &hr);
            lam++;
            steps++;
        }
        lam_out[k] = found;
        if (steps > most)
            most = steps;
    }
    free(hare);
    free(tort);
    free(d);
    return most;
}

uint64_t oracle_exact_cycle(const ts_program *p, const ts_line *mem, const ts_line *rr,
                            unsigned lane)
{
    const uint32_t n = p->ncells, bits = n - TS_CELL_FIRST_FREE + 1;
    if (bits > 24)
        return 0;
    decoded *d = decode(p);
    uint8_t *m = malloc(n), r;
    lane_state(p, mem, rr, lane, m, &r);
    int64_t *seen = malloc(sizeof *seen << bits);
    for (size_t i = 0; i < (size_t)1 << bits; i++)
        seen[i] = -1;
    uint64_t lam = 0;
    for (int64_t t = 0;; t++) {
        size_t s = r;
        for (uint32_t c = TS_CELL_FIRST_FREE; c < n; c++)
            s |= (size_t)m[c] << (c - TS_CELL_FIRST_FREE + 1);
        if (seen[s] >= 0) {
            lam = (uint64_t)(t - seen[s]);
            break;
        }
        seen[s] = t;
        cube_pass(p, d, 1, m, &r);
    }
    free(seen);
    free(m);
    free(d);
    return lam;
}

/* ---- closed forms ------------------------------------------------------------------------- */

/* reference src/gen.c:686-700: B3/S23 on a w x h torus */
void oracle_life_scalar(uint8_t *board, int w, int h)
{
    uint8_t *next = malloc((size_t)w * (size_t)h);
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
            int alive = 0;
            for (int dr = -1; dr <= 1; dr++)
                for (int dc = -1; dc <= 1; dc++)
                    if (dr || dc)
                        alive += board[((r + dr + h) % h) * w + (c + dc + w) % w];
            next[r * w + c] = (uint8_t)(alive == 3 || (alive == 2 && board[r * w + c]));
        }
    memcpy(board, next, (size_t)w * (size_t)h);
    free(next);
}

/* reference src/benes.c:77-86 */
void oracle_benes_apply(const uint8_t *ctrl, uint64_t *v)
{
    uint64_t next[TS_CUBE_SIZE];
    for (int s = 0; s < 2 * TS_CUBE_DIMS - 1; s++) {
        unsigned d = s < TS_CUBE_DIMS ? (unsigned)s : (unsigned)(2 * TS_CUBE_DIMS - 2 - s);
        for (unsigned k = 0; k < TS_CUBE_SIZE; k++)
            next[k] = ctrl[s * TS_CUBE_SIZE + k] ? v[k ^ (1u << d)] : v[k];
        memcpy(v, next, sizeof next);
    }
}

/* reference src/gen.c:79-121: x_i ^= !x_(i+1) & x_(i+2), bit i <- bit (m i + 3) mod n with m the
   first of 5, 7, 11, .. coprime to n, XOR splitmix64-finalised round constant */
uint64_t oracle_toyhash(uint64_t x, int n, int rounds)
{
    static const int primes[] = {5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};
    int m = 0;
    for (size_t i = 0; i < sizeof primes / sizeof primes[0] && !m; i++) {
        int a = primes[i], b = n;
        while (b) {
            int t = a % b;
            a = b;
            b = t;
        }
        if (a == 1)
            m = primes[i];
    }
    for (int r = 0; r < rounds; r++) {
        uint64_t y = 0, z = 0;
        for (int i = 0; i < n; i++) {
            uint64_t xi = x >> i & 1, x1 = x >> (i + 1) % n & 1, x2 = x >> (i + 2) % n & 1;
            y |= (xi ^ ((x1 ^ 1) & x2)) << i;
        }
        for (int i = 0; i < n; i++)
            z |= (y >> (m * i + 3) % n & 1) << i;
        uint64_t rc = (uint64_t)(r + 1) * 0x9E3779B97F4A7C15ull;
        rc = (rc ^ (rc >> 30)) * 0xBF58476D1CE4E5B9ull;
        rc = (rc ^ (rc >> 27)) * 0x94D049BB133111EBull;
        rc ^= rc >> 31;
        x = z ^ (n == 64 ? rc : rc & ((1ull << n) - 1));
    }
    return x;
}

/* B3/S23 on a torus of h rows x 64*ww columns, 64 cells per word (cell c of a row is bit c % 64 of
   word c / 64): the neighbour count is summed bit-parallel with half/full adders into bits
   (ones, twos, fours); pinned against oracle_life_scalar. */
void oracle_life_words(uint64_t *b, int ww, int h, int gens)
{
    uint64_t *next = malloc((size_t)ww * (size_t)h * sizeof *next);
    for (int g = 0; g < gens; g++) {
        for (int r = 0; r < h; r++) {
            const uint64_t *up = b + (size_t)((r + h - 1) % h) * ww, *mid = b + (size_t)r * ww,
                           *dn = b + (size_t)((r + 1) % h) * ww;
            for (int i = 0; i < ww; i++) {
                int il = (i + ww - 1) % ww, ir = (i + 1) % ww;
                uint64_t n[8];
                const uint64_t *rows[3] = {up, mid, dn};
                int k = 0;
                for (int j = 0; j < 3; j++) {
                    uint64_t x = rows[j][i];
                    uint64_t west = x << 1 | rows[j][il] >> 63;  /* cell c-1 */
                    uint64_t east = x >> 1 | rows[j][ir] << 63;  /* cell c+1 */
                    n[k++] = west;
                    n[k++] = east;
                    if (j != 1)
                        n[k++] = x;
                }
                uint64_t ones = 0, twos = 0, fours = 0;
                for (int j = 0; j < 8; j++) {
                    uint64_t c1 = ones & n[j];
                    ones ^= n[j];
                    uint64_t c2 = twos & c1;
                    twos ^= c1;
                    fours |= c2;  /* 4 or more: dead either way */
                }
                next[(size_t)r * ww + i] = ~fours & twos & (ones | mid[i]);
            }
        }
        memcpy(b, next, (size_t)ww * (size_t)h * sizeof *next);
    }
    free(next);
}
