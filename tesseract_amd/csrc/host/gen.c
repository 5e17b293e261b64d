/* Workload generators (reference src/gen.c): array multiplier, toy hash, Game of Life on one
   machine and on a grid of machines, running count over a cube, Walsh-Hadamard transform, XOR
   convolution, plus the adder and the cube library's PERM8/SORT8 built from gates.

   mul, hash, life, life_cube, rank, wht and xorconv reproduce the reference's programs instruction
   for instruction (same gate sequences, same order of first use of every cell). */
#include <stdlib.h>
#include <string.h>

#include "tesseract_host.h"

static uint32_t named(ts_builder *b, const char *fmt, int i, int j)
{
    char name[64];
    snprintf(name, sizeof name, fmt, i, j);
    return ts_b_cell(b, name);
}

/* pc[0..2n) = a * bv, rows of AND partial products summed by half/full adders
   (reference src/gen.c:38-58) */
static void multiply(ts_builder *b, const uint32_t *pc, const uint32_t *a, const uint32_t *bv, int n)
{
    uint32_t pp = ts_b_cell(b, "pp"), carry = ts_b_cell(b, "carry"), zero = TS_CELL_ZERO;
    for (int i = 0; i < n; i++)
        ts_b_gate2(b, TS_G_AND, pc[i], a[i], bv[0]);
    for (int k = n; k < 2 * n; k++)
        ts_b_gate(b, TS_G_BUF, pc[k], &zero, 1);
    for (int j = 1; j < n; j++)
        for (int i = 0; i < n; i++) {
            ts_b_gate2(b, TS_G_AND, pp, a[i], bv[j]);
            if (i == 0)
                ts_b_half_adder(b, pc[j], carry, pc[j], pp);
            else
                ts_b_full_adder(b, pc[i + j], i == n - 1 ? pc[j + n] : carry, pc[i + j], pp, carry);
        }
}

void ts_gen_mul(ts_program *p, int n)
{
    ts_builder *b = ts_b_new();
    uint32_t a[64], bv[64], pc[128];
    for (int i = 0; i < n; i++)
        ts_b_input(b, a[i] = named(b, "a%d", i, 0));
    for (int i = 0; i < n; i++)
        ts_b_input(b, bv[i] = named(b, "b%d", i, 0));
    for (int k = 0; k < 2 * n; k++)
        pc[k] = named(b, "p%d", k, 0);
    multiply(b, pc, a, bv, n);
    for (int k = 0; k < 2 * n; k++)
        ts_b_output(b, pc[k]);
    ts_b_finish(b, p);
}

void ts_gen_adder(ts_program *p, int bits)
{
    ts_builder *b = ts_b_new();
    uint32_t a[64], bv[64], s[64];
    for (int i = bits - 1; i >= 0; i--)
        ts_b_input(b, a[i] = named(b, "a%d", i, 0));
    for (int i = bits - 1; i >= 0; i--)
        ts_b_input(b, bv[i] = named(b, "b%d", i, 0));
    uint32_t c = ts_b_cell(b, "c");
    ts_b_input(b, c);
    for (int i = bits - 1; i >= 0; i--)
        ts_b_output(b, s[i] = named(b, "s%d", i, 0));
    ts_b_output(b, c);
    for (int i = 0; i < bits; i++)
        ts_b_full_adder(b, s[i], c, a[i], bv[i], c);
    ts_b_finish(b, p);
}

/* ---- toy hash (reference src/gen.c:79-162) ------------------------------------------------- */

static int gcd(int a, int b)
{
    while (b) {
        int t = a % b;
        a = b;
        b = t;
    }
    return a;
}

/* The bit permutation i -> (m i + 3) mod n with the first small prime m coprime to n. */
static int perm_multiplier(int n)
{
    static const int primes[] = {5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};
    for (size_t i = 0; i < sizeof primes / sizeof primes[0]; i++)
        if (gcd(primes[i], n) == 1)
            return primes[i];
    return 0;
}

/* splitmix64's finaliser applied to (r + 1) * golden ratio, cut to n bits */
static uint64_t round_constant(int r, int n)
{
    uint64_t z = (uint64_t)(r + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return n == 64 ? z : z & ((1ull << n) - 1);
}

uint64_t ts_toyhash(uint64_t x, int n, int rounds)
{
    int m = perm_multiplier(n);
    for (int r = 0; r < rounds; r++) {
        uint64_t y = 0, z = 0;
        for (int i = 0; i < n; i++) {
            uint64_t xi = x >> i & 1, x1 = x >> (i + 1) % n & 1, x2 = x >> (i + 2) % n & 1;
            y |= (xi ^ ((x1 ^ 1) & x2)) << i;
        }
        for (int i = 0; i < n; i++)
            z |= (y >> (m * i + 3) % n & 1) << i;
        x = z ^ round_constant(r, n);
    }
    return x;
}

int ts_gen_hash(ts_program *p, int n, int rounds)
{
    int m = perm_multiplier(n);
    if (!m || n > 64)
        return 0;
    ts_builder *b = ts_b_new();
    uint32_t cur[64], nxt[64];
    for (int i = 0; i < n; i++)
        ts_b_input(b, cur[i] = named(b, "x%d", i, 0));
    uint32_t nt = ts_b_cell(b, "nt"), tt = ts_b_cell(b, "tt");
    for (int r = 0; r < rounds; r++) {
        int bank = r % 2;
        uint64_t rc = round_constant(r, n);
        for (int i = 0; i < n; i++) {
            uint32_t y = named(b, "y%d_%d", bank, i);
            ts_b_gate(b, TS_G_NOT, nt, &cur[(i + 1) % n], 1);
            ts_b_gate2(b, TS_G_AND, tt, nt, cur[(i + 2) % n]);
            ts_b_gate2(b, TS_G_XOR, y, cur[i], tt);
        }
        for (int i = 0; i < n; i++) {
            uint32_t src = named(b, "y%d_%d", bank, (m * i + 3) % n);
            if (rc >> i & 1) {
                nxt[i] = named(b, "k%d_%d", bank, i);
                ts_b_gate(b, TS_G_NOT, nxt[i], &src, 1);
            } else {
                nxt[i] = src;
            }
        }
        memcpy(cur, nxt, (size_t)n * sizeof *cur);
    }
    for (int i = 0; i < n; i++) {
        uint32_t h = named(b, "h%d", i, 0);
        ts_b_gate(b, TS_G_BUF, h, &cur[i], 1);
        ts_b_output(b, h);
    }
    ts_b_finish(b, p);
    return 1;
}

/* ---- Game of Life (reference src/gen.c:164-219) -------------------------------------------- */

/* One generation of every cell of a w x h board; halo[(r+1)(w+2) + c+1] is the cell at row r,
   column c for r = -1..h, c = -1..w. The 8 neighbours are counted with adders into bits
   c0 (1), c1 (2), u and v (4 each); the cell lives next iff c1 & !(u | v) & (c0 | self). */
static void life_generation(ts_builder *b, const uint32_t *board, const uint32_t *halo, int w, int h)
{
    uint32_t s0 = ts_b_cell(b, "s0"), k0 = ts_b_cell(b, "k0"), s1 = ts_b_cell(b, "s1"),
             k1 = ts_b_cell(b, "k1"), s2 = ts_b_cell(b, "s2"), k2 = ts_b_cell(b, "k2"),
             c0 = ts_b_cell(b, "c0"), k3 = ts_b_cell(b, "k3"), t = ts_b_cell(b, "t"),
             u = ts_b_cell(b, "u"), c1 = ts_b_cell(b, "c1"), v = ts_b_cell(b, "v"),
             r1 = ts_b_cell(b, "r1"), r2 = ts_b_cell(b, "r2");
    int W = w + 2;
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
            uint32_t nb[8];
            int k = 0;
            for (int dr = -1; dr <= 1; dr++)
                for (int dc = -1; dc <= 1; dc++)
                    if (dr || dc)
                        nb[k++] = halo[(r + 1 + dr) * W + c + 1 + dc];
            ts_b_full_adder(b, s0, k0, nb[0], nb[1], nb[2]);
            ts_b_full_adder(b, s1, k1, nb[3], nb[4], nb[5]);
            ts_b_half_adder(b, s2, k2, nb[6], nb[7]);
            ts_b_full_adder(b, c0, k3, s0, s1, s2);
            ts_b_full_adder(b, t, u, k0, k1, k2);
            ts_b_half_adder(b, c1, v, t, k3);
            uint32_t self = board[r * w + c], next = named(b, "n%d_%d", r, c);
            uint32_t all[3] = {c1, r1, r2};
            ts_b_gate2(b, TS_G_OR, r1, c0, self);
            ts_b_gate2(b, TS_G_NOR, r2, u, v);
            ts_b_gate(b, TS_G_AND, next, all, 3);
            ts_b_reg(b, self, next);
        }
}

static uint32_t *board_cells(ts_builder *b, int w, int h, uint32_t *board)
{
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++)
            board[r * w + c] = named(b, "x%d_%d", r, c);
    return malloc((size_t)(w + 2) * (size_t)(h + 2) * sizeof(uint32_t));
}

void ts_gen_life(ts_program *p, int w, int h, uint32_t *board)
{
    ts_builder *b = ts_b_new();
    uint32_t *halo = board_cells(b, w, h, board);
    for (int r = -1; r <= h; r++)
        for (int c = -1; c <= w; c++)
            halo[(r + 1) * (w + 2) + c + 1] = board[((r + h) % h) * w + (c + w) % w];
    life_generation(b, board, halo, w, h);
    free(halo);
    ts_b_finish(b, p);
}

/* ---- grids over the cube (reference src/gen.c:221-319) -------------------------------------- */

unsigned ts_grid_machine(unsigned row, unsigned col, int cbits)
{
    return (row ^ row >> 1) << cbits | (col ^ col >> 1);
}

/* Gray-code axis over dims lo..lo+k-1: next[j] / prev[j] is 1 where the next / previous position
   on the axis is across dimension lo + j (reference src/gen.c:229-262). */
static void grid_masks(ts_builder *b, int lo, int k, uint32_t *next, uint32_t *prev, const char *tag)
{
    char name[32];
    uint32_t id[TS_CUBE_MAX_DIMS], v[4], zero = TS_CELL_ZERO;
    for (int j = 0; j < k; j++) {
        id[j] = named(b, "ID%d", lo + j, 0);
        snprintf(name, sizeof name, "%s_n%d", tag, j);
        next[j] = ts_b_cell(b, name);
        snprintf(name, sizeof name, "%s_v%d", tag, j);
        prev[j] = ts_b_cell(b, name);
    }
    for (int i = 0; i < 4; i++) {
        snprintf(name, sizeof name, "%s_%c", tag, "pqol"[i]);
        v[i] = ts_b_cell(b, name);
    }
    uint32_t odd = v[0], even = v[1], below = v[2], low = v[3];
    ts_b_gate(b, TS_G_XOR, odd, id, k);
    ts_b_gate(b, TS_G_NOT, even, &odd, 1);
    ts_b_gate(b, TS_G_BUF, next[0], &even, 1);
    ts_b_gate(b, TS_G_BUF, prev[0], &odd, 1);
    ts_b_gate(b, TS_G_BUF, below, &zero, 1);
    for (int j = 0; j < k; j++) {
        ts_b_gate(b, TS_G_NOT, low, &below, 1);
        ts_b_gate2(b, TS_G_AND, low, low, id[j]);
        if (j + 1 < k) {
            ts_b_gate2(b, TS_G_AND, next[j + 1], low, odd);
            ts_b_gate2(b, TS_G_AND, prev[j + 1], low, even);
        } else {
            ts_b_gate2(b, TS_G_AND, low, low, odd);
            ts_b_gate2(b, TS_G_OR, next[j], next[j], low);
        }
        ts_b_gate2(b, TS_G_OR, below, below, id[j]);
    }
    ts_b_gate(b, TS_G_NOT, low, &below, 1);
    ts_b_gate2(b, TS_G_OR, prev[k - 1], prev[k - 1], low);
}

/* new cell `name` = src of the machine mask[0..k) points to, one link per dimension */
static uint32_t grid_gather(ts_builder *b, const char *name, uint32_t src, int lo, int k,
                            const uint32_t *mask)
{
    uint32_t t[TS_CUBE_MAX_DIMS], out = ts_b_cell(b, name);
    char link[48];
    for (int j = 0; j < k; j++) {
        snprintf(link, sizeof link, "lk%u_%d", (unsigned)src, lo + j);
        uint32_t l = ts_b_has(b, link) ? ts_b_cell(b, link) : ts_b_link(b, link, src, lo + j);
        t[j] = named(b, "gt%d", j, 0);
        ts_b_gate2(b, TS_G_NAND, t[j], mask[j], l);
    }
    ts_b_gate(b, TS_G_NAND, out, t, k);
    return out;
}

void ts_gen_life_cube(ts_program *p, int w, int h, int rbits, int cbits, uint32_t *board)
{
    ts_builder *b = ts_b_new();
    int W = w + 2;
    uint32_t *halo = board_cells(b, w, h, board), right[TS_CUBE_MAX_DIMS], left[TS_CUBE_MAX_DIMS],
             down[TS_CUBE_MAX_DIMS], up[TS_CUBE_MAX_DIMS];
    char name[32];
    if (cbits)
        grid_masks(b, 0, cbits, right, left, "gc");
    if (rbits)
        grid_masks(b, cbits, rbits, down, up, "gr");
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++)
            halo[(r + 1) * W + c + 1] = board[r * w + c];
    for (int r = 0; r < h; r++) {
        uint32_t last = board[r * w + w - 1], first = board[r * w];
        snprintf(name, sizeof name, "hw%d", r);
        halo[(r + 1) * W] = cbits ? grid_gather(b, name, last, 0, cbits, left) : last;
        snprintf(name, sizeof name, "he%d", r);
        halo[(r + 1) * W + w + 1] = cbits ? grid_gather(b, name, first, 0, cbits, right) : first;
    }
    for (int c = 0; c < W; c++) {
        uint32_t last = halo[h * W + c], first = halo[W + c];
        snprintf(name, sizeof name, "hn%d", c);
        halo[c] = rbits ? grid_gather(b, name, last, cbits, rbits, up) : last;
        snprintf(name, sizeof name, "hs%d", c);
        halo[(h + 1) * W + c] = rbits ? grid_gather(b, name, first, cbits, rbits, down) : first;
    }
    life_generation(b, board, halo, w, h);
    free(halo);
    ts_b_finish(b, p);
}

/* ---- running count over a cube (reference src/gen.c:325-370) -------------------------------- */

void ts_gen_rank(ts_program *p, int dims)
{
    ts_builder *b = ts_b_new();
    int w = dims + 1;
    uint32_t x = ts_b_cell(b, "x"), pc[TS_CUBE_MAX_DIMS + 1], t[2][TS_CUBE_MAX_DIMS + 1],
             add[TS_CUBE_MAX_DIMS + 1];
    uint32_t zero = TS_CELL_ZERO, c = ts_b_cell(b, "rk_c"), a = ts_b_cell(b, "rk_a");
    char name[32];
    ts_b_input(b, x);
    for (int i = 0; i < w; i++) {
        pc[i] = named(b, "p%d", i, 0);
        t[dims % 2][i] = named(b, "t%d", i, 0);
        t[1 - dims % 2][i] = named(b, "u%d", i, 0);
    }
    for (int i = 0; i < w; i++) {
        ts_b_gate(b, TS_G_BUF, pc[i], &zero, 1);
        ts_b_gate(b, TS_G_BUF, t[0][i], i ? &zero : &x, 1);
        ts_b_gate(b, TS_G_BUF, t[1][i], &zero, 1);
    }
    for (int d = 0; d < dims; d++) {
        const uint32_t *cur = t[d % 2], *nxt = t[1 - d % 2];
        uint32_t id = named(b, "ID%d", d, 0), n[TS_CUBE_MAX_DIMS + 1];
        for (int i = 0; i <= d; i++) {
            snprintf(name, sizeof name, "rk%d_%d", d, i);
            n[i] = ts_b_link(b, name, cur[i], d);
        }
        ts_b_gate(b, TS_G_BUF, c, &zero, 1);
        for (int i = 0; i <= d; i++)
            ts_b_full_adder(b, nxt[i], c, cur[i], n[i], c);
        ts_b_gate(b, TS_G_BUF, nxt[d + 1], &c, 1);
        ts_b_gate(b, TS_G_BUF, c, &zero, 1);
        for (int i = 0; i < w; i++) {
            if (i <= d)
                ts_b_gate2(b, TS_G_AND, add[i] = a, n[i], id);
            else
                add[i] = zero;
            ts_b_full_adder(b, pc[i], c, pc[i], add[i], c);
        }
    }
    for (int i = 0; i < w; i++)
        ts_b_output(b, pc[i]);
    for (int i = 0; i < w; i++)
        ts_b_output(b, t[dims % 2][i]);
    ts_b_finish(b, p);
}

/* ---- transforms over the 512 machines (reference src/gen.c:372-449) ------------------------ */

/* w <- Walsh-Hadamard transform mod 2^bits: round d, u = u + v in the machine with bit d clear,
   v = u - v in the other, as x + n + c with x = w ^ IDd and carry in IDd. */
static void wht(ts_builder *b, const uint32_t *w, int bits, const char *tag)
{
    char name[64];
    snprintf(name, sizeof name, "%s_x", tag);
    uint32_t x = ts_b_cell(b, name);
    snprintf(name, sizeof name, "%s_c", tag);
    uint32_t c = ts_b_cell(b, name);
    for (int d = 0; d < TS_CUBE_DIMS; d++) {
        snprintf(name, sizeof name, "ID%d", d);
        uint32_t id = ts_b_cell(b, name);
        ts_b_gate(b, TS_G_BUF, c, &id, 1);
        for (int i = 0; i < bits; i++) {
            snprintf(name, sizeof name, "%s_n%d_%d", tag, d, i);
            uint32_t n = ts_b_link(b, name, w[i], d);
            ts_b_gate2(b, TS_G_XOR, x, w[i], id);
            ts_b_full_adder(b, w[i], c, x, n, c);
        }
    }
}

void ts_gen_wht(ts_program *p, int bits)
{
    ts_builder *b = ts_b_new();
    uint32_t f = ts_b_cell(b, "f"), one = TS_CELL_ONE;
    uint32_t *w = malloc((size_t)bits * sizeof *w);
    ts_b_input(b, f);
    for (int i = 0; i < bits; i++) {
        w[i] = named(b, "w%d", i, 0);
        ts_b_gate(b, TS_G_BUF, w[i], i == 0 ? &one : &f, 1);
    }
    wht(b, w, bits, "a");
    for (int i = 0; i < bits; i++)
        ts_b_output(b, w[i]);
    ts_b_finish(b, p);
    free(w);
}

/* acc = x * y mod 2^bits */
static void multiply_low(ts_builder *b, uint32_t *acc, const uint32_t *x, const uint32_t *y, int bits)
{
    uint32_t pp = ts_b_cell(b, "m_pp"), carry = ts_b_cell(b, "m_carry");
    for (int k = 0; k < bits; k++)
        ts_b_gate2(b, TS_G_AND, acc[k], x[k], y[0]);
    for (int j = 1; j < bits; j++)
        for (int k = j; k < bits; k++) {
            ts_b_gate2(b, TS_G_AND, pp, x[k - j], y[j]);
            if (k == j)
                ts_b_half_adder(b, acc[k], carry, acc[k], pp);
            else
                ts_b_full_adder(b, acc[k], carry, acc[k], pp, carry);
        }
}

void ts_gen_xorconv(ts_program *p)
{
    enum { BITS = 20 };
    ts_builder *b = ts_b_new();
    uint32_t f = ts_b_cell(b, "f"), g = ts_b_cell(b, "g"), zero = TS_CELL_ZERO;
    uint32_t fw[BITS], gw[BITS], pw[BITS];
    ts_b_input(b, f);
    ts_b_input(b, g);
    for (int i = 0; i < BITS; i++) {
        fw[i] = named(b, "F%d", i, 0);
        ts_b_gate(b, TS_G_BUF, fw[i], i == 0 ? &f : &zero, 1);
        gw[i] = named(b, "G%d", i, 0);
        ts_b_gate(b, TS_G_BUF, gw[i], i == 0 ? &g : &zero, 1);
        pw[i] = named(b, "P%d", i, 0);
    }
    wht(b, fw, BITS, "F");
    wht(b, gw, BITS, "G");
    multiply_low(b, pw, fw, gw, BITS);
    wht(b, pw, BITS, "P");
    for (int i = 0; i < 10; i++) {
        uint32_t hc = named(b, "h%d", i, 0);
        ts_b_gate(b, TS_G_BUF, hc, &pw[9 + i], 1);
        ts_b_output(b, hc);
    }
    ts_b_finish(b, p);
}

/* ---- the cube library's data movers, built from gates (semantics of lib/cube.inc:389-575) -- */

int ts_benes_dim(int round)
{
    return round < TS_CUBE_DIMS ? round : 2 * TS_CUBE_DIMS - 2 - round;
}

/* v = c ? v of the machine across d : v, 8 bits; fresh links per round */
static void swap8(ts_builder *b, const uint32_t *v, uint32_t c, int d, int round)
{
    char name[48];
    uint32_t l[8], nc = named(b, "pm_nc%d", round, 0);
    for (int i = 0; i < 8; i++) {
        snprintf(name, sizeof name, "pm_l%d_%d", round, i);
        l[i] = ts_b_link(b, name, v[i], d);
    }
    ts_b_gate(b, TS_G_NOT, nc, &c, 1);
    for (int i = 0; i < 8; i++)
        ts_b_mux(b, v[i], c, nc, l[i], v[i]);
}

void ts_gen_perm8(ts_program *p)
{
    ts_builder *b = ts_b_new();
    uint32_t v[8], c[TS_BENES_ROUNDS];
    for (int i = 0; i < 8; i++)
        ts_b_input(b, v[i] = named(b, "v%d", i, 0));
    for (int s = 0; s < TS_BENES_ROUNDS; s++)
        ts_b_input(b, c[s] = named(b, "c%d", s, 0));
    for (int s = 0; s < TS_BENES_ROUNDS; s++)
        swap8(b, v, c[s], ts_benes_dim(s), s);
    for (int i = 0; i < 8; i++)
        ts_b_output(b, v[i]);
    ts_b_finish(b, p);
}

/* compare-exchange across dim j: the machine with (bit j of its number) ^ dir = 1 keeps the
   larger value, the other the smaller (lib/cube.inc:389-407) */
static void compare_exchange8(ts_builder *b, const uint32_t *v, int j, uint32_t dir, int round)
{
    char name[48];
    uint32_t n[8], one = TS_CELL_ONE;
    uint32_t g = named(b, "so_g%d", round, 0), d = ts_b_cell(b, "so_d"), nd = ts_b_cell(b, "so_nd");
    uint32_t want = ts_b_cell(b, "so_w"), keep = ts_b_cell(b, "so_k"), nkeep = ts_b_cell(b, "so_nk");
    for (int i = 0; i < 8; i++) {
        snprintf(name, sizeof name, "so_l%d_%d", round, i);
        n[i] = ts_b_link(b, name, v[i], j);
    }
    /* g = v >= n: from the lowest bit up, a differing bit decides */
    ts_b_gate(b, TS_G_BUF, g, &one, 1);
    for (int i = 0; i < 8; i++) {
        ts_b_gate2(b, TS_G_XOR, d, v[i], n[i]);
        ts_b_gate(b, TS_G_NOT, nd, &d, 1);
        ts_b_mux(b, g, d, nd, v[i], g);
    }
    ts_b_gate2(b, TS_G_XOR, want, named(b, "ID%d", j, 0), dir);
    ts_b_gate2(b, TS_G_XNOR, keep, want, g);
    ts_b_gate(b, TS_G_NOT, nkeep, &keep, 1);
    for (int i = 0; i < 8; i++)
        ts_b_mux(b, v[i], keep, nkeep, v[i], n[i]);
}

void ts_gen_sort8(ts_program *p)
{
    ts_builder *b = ts_b_new();
    uint32_t v[8];
    for (int i = 0; i < 8; i++)
        ts_b_input(b, v[i] = named(b, "v%d", i, 0));
    int round = 0;
    for (int blk = 1; blk <= TS_CUBE_DIMS; blk++) {
        uint32_t dir = blk < TS_CUBE_DIMS ? named(b, "ID%d", blk, 0) : TS_CELL_ZERO;
        for (int j = blk - 1; j >= 0; j--)
            compare_exchange8(b, v, j, dir, round++);
    }
    for (int i = 0; i < 8; i++)
        ts_b_output(b, v[i]);
    ts_b_finish(b, p);
}

/* ---- polynomial product mod (x^512 + 1, 12289) by the number-theoretic transform ------------
   (reference src/gen.c:456-669). Machine k of a 512-machine cube holds coefficient k; numbers mod
   q = 12289 are 14-bit vectors. The butterflies of round d pair machines k and k ^ 2^d through
   links of dimension d; the twiddle of each machine and round is an input. The gate order below
   is the reference's, so the program is the reference's instruction for instruction. */

enum { NTT_BITS = 14, NTT_Q = 12289, NO_CARRY_OUT = -1 };

/* out = x + y + cin over n bits, the carry into cout unless cout < 0 */
static void adder_n(ts_builder *b, const uint32_t *out, const uint32_t *x, const uint32_t *y,
                    uint32_t cin, int n, int64_t cout)
{
    uint32_t c = ts_b_cell(b, "add_c");
    ts_b_gate(b, TS_G_BUF, c, &cin, 1);
    for (int i = 0; i < n; i++)
        ts_b_full_adder(b, out[i], c, x[i], y[i], c);
    if (cout >= 0) {
        uint32_t co = (uint32_t)cout;
        ts_b_gate(b, TS_G_BUF, co, &c, 1);
    }
}

/* out = sel ? x1 : x0, bit by bit through x0 ^ ((x0 ^ x1) & sel) */
static void select_n(ts_builder *b, const uint32_t *out, uint32_t sel, const uint32_t *x1,
                     const uint32_t *x0, int n)
{
    uint32_t t = ts_b_cell(b, "mux_t");
    for (int i = 0; i < n; i++) {
        ts_b_gate2(b, TS_G_XOR, t, x0[i], x1[i]);
        ts_b_gate2(b, TS_G_AND, t, t, sel);
        ts_b_gate2(b, TS_G_XOR, out[i], x0[i], t);
    }
}

/* r = x y 2^-14 mod q (Montgomery reduction), every value in 0 .. q - 1. With t = x y,
   m = t q^-1 mod 2^14 = t (1 + 2^12) mod 2^14 (q = 3 2^12 + 1), and t - m q is divisible by
   2^14: r = t / 2^14 - (m q) / 2^14, where (m q) / 2^14 = (m + m / 2^12 + 2 m) / 4 rounded down
   (the low 12 bits of m q are those of m), plus q when that difference is negative. */
static void montgomery(ts_builder *b, const uint32_t *r, const uint32_t *x, const uint32_t *y)
{
    const uint32_t zero = TS_CELL_ZERO, one = TS_CELL_ONE;
    uint32_t k = ts_b_cell(b, "mm_k");
    uint32_t t[2 * NTT_BITS], m[NTT_BITS], s[16], h[15], u[16], v[16], th[15], ns[15], qm[NTT_BITS];
    for (int i = 0; i < 2 * NTT_BITS; i++)
        t[i] = named(b, "mm_t%d", i, 0);
    for (int i = 0; i < 16; i++)
        s[i] = named(b, "mm_s%d", i, 0);
    for (int i = 0; i < 15; i++)
        h[i] = named(b, "mm_h%d", i, 0);
    multiply(b, t, x, y, NTT_BITS);
    memcpy(m, t, 12 * sizeof *m);
    m[12] = ts_b_cell(b, "mm_m12");
    m[13] = ts_b_cell(b, "mm_m13");
    ts_b_gate2(b, TS_G_AND, k, t[12], t[0]);  /* bits 12, 13 of t + (t << 12) */
    ts_b_gate2(b, TS_G_XOR, m[12], t[12], t[0]);
    ts_b_gate2(b, TS_G_XOR, m[13], t[13], t[1]);
    ts_b_gate2(b, TS_G_XOR, m[13], m[13], k);
    for (int i = 0; i < 16; i++) {  /* s = m + (m >> 12) */
        u[i] = i < NTT_BITS ? m[i] : zero;
        v[i] = i < 2 ? m[12 + i] : zero;
    }
    adder_n(b, s, u, v, zero, 16, NO_CARRY_OUT);
    for (int i = 0; i < 16; i++)  /* s += 2 m */
        v[i] = (i >= 1 && i <= NTT_BITS) ? m[i - 1] : zero;
    adder_n(b, s, s, v, zero, 16, NO_CARRY_OUT);
    for (int i = 0; i < 15; i++) {  /* h = (t >> 14) + ~(s >> 2) + 1, 15 bits */
        th[i] = i < NTT_BITS ? t[NTT_BITS + i] : zero;
        if (i < NTT_BITS) {
            ns[i] = named(b, "mm_n%d", i, 0);
            ts_b_gate(b, TS_G_NOT, ns[i], &s[2 + i], 1);
        } else {
            ns[i] = one;
        }
    }
    adder_n(b, h, th, ns, one, 15, NO_CARRY_OUT);
    for (int i = 0; i < NTT_BITS; i++)  /* r = h + (h < 0 ? q : 0) */
        qm[i] = (NTT_Q >> i) & 1 ? h[14] : zero;
    adder_n(b, r, h, qm, zero, NTT_BITS, NO_CARRY_OUT);
}

/* out = sel ? x - y : x + y mod q */
static void add_or_sub_q(ts_builder *b, const uint32_t *out, uint32_t sel, const uint32_t *x,
                         const uint32_t *y)
{
    const uint32_t zero = TS_CELL_ZERO, one = TS_CELL_ONE;
    uint32_t carry = ts_b_cell(b, "as_cy");
    uint32_t s[15], d[15], c[15], xx[15], w[15];
    for (int i = 0; i < 15; i++) {
        s[i] = named(b, "as_s%d", i, 0);
        d[i] = named(b, "as_d%d", i, 0);
        xx[i] = i < NTT_BITS ? x[i] : zero;
        c[i] = (NTT_Q >> i) & 1 ? sel : zero;
    }
    adder_n(b, s, xx, c, zero, 15, NO_CARRY_OUT);  /* x, plus q when subtracting */
    for (int i = 0; i < NTT_BITS; i++) {
        w[i] = named(b, "as_v%d", i, 0);
        ts_b_gate2(b, TS_G_XOR, w[i], y[i], sel);
    }
    w[NTT_BITS] = sel;
    adder_n(b, s, s, w, sel, 15, NO_CARRY_OUT);  /* + y, or + ~y + 1: below 2 q */
    for (int i = 0; i < 15; i++)
        c[i] = ((32768 - NTT_Q) >> i) & 1 ? one : zero;
    adder_n(b, d, s, c, zero, 15, carry);  /* s - q; the carry says s >= q */
    select_n(b, out, carry, d, s, NTT_BITS);
}

/* a <- its negacyclic transform: rounds d = 8 .. 0 of Cooley-Tukey butterflies between machines
   u and u + 2^d, (a_u, a_v) -> (a_u + z a_v, a_u - z a_v), z from input z[d] */
static void ntt_forward(ts_builder *b, const uint32_t *a, uint32_t z[TS_CUBE_DIMS][NTT_BITS],
                        const char *tag)
{
    uint32_t t[NTT_BITS], x[NTT_BITS], y[NTT_BITS], la[NTT_BITS], lt[NTT_BITS];
    char name[64];
    for (int i = 0; i < NTT_BITS; i++) {
        t[i] = named(b, "nt_t%d", i, 0);
        x[i] = named(b, "nt_x%d", i, 0);
        y[i] = named(b, "nt_y%d", i, 0);
    }
    for (int d = TS_CUBE_DIMS - 1; d >= 0; d--) {
        uint32_t id = named(b, "ID%d", d, 0);
        montgomery(b, t, a, z[d]);  /* z a, used where bit d is set */
        for (int i = 0; i < NTT_BITS; i++) {
            snprintf(name, sizeof name, "%s_a%d_%d", tag, d, i);
            la[i] = ts_b_link(b, name, a[i], d);
            snprintf(name, sizeof name, "%s_t%d_%d", tag, d, i);
            lt[i] = ts_b_link(b, name, t[i], d);
        }
        select_n(b, x, id, la, a, NTT_BITS);  /* a_u */
        select_n(b, y, id, t, lt, NTT_BITS);  /* z a_v */
        add_or_sub_q(b, a, id, x, y);
    }
}

/* the way back without the factor 1/512: rounds d = 0 .. 8 of Gentleman-Sande butterflies,
   (a_u, a_v) -> (a_u + a_v, (a_u - a_v) w), w from input w[d] (1 in machine u) */
static void ntt_inverse(ts_builder *b, const uint32_t *a, uint32_t w[TS_CUBE_DIMS][NTT_BITS],
                        const char *tag)
{
    uint32_t x[NTT_BITS], y[NTT_BITS], la[NTT_BITS];
    char name[64];
    for (int i = 0; i < NTT_BITS; i++) {
        x[i] = named(b, "nt_x%d", i, 0);
        y[i] = named(b, "nt_y%d", i, 0);
    }
    for (int d = 0; d < TS_CUBE_DIMS; d++) {
        uint32_t id = named(b, "ID%d", d, 0);
        for (int i = 0; i < NTT_BITS; i++) {
            snprintf(name, sizeof name, "%s_a%d_%d", tag, d, i);
            la[i] = ts_b_link(b, name, a[i], d);
        }
        select_n(b, x, id, la, a, NTT_BITS);  /* a_u */
        select_n(b, y, id, a, la, NTT_BITS);  /* a_v */
        add_or_sub_q(b, a, id, x, y);
        montgomery(b, a, a, w[d]);
    }
}

static uint32_t pow_mod_q(uint32_t x, uint32_t e)
{
    uint64_t r = 1, p = x % NTT_Q;
    while (e) {
        if (e & 1)
            r = r * p % NTT_Q;
        p = p * p % NTT_Q;
        e >>= 1;
    }
    return (uint32_t)r;
}

static unsigned reverse9(unsigned k)
{
    unsigned r = 0;
    for (int i = 0; i < TS_CUBE_DIMS; i++)
        if (k >> i & 1)
            r |= 1u << (TS_CUBE_DIMS - 1 - i);
    return r;
}

/* The twiddle inputs of machine k in round d, in Montgomery form (times 2^14 mod q): forward
   zf = psi^rev9(2^(8-d) + k / 2^(d+1)), psi a primitive 1024th root of 1; inverse zi = zf^-1 on
   the machines with bit d set, 1 elsewhere (reference src/gen.c:617-632). */
void ts_ntt_twiddles(uint16_t zf[TS_CUBE_DIMS][TS_CUBE_SIZE], uint16_t zi[TS_CUBE_DIMS][TS_CUBE_SIZE])
{
    const uint32_t mont = (1u << NTT_BITS) % NTT_Q;
    uint32_t g = 2;  /* the least generator: q - 1 = 2^12 * 3 */
    while (pow_mod_q(g, (NTT_Q - 1) / 2) == 1 || pow_mod_q(g, (NTT_Q - 1) / 3) == 1)
        g++;
    const uint32_t psi = pow_mod_q(g, (NTT_Q - 1) / 1024);
    for (int d = 0; d < TS_CUBE_DIMS; d++)
        for (unsigned k = 0; k < TS_CUBE_SIZE; k++) {
            uint32_t z = pow_mod_q(psi, reverse9((256u >> d) + (k >> (d + 1))));
            zf[d][k] = (uint16_t)((uint64_t)z * mont % NTT_Q);
            zi[d][k] = (uint16_t)((k >> d) & 1 ? (uint64_t)pow_mod_q(z, NTT_Q - 2) * mont % NTT_Q
                                               : mont);
        }
}

void ts_gen_polymul(ts_program *p)
{
    ts_builder *b = ts_b_new();
    uint32_t a[NTT_BITS], bv[NTT_BITS], c[NTT_BITS], wa[NTT_BITS], wb[NTT_BITS], f[NTT_BITS];
    uint32_t zf[TS_CUBE_DIMS][NTT_BITS], zi[TS_CUBE_DIMS][NTT_BITS];
    for (int i = 0; i < NTT_BITS; i++)
        ts_b_input(b, a[i] = named(b, "a%d", i, 0));
    for (int i = 0; i < NTT_BITS; i++)
        ts_b_input(b, bv[i] = named(b, "b%d", i, 0));
    for (int d = 0; d < TS_CUBE_DIMS; d++)
        for (int i = 0; i < NTT_BITS; i++)
            ts_b_input(b, zf[d][i] = named(b, "zf%d_%d", d, i));
    for (int d = 0; d < TS_CUBE_DIMS; d++)
        for (int i = 0; i < NTT_BITS; i++)
            ts_b_input(b, zi[d][i] = named(b, "zi%d_%d", d, i));
    /* 2^28 / 512 mod q: two Montgomery factors 2^-14 and the transform's 512 undone at the end */
    const uint32_t fc = (uint32_t)((uint64_t)pow_mod_q((1u << NTT_BITS) % NTT_Q, 2) *
                                   pow_mod_q(TS_CUBE_SIZE, NTT_Q - 2) % NTT_Q);
    for (int i = 0; i < NTT_BITS; i++) {
        c[i] = named(b, "c%d", i, 0);
        wa[i] = named(b, "wa%d", i, 0);
        ts_b_gate(b, TS_G_BUF, wa[i], &a[i], 1);
        wb[i] = named(b, "wb%d", i, 0);
        ts_b_gate(b, TS_G_BUF, wb[i], &bv[i], 1);
        f[i] = (fc >> i) & 1 ? TS_CELL_ONE : TS_CELL_ZERO;
    }
    ntt_forward(b, wa, zf, "fa");
    ntt_forward(b, wb, zf, "fb");
    montgomery(b, wa, wa, wb);  /* pointwise product */
    ntt_inverse(b, wa, zi, "ia");
    montgomery(b, c, wa, f);
    for (int i = 0; i < NTT_BITS; i++)
        ts_b_output(b, c[i]);
    ts_b_finish(b, p);
}
