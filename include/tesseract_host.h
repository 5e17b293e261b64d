/* Host side of the engine: everything that produces or describes the programs the GPU engine runs.
   Program helpers, the program writer and the .bench netlist loader, verify/search specifications,
   random programs for the cycle tester, and Beneš routing. No device code.

   Each function names the reference routine whose behaviour it reproduces. */
#ifndef TESSERACT_HOST_H
#define TESSERACT_HOST_H

#include <stdio.h>

#include "tesseract_isa.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- program helpers (reference src/asm.c:767-820) ------------------------------------------ */

void     ts_program_free(ts_program *p);
int      ts_program_find_cell(const ts_program *p, const char *name, uint32_t *addr);
int      ts_program_readonly(const ts_program *p, uint32_t cell);
/* The plain/constant/id cell a read of `cell` reaches through links; *mask = XOR of the dims. */
uint32_t ts_program_resolve(const ts_program *p, uint32_t cell, unsigned *mask);
/* 9, or one more than the highest dimension an id or link cell names. */
unsigned ts_program_dims(const ts_program *p);
int      ts_program_id_name(const char *name);  /* ID0..ID15 -> bit, else -1 */
unsigned ts_addr_bits(uint32_t ncells);
/* A deep copy (so a caller's program can be freed while the copy lives). */
int      ts_program_copy(ts_program *dst, const ts_program *src);
/* A program from flat arrays (the form program files store), in memory ts_program_free releases:
   code[i] = addr | op << 31; names = ncells NUL-terminated names one after another; kind, dim, src
   per cell, or all three NULL for a program without id and link cells. Checks that every address
   is a cell and every special cell is well formed; returns 0 with a message otherwise. */
int      ts_program_from_arrays(ts_program *out, const uint32_t *code, size_t len, uint32_t ncells,
                                const char *names, size_t names_bytes, const uint32_t *inputs,
                                size_t ninputs, const uint32_t *outputs, size_t noutputs,
                                const uint8_t *kind, const uint8_t *dim, const uint32_t *src,
                                char *err, size_t errlen);

/* ---- the program writer: cells, gates, registers (the reference's src/build.h interface) ------ */

enum { TS_G_BUF, TS_G_NOT, TS_G_AND, TS_G_NAND, TS_G_OR, TS_G_NOR, TS_G_XOR, TS_G_XNOR, TS_G_COUNT };

typedef struct ts_builder ts_builder;

/* NULL if out of memory. A misuse (an unknown gate, a write to a read-only cell, a cell that does
   not exist) or running out of memory is recorded and reported by ts_b_finish. */
ts_builder *ts_b_new(void);
uint32_t    ts_b_cell(ts_builder *b, const char *name);  /* created on first use; ID0..ID15 are id cells */
int         ts_b_has(const ts_builder *b, const char *name);
void        ts_b_input(ts_builder *b, uint32_t cell);
void        ts_b_output(ts_builder *b, uint32_t cell);
/* out = gate(in[0..n)); the gate reads all its inputs before it writes out (out may be an input) */
void        ts_b_gate(ts_builder *b, int type, uint32_t out, const uint32_t *in, int n);
void        ts_b_gate2(ts_builder *b, int type, uint32_t out, uint32_t x, uint32_t y);
/* q <- d at the end of every pass, all registers at once */
void        ts_b_reg(ts_builder *b, uint32_t q, uint32_t d);
/* s = x ^ y, k = x & y (11 instructions; s may be x or y) */
void        ts_b_half_adder(ts_builder *b, uint32_t s, uint32_t k, uint32_t x, uint32_t y);
/* s = x ^ y ^ c, k = the carry (19 instructions; s may be x or y, k may be c) */
void        ts_b_full_adder(ts_builder *b, uint32_t s, uint32_t k, uint32_t x, uint32_t y, uint32_t c);
/* the new read-only cell `name` reads `src` of the machine across dimension dim */
uint32_t    ts_b_link(ts_builder *b, const char *name, uint32_t src, int dim);
size_t      ts_b_len(const ts_builder *b);
/* Raw instructions on any existing cell (a STO to a read-only cell is kept; machines ignore it). */
void        ts_b_nand(ts_builder *b, uint32_t cell);
void        ts_b_sto(ts_builder *b, uint32_t cell);
/* Emits the registers, fills out and frees the builder; 0 with the first recorded misuse in err. */
int         ts_b_finish(ts_builder *b, ts_program *out, char *err, size_t errlen);
const char *ts_gate_name(int type);
int         ts_gate_type(const char *name);  /* -1 if unknown; BUF and BUFF both mean TS_G_BUF */

/* ---- specifications (reference src/spec.h) --------------------------------------------------- */

typedef struct ts_spec ts_spec;

ts_spec *ts_spec_parse(const char *text, const ts_program *p, int inputs_allowed, char *err,
                       size_t errlen);
void     ts_spec_free(ts_spec *s);
int      ts_spec_width(const ts_spec *s);
uint32_t ts_spec_out_cell(const ts_spec *s, int bit);
uint64_t ts_spec_value(const ts_spec *s, uint64_t x);
int      ts_vec_resolve(const ts_program *p, const uint32_t *list, size_t n, const char *name,
                        int *idx);

/* The lowered right-hand side, for engines that evaluate it themselves: an RPN program over
   64-bit words. Variables are input vectors; bit b of variable k is bit var_pos[k][b] of the
   assignment x. */
enum { TS_T_NUM, TS_T_VAR, TS_T_ADD, TS_T_SUB, TS_T_MUL, TS_T_AND, TS_T_OR, TS_T_XOR, TS_T_SHL,
       TS_T_SHR, TS_T_NOT, TS_T_NEG };
typedef struct {
    int      op;
    uint64_t v;
} ts_spec_item;
int                 ts_spec_nrpn(const ts_spec *s);
const ts_spec_item *ts_spec_rpn(const ts_spec *s);
int                 ts_spec_nvars(const ts_spec *s);
int                 ts_spec_var_bits(const ts_spec *s, int k);
int                 ts_spec_var_pos(const ts_spec *s, int k, int bit);

/* ---- enumeration results (reference src/enum.h) ----------------------------------------------- */

enum { TS_EN_VERIFY = 0, TS_EN_EQUIV = 1, TS_EN_SEARCH = 2 };
#define TS_EN_SHOW   16
#define TS_EN_MAX_IN 46
#define TS_EN_CHUNK  64  /* passes of 512 assignments per hand-out (reference src/enum.c:11) */

/* Layout-identical to the reference EnumResult (src/enum.h:26-35). */
typedef struct {
    uint64_t space;
    uint64_t tried;
    uint64_t hits;
    uint64_t hit[TS_EN_SHOW];
    int      nhit;
    double   seconds;
    uint64_t ops;
    int      threads;
} ts_enum_result;

void ts_enum_add_hit(ts_enum_result *r, uint64_t x);
/* equiv: in_map[k] / out_map[k] = index in b's inputs / outputs of a's input / output k, by
   case-insensitive name. Returns 0 with the reference's message on a mismatch. */
int  ts_enum_match(const ts_program *a, const ts_program *b, int *in_map, int *out_map, char *err,
                   size_t errlen);

/* ---- random programs and start states (reference src/tester.c:51-91) ------------------------ */

uint64_t ts_splitmix64(uint64_t *state);
void     ts_random_program(ts_program *p, uint32_t cells, size_t len, uint64_t seed);
uint64_t ts_t1_program_seed(uint64_t seed, uint32_t cells, uint64_t k);
/* Start states for `groups` groups of 512 machines: mem[g * ncells + cell] and rr[g] lines.
   Group 0 is the reference's t1_start(seed) exactly; later groups continue the same splitmix64
   stream (cells, then RR, group after group). Read-only cells are left as they are. */
void     ts_t1_start(const ts_program *p, uint64_t seed, unsigned groups, ts_line *mem,
                     ts_line *rr);
#define TS_T1_BUCKETS 64
int      ts_t1_bucket(uint64_t lam);

/* ---- tester 1 summaries (reference src/tester.h:25-57) ----------------------------------------- */

#define TS_T1_SIZES 32

/* Layout-identical to the reference T1Size (src/tester.h:30-37): the results of one machine size. */
typedef struct {
    uint32_t cells;
    size_t   len;
    uint64_t programs, starts, passes;
    uint64_t unfinished;             /* programs cut off by a time limit, not in hist */
    uint64_t hist[TS_T1_BUCKETS];
    uint64_t longest, longest_seed;  /* longest cycle found, and the seed of the program with it */
} ts_t1_size;

/* Layout-identical to the reference T1Config (src/tester.h:39-50). threads, idle and excluded
   choose host processors and mean nothing to a GPU run. */
typedef struct {
    uint32_t             cells[TS_T1_SIZES];
    int                  nsizes;
    double               ratio;       /* instructions per cell */
    uint64_t             programs;    /* per size; 0 = until `seconds` are up */
    double               seconds;
    uint64_t             max_passes;
    uint64_t             seed;
    int                  threads;
    int                  idle;
    const unsigned char *excluded;
} ts_t1_config;

/* round(cells * ratio), at least 1 (reference src/tester.c:176-180) */
size_t ts_t1_program_len(const ts_t1_config *cfg, uint32_t cells);
/* to += from; the longest cycle moves only if strictly longer (reference src/tester.c:228-240) */
void   ts_t1_add_size(ts_t1_size *to, const ts_t1_size *from);
/* The raw, machine-readable form of a run (reference t1_print_raw, src/tester.c:355-370): a
   "T1RUN host threads seconds max_passes" line, then one "T1 ..." line per size. Returns a
   malloc'd string (free it), NULL if out of memory. */
char  *ts_t1_raw(const ts_t1_size *s, int n, const char *host, double seconds, int threads,
                 uint64_t max_passes);
/* Adds up raw outputs read from files (reference t1_merge, src/tester.c:372-431): sizes with the
   same (cells, len) are summed, threads summed, seconds the largest, max_passes the last seen;
   lines that are neither form are skipped. Returns the number of sizes, -1 if a file cannot be
   read. s has room for TS_T1_SIZES. */
int    ts_t1_merge(char **files, int nfiles, ts_t1_size *s, double *seconds, int *threads,
                   uint64_t *max_passes);

/* ---- raw enumeration results: shares run apart, merged (reference src/work.c:341-371,435-441) -- */

/* "EN mode space tried hits seconds threads ops" and one "ENHIT x" (hex) per shown hit, the
   reference's --raw output of verify / equiv / search. Returns a malloc'd string (free it). */
char *ts_enum_raw(int mode, const ts_enum_result *r);
/* Adds up such outputs read from files into r: tried, hits, ops and threads summed, seconds the
   largest, space the last seen, hits kept as the 16 smallest. Returns 0 with a message if a file
   cannot be opened. */
int   ts_enum_merge(char **files, int nfiles, ts_enum_result *r, char *err, size_t errlen);

/* ---- Beneš routing (reference src/benes.h) --------------------------------------------------- */

#define TS_BENES_ROUNDS (2 * TS_CUBE_DIMS - 1)
int  ts_benes_dim(int round);
void ts_benes_route(const uint16_t *dest, uint8_t ctrl[TS_BENES_ROUNDS][TS_CUBE_SIZE]);

/* ---- ISCAS .bench netlists (reference src/net.c) ---------------------------------------------- */

int  ts_net_parse(const char *text, const char *name, ts_program *out, char *err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif
