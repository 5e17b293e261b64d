/* ORACLE — test infrastructure only.

   A CPU restatement of the reference's golden model and of the closed forms its tests check
   against. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and
   only as the checker: the product (tesseract_amd) never calls into it.

   Every function cites the reference routine it restates. How each one is pinned:
     oracle_run_lines      ref_step/ref_cube_step semantics (src/ref.c:49-67,117-130); pinned by
                           tests against the reference's own build (oracle/_ref) on random
                           programs with id and link cells, and by closed forms (adders,
                           multipliers, toyhash, Life) that a dropped term or a wrong index fails.
     oracle_steps_lines    the same machine one instruction at a time (ref_cube_step, src/ref.c:
                           119-130), from any pc, wrapping into the next pass; pinned against
                           oracle/_ref's ref_cube_step and against oracle_run_lines for whole
                           passes.
     oracle_enum           enum_run's assignment mapping and hit rules (src/enum.c:150-213,
                           src/enum.h:3-5); pinned by the reference's known answers (adder8
                           131,072 tried / 0 hits and 65,536 hits with hit[] = 1,3,..,31), shares
                           that add up, and brute-force toyhash preimage counts.
     oracle_brent_lines    t1_cycles' schedule (src/tester.c:96-164) on the reference machine;
                           pinned by exact cycle lengths from exhaustive state enumeration on tiny
                           programs (oracle_exact_cycle) and by oracle/_ref's t1_cycles.
     oracle_life_scalar    life_scalar (src/gen.c:686-700): textbook B3/S23 on a torus; pinned by
                           known oscillators (blinker period 2, block still life, glider shift).
     oracle_life_words     the same rule bit-parallel, 64 cells per word, for full-size boards;
                           pinned against oracle_life_scalar on random boards of several sizes.
     oracle_benes_apply    benes_apply (src/benes.c:77-86); pinned by routing a permutation and
                           checking every value lands at dest[k].
     oracle_toyhash        toyhash (src/gen.c:107-121); pinned against oracle/_ref's toyhash and the
                           published one-preimage target 0x165c2a07 over 2^32 (README.md:67). */
#ifndef TESSERACT_ORACLE_H
#define TESSERACT_ORACLE_H

#include "tesseract_isa.h"

/* Runs `passes` passes of every machine of the state on the reference machine (one byte per
   cell, one instruction at a time, cubes in lockstep). groups = number of 512-machine groups; it
   must be a multiple of the program's cube size in groups (2^(dims-9)); consecutive blocks form
   independent cubes. mem[g * ncells + c] / rr[g] are 512-machine lines, updated in place. */
int64_t  oracle_steps_lines(const ts_program *p, unsigned groups, ts_line *mem, ts_line *rr,
                            uint64_t pc, uint64_t steps);
int      oracle_run_lines(const ts_program *p, unsigned groups, ts_line *mem, ts_line *rr,
                          uint64_t passes);

/* Power-on state (all 0, ONE = 1, id cells = bit d of the machine number, RR = 0). */
void     oracle_reset_lines(const ts_program *p, unsigned groups, ts_line *mem, ts_line *rr);

/* The reference's check protocol (src/check.c:66-77): every writable cell and RR of every group
   from splitmix64(seed), group -> cell -> 8 words, then the 8 RR words; id cells keep their
   pattern. */
void     oracle_random_lines(const ts_program *p, unsigned groups, uint64_t seed, ts_line *mem,
                             ts_line *rr);

/* Exhaustive enumeration on the reference machine, assignment by assignment.
   mode 0 verify: outputs = spec (spec_value given as a table over x, width `width`, out_cells the
   cells of the left side lowest bit first); mode 1 equiv: a's outputs vs b's outputs (b_in[k],
   b_out[k] = b's cell for a's input / output k); mode 2 search: outputs (out_cells) == target.
   Hits: count and the 16 smallest x. Shares as in enum_run (chunks of 64 passes of 512). */
typedef struct {
    uint64_t space, tried, hits, hit[16];
    int      nhit;
} oracle_enum_result;

int      oracle_enum(int mode, const ts_program *a, const ts_program *b, const uint32_t *b_in,
                     const uint32_t *b_out, const uint32_t *out_cells, int width,
                     const uint64_t *expect, uint64_t target, unsigned ticks, int part,
                     int part_end, int parts, oracle_enum_result *r);

/* Brent's cycle search for every machine of one group from the given start state, with the
   reference's schedule; lam[512] = cycle length or 0 if not found within max_passes. */
uint64_t oracle_brent_lines(const ts_program *p, const ts_line *mem, const ts_line *rr,
                            uint64_t max_passes, uint64_t *lam);

/* Exact cycle length of one machine by remembering every state (at most 2^24 states). */
uint64_t oracle_exact_cycle(const ts_program *p, const ts_line *mem, const ts_line *rr,
                            unsigned lane);

void     oracle_life_scalar(uint8_t *board, int w, int h);
/* The same rule on 64 cells per word (cell c of a row is bit c % 64 of word c / 64). */
void     oracle_life_words(uint64_t *board, int words_per_row, int h, int gens);
void     oracle_benes_apply(const uint8_t *ctrl /* [17][512] */, uint64_t *v /* [512] */);
uint64_t oracle_toyhash(uint64_t x, int bits, int rounds);

#endif
