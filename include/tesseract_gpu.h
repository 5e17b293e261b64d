/* The GPU engine: the C ABI a host registers as `--engine gpu`.

   Plain C types only: programs are the layout-identical ts_program (reference src/isa.h:36-46),
   state crosses as the reference's 64-byte lines (src/lanes.h:13-15), results as ts_enum_result
   (src/enum.h:26-35). Every function returns 1 on success and 0 with a message in err on failure;
   nothing in the library calls exit() or asserts on caller input.

   Threading: one gpu_engine per device owns its stream; gpu_code is immutable after gpu_compile
   and may be shared; a gpu_state has one owner. Calls on one engine are serialised by the caller.
   A gpu_code keeps no pointer into the caller's program. */
#ifndef TESSERACT_GPU_H
#define TESSERACT_GPU_H

#include "tesseract_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpu_engine gpu_engine;
typedef struct gpu_code   gpu_code;
typedef struct gpu_state  gpu_state;

/* Opens device `device` (0 = the first GPU). There is no CPU fallback: without a usable GPU this
   fails with the HIP error. */
int  gpu_engine_open(int device, gpu_engine **out, char *err, size_t errlen);
void gpu_engine_close(gpu_engine *e);
/* "gfx950, 256 CUs, ..." */
const char *gpu_engine_describe(const gpu_engine *e);

/* Tuning knobs of one compiled program; zero fields take the defaults. */
typedef struct {
    unsigned block;        /* threads per workgroup (multiple of 64) */
    unsigned slice_nodes;  /* AIG nodes per slice for segmented programs (0 = by cube size) */
    unsigned lut3;         /* 1 = map to v_bitop3_b32 (default), 2 = plain AND/NOT */
    unsigned grid;         /* workgroups for enumeration kernels (0 = default: 256 per CU) */
    unsigned graph_passes; /* segmented programs: passes per captured HIP graph (default 32,
                              1 = launch every segment directly) */
    unsigned tape;         /* enumerations: 1 = run the precompiled tape interpreter instead of
                              compiling a kernel (equiv and search with one tick; no hipRTC) */
    unsigned waves;        /* waves per SIMD the register allocation must allow (0 = default:
                              2 for a resident program, which caps it at 256 VGPRs, the
                              compiler's choice for segments and enumerations; 1 = the
                              compiler's choice everywhere) */
    unsigned max_waves;    /* enumerations: at most this many waves per SIMD run at once (each
                              workgroup reserves LDS it does not use; 0 = no cap) */
    unsigned sched;        /* the compiler's instruction scheduling: 1 = its default,
                              3 = max-ilp, 4 = iterative-minreg; 0 = auto: an enumeration of at
                              least 2^40 1-bit ops is compiled with 1 and 3, each timed on a sample
                              of its words, and the faster kept; otherwise 1. 2 (iterative-ilp) is
                              refused with an error: the compiler itself faults on some kernels
                              under it (the reference's PERM8, Life's segment kernel, the mul16
                              verify against a * b + (a & 1)) */
    unsigned group;        /* enumerations: 1 = a thread takes one word of 32 assignments at a
                              time; 2 = groups of 32 words, the part of the program no lane bit
                              (x bits 0..4) reaches evaluated once per group, bit-sliced over its
                              words; 0 = auto (groups for one tick, at least 24 inputs and at
                              least 32 such LUTs) */
    unsigned lds_links;    /* cubes of 128 to 1,024 words (2^12 to 2^15 machines): 1 = each cube
                              runs in one workgroup of that many threads, reads across the waves
                              of its cube through LDS (a barrier pair per level of such reads),
                              the pass loop in-kernel with the state in registers; 2 = as larger
                              cubes: segments at reads of values computed across waves (or the
                              value recomputed when its cone is small), slices of threads per
                              segment; 0 = the form a cost model of one cube's pass prefers (the
                              LUTs one thread evaluates, and the launches) */
} gpu_tune;

typedef struct {
    uint32_t segments;     /* launches per pass */
    uint32_t resident;     /* 1 = the pass loop runs in-kernel with the state in registers */
    uint32_t slices;       /* largest number of slices in a segment */
    uint64_t bitop3;       /* LUT instructions in the generated code (all slices) */
    uint64_t values;       /* values the generated code materialises */
    double   compile_seconds;
    uint32_t dims;         /* cube dimensions */
    uint64_t chain;        /* LUTs one thread evaluates in a pass: the resident kernel's pass, or
                              the largest slice of each segment, summed */
} gpu_code_info;

/* Compiles p to gfx950 code through hipRTC: the counterpart of
   jit_compile(const Program*, int literal, JitCode*, char **listing, char*, size_t)
   (reference src/jit.h:46). listing, if not NULL, receives the generated HIP source (free()). */
int  gpu_compile(gpu_engine *e, const ts_program *p, const gpu_tune *tune, gpu_code **out,
                 char **listing, char *err, size_t errlen);
void gpu_code_free(gpu_code *c);
int  gpu_code_stats(const gpu_code *c, gpu_code_info *info);

/* Device state for `groups` groups of 512 machines, in the power-on state (reference
   lane_init/lane_reset, src/lanes.c:110-192; lane_cube_init, src/lanes.c:341-370). For a cube
   program groups must be a multiple of its 2^(dims-9) groups; consecutive blocks are separate
   cubes. */
int  gpu_state_new(gpu_code *c, unsigned groups, gpu_state **out, char *err, size_t errlen);
void gpu_state_free(gpu_state *s);
int  gpu_state_reset(gpu_state *s, char *err, size_t errlen);
/* Copies ngroups groups from / to the reference layout: mem[(g - group0) * ncells + cell] and
   rr[g - group0] are lines, machine k of group g being bit k % 64 of w[k / 64] (the memory a
   LaneMachine holds, src/lanes.h:38-51). put ignores read-only cells; get returns ZERO as 0, ONE
   as all ones, id cells as their pattern and link cells as 0, as a LaneMachine holds them. */
int  gpu_state_put(gpu_state *s, unsigned group0, unsigned ngroups, const ts_line *mem,
                   const ts_line *rr, char *err, size_t errlen);
int  gpu_state_get(gpu_state *s, unsigned group0, unsigned ngroups, ts_line *mem, ts_line *rr,
                   char *err, size_t errlen);

/* `passes` whole passes of every machine of the state, cubes in lockstep: the counterpart of
   jit_run (src/jit.h:54) and jit_cube_run (src/jit.h:59). */
int  gpu_run(gpu_code *c, gpu_state *s, uint64_t passes, char *err, size_t errlen);
/* One machine's cell, machine = group * 512 + lane, cell == ncells for RR: the counterparts of
   lane_get / lane_set (src/lanes.h:60-63) and ref_get / ref_set (src/ref.h:28-29) for `run
   --table`. peek reads any cell, a link cell through its chain from the neighbouring machine;
   poke ignores read-only cells (ZERO, ONE, id, link) as the reference does. */
int  gpu_state_peek(gpu_state *s, uint64_t machine, uint32_t cell, uint8_t *v, char *err,
                    size_t errlen);
int  gpu_state_poke(gpu_state *s, uint64_t machine, uint32_t cell, uint8_t v, char *err,
                    size_t errlen);
/* `steps` instructions in every machine, starting at the state's next instruction (pc) and
   wrapping into further passes: whole passes run as gpu_run, a partial pass as the sub-program of
   its instructions (compiled on first use, kept with the code). The counterpart of the reference's
   single steps (ref_cube_step, src/ref.c:119-130) and of finishing a started pass
   (src/jit.c:743-746, src/check.c:89-107). gpu_run refuses a state whose pass is in progress;
   gpu_state_reset sets pc to 0. */
int  gpu_run_steps(gpu_code *c, gpu_state *s, uint64_t steps, char *err, size_t errlen);
uint64_t gpu_state_pc(const gpu_state *s);
/* The device's measured rate of 32-bit v_bitop3_b32 lane-operations per second (a calibration
   launch that fills every SIMD), for stating results against the real integer-VALU roofline. */
int  gpu_valu_rate(gpu_engine *e, double *lane_ops_per_s, char *err, size_t errlen);
/* Bytes per second read + written by a plain device-to-device copy of `bytes` (a multiple of 16),
   best of three launches. No counterpart in the reference: the HBM side of the engine's roofline,
   and a copy of known size whose HBM counter readings scale those of the other kernels. */
int  gpu_hbm_rate(gpu_engine *e, size_t bytes, double *bytes_per_s, char *err, size_t errlen);
/* Milliseconds of device time the last gpu_run / gpu_enum_run took (HIP events on the engine's
   stream). */
double gpu_last_ms(const gpu_engine *e);

/* Exhaustive enumeration on the device: the counterpart of enum_run (src/enum.h:37) with the same
   validation messages, the same tried / hits / hit[16] / ops and the same share rule for
   --part (chunks of 64 passes of 512 assignments, chunk c in share c % parts,
   src/enum.c:232-234). The spec is the text the reference parses (src/spec.h); NULL for equiv. */
typedef struct {
    int               mode;      /* TS_EN_VERIFY, TS_EN_EQUIV, TS_EN_SEARCH */
    const ts_program *a, *b;     /* b: equiv only, inputs and outputs matched by name */
    const char       *spec;
    unsigned          ticks;     /* passes per assignment, >= 1 */
    int               part, part_end, parts;
    const gpu_tune   *tune;      /* NULL = defaults */
} gpu_enum;

int  gpu_enum_run(gpu_engine *e, const gpu_enum *en, ts_enum_result *r, char *err, size_t errlen);

/* gpu_enum_run in two steps, so one compiled enumeration can run many times: gpu_enum_compile
   validates and builds the kernel (the reference's enum_run does jit_compile first,
   src/enum.c:276-288); gpu_enum_exec runs the whole enumeration. The gpu_enum_code keeps copies of
   what it needs from the programs. */
typedef struct gpu_enum_code gpu_enum_code;
int  gpu_enum_compile(gpu_engine *e, const gpu_enum *en, gpu_enum_code **out, char *err,
                      size_t errlen);
int  gpu_enum_exec(gpu_enum_code *c, ts_enum_result *r, char *err, size_t errlen);
void gpu_enum_free(gpu_enum_code *c);
/* How a compiled enumeration was built: its compile time, the tape tier or a kernel, the
   scheduling strategy kept (gpu_tune.sched numbering) and, when it was chosen by timing, each
   candidate's time on the sample (ms, 0 = not tried). */
typedef struct {
    double   compile_seconds;
    unsigned tape;
    unsigned sched;
    double   sched_ms[5];
    uint64_t sample_words;
} gpu_enum_info;
int  gpu_enum_stats(const gpu_enum_code *c, gpu_enum_info *info);
/* gpu_enum_exec that allocates nothing: no counterpart in the reference. For callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc). The hit counters and the hit list live in a fixed file-scope device array of
   2 + 2 x 4,096 words per device (64 KiB), held until the process ends; list_words (at most 4,096,
   0 = 4,096) bounds the hit words one launch lists before its words are re-run by halves, as
   gpu_enum_exec does with its own list of 65,536. Larger values fail with a message. Results
   are exactly gpu_enum_exec's. Every call on one device shares the array, so calls that overlap
   on one device race; the array keeps the last call's contents. */
int  gpu_enum_exec_noalloc(gpu_enum_code *c, unsigned list_words, ts_enum_result *r, char *err,
                           size_t errlen);

/* Batched cycle search: many link-free programs, each on `groups` x 512 machines started from the
   reference's t1_start stream of its seed (group 0 exactly t1_start, src/tester.c:83-91; later
   groups continue the stream), each machine's cycle length found with the reference's Brent
   schedule (src/tester.c:96-164); a group stops once all its machines found theirs, or after
   `max_passes` passes, as the reference's does. The counterpart of tester 1's worker
   pool (src/tester.c:182-279): programs are data for a persistent interpreter kernel fed by a device
   work queue, not compiled one by one. stats[i] is program i's histogram (bucket b: cycle length
   in (2^(b-1), 2^b], src/tester.c:42-48; bucket 63: none within max_passes). plain = 1 runs the same
   kernel one launch per program, the baseline a batch is compared with. A work item is one wave
   of 64 words (2,048 machines); a program of one or two groups gets items of 16 or 32 words, and
   a caller may ask for narrower items per program. */
typedef struct {
    const ts_program *prog;
    uint64_t          seed;
    uint64_t         *lam;  /* optional (NULL = not wanted): groups x 512 entries, machine
                               512 g + k's cycle length, 0 if none within max_passes (the
                               reference's t1_cycles lam_out per group, src/tester.c:96) */
    const ts_line    *mem;  /* optional start state instead of t1_start(seed): groups x ncells
                               lines (group-major, the reference's LaneMachine.mem) and rr, groups
                               lines; cells ZERO and ONE keep their constants */
    const ts_line    *rr;
    unsigned          lam_group, lam_groups;  /* lam covers groups [lam_group, lam_group +
                               lam_groups) only (lam_groups x 512 entries); lam_groups 0 = all */
    unsigned          residency; /* 0 = by the tune's cell thresholds; 1 = the smallest LDS class
                               that holds the program (HBM if none does); 2 = HBM; 3, 4, 5 = LDS
                               class 0, 1, 2 (small, medium, large), whose launch sizes its LDS for
                               its largest item: a caller's own grouping (at most one CU's LDS) */
    unsigned          win;       /* slots per window: 0 = the cost model's choice, 8 or 16 */
    unsigned          words;     /* words (of 32 machines) per work item: 16, 32 or 64; 0 = by
                               groups (16 for one group, 32 for two, else 64). A narrower item
                               needs less LDS, so more waves of an LDS class share a CU (each with
                               idle lanes); the classes' cell thresholds count cells of the
                               default width, a program's need is its bytes */
} gpu_batch_item;

typedef struct {
    uint64_t hist[64];
    uint64_t lam_sum;   /* sum of the cycle lengths found */
    uint64_t longest;
    uint64_t passes;    /* passes run, summed over the groups (the reference's T1Size.passes) */
    uint64_t unfinished; /* 1: stopped by the tune's deadline before its search ended (hist,
                            lam_sum, longest and lam then hold nothing to use; passes the work
                            done); a program never started has passes 0 */
} gpu_batch_stats;

typedef struct {
    unsigned groups;        /* 512-machine groups per program (default 2048 = 2^20 machines) */
    unsigned max_passes;    /* default 64 */
    unsigned plain;
    unsigned lds_cells;     /* programs with ncells + 1 <= this keep their machines in LDS (default:
                               as many bytes as 130 cells of 2,048 machines, more when few work
                               items share a CU) */
    unsigned waves_per_cu;  /* persistent waves per CU per class, LDS permitting (default 16) */
    unsigned lds_large_cells; /* programs too large for lds_cells with ncells + 1 <= this keep
                               their machines in LDS in a class of their own, down to two waves per
                               CU (default and upper bound: what two items per CU allow; a value
                               <= lds_cells turns the class off) */
    double   seconds;       /* > 0: the run stops this long after its launch; the searches in
                               flight end at their next 64th pass (the reference's stop rule,
                               src/tester.c:145-148) as unfinished, the rest never start */
    unsigned sink_branch;   /* bit c set: residency class c (LDS small, LDS, LDS large, HBM)
                               branches around the loads and stores of slots without a NAND or a
                               STO; clear: such slots load and store a sink cell (no branch) */
} gpu_batch_tune;

typedef struct {
    double   ms;            /* device time of the batch */
    uint64_t ops;           /* 1-bit ops: instructions x passes x machines, summed */
    uint64_t launches;
    unsigned lds_cells;
    double   class_ms[4];   /* batched: from the start to the end of each residency class's launch
                               (LDS small, LDS, LDS large, HBM); they run concurrently */
    uint64_t class_ops[4];
} gpu_batch_info;

int  gpu_batch_run(gpu_engine *e, const gpu_batch_item *items, size_t n, const gpu_batch_tune *tune,
                   gpu_batch_stats *stats, gpu_batch_info *info, char *err, size_t errlen);

/* Tester 1 on the device: replaces t1_run (src/tester.c:242-279). Program k of size index si is
   random_program(cells[si], ts_t1_program_len(cfg, cells[si]), t1_program_seed(seed, cells[si], k))
   run from t1_start of that seed on 512 machines, each machine's cycle length searched up to
   max_passes passes (< 2^32) with the reference's Brent schedule; out[si] (nsizes entries) gets
   the per-size summary the reference's workers add up (src/tester.c:182-226). Sizes take turns in
   the reference's order, and the longest cycle's seed is the first program, in that order, that
   reaches it (the reference's single-thread result). programs > 0: that many per size, in
   batches (gpu_batch_run with one group per program) of at most 4,096 per size; programs = 0:
   rounds of programs enough for two waves of every wave slot of the device, until `seconds`
   have passed: the searches in flight then stop at their next 64th pass and count as
   unfinished with the passes they ran, as the reference's workers do (src/tester.c:208-216).
   Each size's machines stay in LDS or in HBM, whichever ran a sample of the first round's
   programs of that size (at most 32 passes each) at more 1-bit ops/s (gpu_t1_choices).
   *seconds gets the wall time.
   threads, idle and excluded are ignored. */
int  gpu_t1_run(gpu_engine *e, const ts_t1_config *cfg, ts_t1_size *out, double *seconds, char *err,
                size_t errlen);

/* The residency the last gpu_t1_run chose per size, with the sample rates it measured (no
   counterpart in the reference: its threads have one memory). Returns the number of sizes. */
typedef struct {
    uint32_t cells;
    unsigned residency;   /* 1 = LDS, 2 = HBM */
    double   rate[2];     /* 1-bit ops/s of the sample in LDS and in HBM (0 = not run: too large
                             for LDS) */
    uint64_t sample_programs;
} gpu_t1_choice;
int  gpu_t1_choices(const gpu_engine *e, gpu_t1_choice *out, int max);

/* Serving: verify / equiv / search requests submitted over time and run with continuous batching.
   A request is copied, compiled on a pool of host threads (hipRTC), admitted to the device when its
   code fits the memory budget and fewer than max_active requests run, and enumerated in chunks of
   chunk_words words; `streams` device slots each run one chunk at a time, taking the next chunk of
   the admitted request with the least work left, so requests that arrive while others run start
   within a chunk and their work overlaps on the device. Each result is exactly what gpu_enum_run
   returns for the request alone. Compiled modules stay resident for repeated circuits. Small
   equiv/search requests skip the compile: their tape (gpu_enum_tape) runs on the precompiled
   interpreter kernel; larger ones, up to about 0.3 s of tape (tape_race_ms), start there too and
   move to their compiled kernel once it is ready. */
typedef struct gpu_server gpu_server;
typedef struct {
    unsigned streams;          /* device slots (default 4) */
    unsigned compile_threads;  /* host compile threads (default 8) */
    uint64_t memory_budget;    /* device bytes the admitted requests may hold: resident code
                                  objects, tapes and the scratch of kernels that spill
                                  (default 1 GiB) */
    uint64_t chunk_words;      /* words (32 assignments each) per launch (default 2^24) */
    unsigned max_active;       /* admitted requests at once (default 64) */
    double   tape_max_ms;      /* equiv/search requests whose estimated device time on the tape
                                  interpreter is at most this run there with no compile; larger
                                  ones, up to tape_race_ms, start there and finish compiled
                                  (default 0 = 50 ms; negative = compile only, no tape) */
    double   tape_race_ms;     /* the estimated tape time up to which a request worth a compile
                                  also starts on the tape meanwhile (default 0 = 300 ms, about a
                                  compile; negative = never) */
} gpu_serve_tune;
typedef struct {
    double   submitted_s, compiled_s, admitted_s, started_s, done_s;  /* steady clock, seconds */
    uint64_t launches;
    uint64_t tier;             /* 0 = compiled kernel, 1 = tape interpreter, 2 = started on the
                                  tape while its kernel compiled, finished compiled */
    /* its footprint when admitted: the kernel's VGPRs and LDS bytes per workgroup, the waves of
       it a CU holds at once (which sizes its chunks' grids), and the device bytes it held (code
       object if loaded for it, tape, scratch of a spilling kernel for its grid share) */
    uint32_t vgprs, lds_bytes, waves_per_cu;
    uint64_t hbm_bytes;
} gpu_request_times;
/* state: 0 queued, 1 compiling, 2 waiting for admission, 3 running, 4 done, 5 failed */
int  gpu_server_open(gpu_engine *e, const gpu_serve_tune *tune, gpu_server **out, char *err,
                     size_t errlen);
int  gpu_server_submit(gpu_server *s, const gpu_enum *request, uint64_t *ticket, char *err,
                       size_t errlen);
int  gpu_server_poll(gpu_server *s, uint64_t ticket, ts_enum_result *r, int *state, char *err,
                     size_t errlen);
int  gpu_server_wait(gpu_server *s, uint64_t ticket, ts_enum_result *r, char *err, size_t errlen);
int  gpu_server_times(gpu_server *s, uint64_t ticket, gpu_request_times *t);
/* Frees a finished (done or failed) request: its copies of the programs and its ticket, which no
   call knows afterwards. A server that runs for long releases each request once its result has
   been read; requests never released are freed by gpu_server_close. Fails for an unknown ticket
   or a request still queued or running. No reference counterpart (the reference has no server). */
int  gpu_server_release(gpu_server *s, uint64_t ticket, char *err, size_t errlen);
/* Requests the server holds: submitted and not yet released. */
uint64_t gpu_server_live(gpu_server *s);
void gpu_server_close(gpu_server *s);

/* Listings without a device (the counterpart of `alpha jit`, reference src/main.c:468-496):
   the HIP source gpu_compile / gpu_enum_run would build, and a hipRTC compile of any such source
   for an architecture ("gfx950"), which needs no GPU. */
int  gpu_source(const ts_program *p, const gpu_tune *tune, char **source, gpu_code_info *info,
                char *err, size_t errlen);
int  gpu_enum_source(const gpu_enum *en, char **source, char *err, size_t errlen);
/* The tape the interpreter tier (gpu_tune.tape) runs for an equiv or search enumeration, built on
   the host: groups of 16 entries of independent AND nodes, 80 words per group, field-major:
   a[16], b[16], d[16] (byte offsets of slots of 256 bytes), ma[16], mb[16] (0 or ~0), entry j
   computing d = (a ^ ma) & (b ^ mb); `count` real nodes per group (the rest read slot 0 and write
   the last slot), input leaves (slot offset | input bit j << 24), the hit word's slot
   (| complemented << 31). Free with gpu_tape_free. max_slots 0 = 640 (160 KB of LDS). */
typedef struct {
    uint32_t *ops, *count, *leaves;
    uint32_t  ngroups, nleaves, root, slots;
    uint64_t  and_nodes;
} gpu_tape;
int  gpu_enum_tape(const gpu_enum *en, unsigned max_slots, gpu_tape *out, char *err, size_t errlen);
void gpu_tape_free(gpu_tape *t);
/* The batch interpreter's records of a program (no device; no reference counterpart: the
   reference compiles each program, this lists what gpu_batch_run runs instead): windows of `win`
   slots (8 or 16), each 2 x win words, the load offsets (bit 31 = NAND) then the store offsets
   (bit 31 = the slot stores nothing), in bytes of cell x row with row = 4 x item_words; the sink
   is cell ncells, named by slots without a NAND or a STO. Writes up to cap words to out and the
   total to *nwords; fails for a window size other than 8 or 16 or item words other than 16, 32
   or 64. */
int  gpu_batch_records(const ts_program *p, unsigned item_words, unsigned win, uint32_t *out,
                       size_t cap, size_t *nwords, char *err, size_t errlen);
int  gpu_rtc_check(const char *source, const char *arch, double *seconds, size_t *code_bytes,
                   char *err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif
