// Lowering of NAND/STO programs to straight-line gfx950 kernels.
//
// A pass of a program is a Boolean function of the state at the start of the pass. Symbolic
// execution of the instruction list turns it into an And-Inverter Graph (AIG): every NAND is an
// AND node with complemented edges, a STO only renames, constants fold, and structural hashing
// shares repeated logic. The graph is then printed as HIP source where every value is one 32-bit
// word = 32 machines (machine k of a word is bit k), so one VALU op evaluates one gate for 32
// machines. gfx950's v_bitop3_b32 evaluates any 3-input function, so 3-input cuts of the graph are
// mapped to one instruction each.
//
// Cube programs read cells of other machines (links). Machine m lives in word m >> 5, bit m & 31;
// a read across dims 0-4 is an in-word bit permutation, across dims 5-10 a cross-lane move inside
// the 64-lane wave, and across dims >= 11 a read of another wave's memory. Values computed inside a
// kernel can only be moved within the wave, so a read across dims >= 11 of a cell written earlier
// in the pass cuts the pass into segments (one launch each); values cross segment boundaries
// through HBM in per-(cell, segment) slots with two parities, so no launch ever reads a location
// another wave of the same launch writes (the reference's segment rule, src/jit.c:521-560,
// generalised from "dims >= 9" to "dims the wave cannot see").
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "tesseract_host.h"

namespace ts {

typedef uint32_t Lit;  // node << 1 | complemented
inline Lit lit_not(Lit x) { return x ^ 1u; }

enum NodeKind : uint8_t {
    N_CONST = 0,  // node 0: false
    N_AND,        // a, b: literals
    N_LOAD,       // a: version id, b: machine mask (value of that version in machine k ^ mask)
    N_INBIT,      // a: bit j of the enumeration assignment x
    N_IDBIT,      // a: bit d of the machine number within its cube
    N_PERM,       // a: literal (an AND), b: machine mask 1..2047 (dims 0-10)
};

struct Node {
    uint8_t  kind;
    uint32_t a, b;
};

struct Aig {
    std::vector<Node> nodes;
    std::unordered_map<uint64_t, uint32_t> and_table;
    std::unordered_map<uint64_t, uint32_t> leaf_table;

    Aig();
    Lit And(Lit x, Lit y);
    Lit Or(Lit x, Lit y) { return lit_not(And(lit_not(x), lit_not(y))); }
    Lit Xor(Lit x, Lit y) { return Or(And(x, lit_not(y)), And(lit_not(x), y)); }
    Lit Leaf(uint8_t kind, uint32_t a, uint32_t b);
    Lit Load(uint32_t version, uint32_t mask) { return Leaf(N_LOAD, version, mask); }
    Lit InBit(uint32_t j) { return Leaf(N_INBIT, j, 0); }
    Lit IdBit(uint32_t d) { return Leaf(N_IDBIT, d, 0); }
    Lit Perm(Lit x, uint32_t mask);
    uint8_t kind_of(Lit x) const { return nodes[x >> 1].kind; }
    bool derived(Lit x) const { uint8_t k = kind_of(x); return k == N_AND || k == N_PERM; }
};

// A value of a cell at a segment boundary: written in segment `seg` of this pass, or (seg = -1)
// the value at the start of the pass.
struct Version {
    uint32_t cell;
    int      seg;
};

struct SegOut {
    uint32_t cell;  // ncells = RR
    Lit      v;
    int32_t  slot;  // storage slot, -1 if nobody reads this version
    // only the caller can read it: the cell's value after the pass, which no later segment reads
    // and which the next pass writes before reading (needed after the last pass of a run only)
    bool     final_only = false;
};

struct Segment {
    Aig                 g;
    std::vector<SegOut> outs;
    // slices: output index ranges, each computed by its own threads (cones duplicated); slices
    // [0, always) hold the outputs every pass needs, the rest only final_only outputs
    std::vector<std::vector<uint32_t>> slices;
    size_t always = 0;
};

// Where a version lives at run time.
enum { WHERE_CUR = 0, WHERE_PREV = 1, WHERE_INIT = 2 };

struct Lowered {
    uint32_t ncells = 0;     // RR is cell ncells
    unsigned dims = 9;       // cube dimensions
    unsigned wave_dims = 0;  // dims < wave_dims are in-thread or in-wave
    std::vector<Version> versions;
    std::vector<uint8_t> ver_where;
    std::vector<int32_t> ver_slot;
    std::vector<Segment> segs;
    std::vector<int>     last_seg;   // per cell incl. RR: last segment writing it, -1 = never
    std::vector<int32_t> final_slot; // per cell incl. RR: slot of the value after a pass, -1 = init
    std::vector<int32_t> init_slot;  // per cell incl. RR: slot in the init area, -1 = none
    uint32_t nslots = 0, ninit = 0;
    bool     resident = false;       // one segment, every read in the workgroup: the pass loop
                                     // runs in-kernel
    // dims < wg_dims are inside one workgroup: wave_dims, or all the cube's dims when a whole cube
    // runs in one workgroup (its reads across waves then go through LDS)
    unsigned wg_dims = 0;
};

struct LowerOptions {
    unsigned cube_words_log2;   // log2(words per cube) = dims - 5
    unsigned slice_nodes;       // target AIG nodes per slice (0 = chosen per segment)
    unsigned wg_words_log2 = 0; // a cube of more than 64 and at most 2^this words runs in one
                                // workgroup (0 = never)
};

// Lowers one pass of p into segments. Returns false with a message for programs it cannot take.
bool lower_program(const ts_program *p, const LowerOptions &opt, Lowered &out, std::string &err);

// ---- code generation --------------------------------------------------------------------------

struct KernelSource {
    std::string text;
    unsigned block = 0;              // workgroup size the kernels need (0 = any)
    // LUTs one thread evaluates in a pass: the resident kernel's pass, or over the segments the
    // largest slice of each (what bounds a pass when few cubes run)
    size_t chain_luts = 0;
    std::vector<std::string> names;  // kernel names in launch order (one per segment)
    size_t and_nodes = 0, bitop3 = 0, values = 0;
    unsigned group = 1;  // enumerations: words a thread takes per step of its loop (1 or 32)
};

struct CodegenOptions {
    unsigned block = 256;
    bool     lut3 = true;  // map 3-input cuts to v_bitop3_b32
    unsigned waves = 0;    // run kernels: minimum waves per SIMD (0 = the compiler's choice)
    unsigned group = 0;    // enumerations: 1 = a word per thread, 2 = groups of 32 words, 0 = auto
};

// Streamed form: one kernel per segment, state in HBM slots.
// Resident form (out.resident): one kernel, the pass loop inside, state in registers.
void codegen_run(const Lowered &lw, const CodegenOptions &opt, KernelSource &ks);

// Exhaustive enumeration kernel for link-free programs (verify/equiv/search).
struct EnumLowering {
    int      mode;            // TS_EN_*
    unsigned n;               // inputs
    unsigned ticks;
    // search: target bits; verify: spec lowered per lane; equiv: two programs
    uint64_t target = 0;
    int      width = 0;
    std::vector<uint32_t> out_cells_a;  // search/verify: spec output cells (lowest bit first)
};

bool codegen_enum(const ts_program *a, const ts_program *b, const int *in_map, const int *out_map,
                  const EnumLowering &el, const CodegenOptions &opt, KernelSource &ks,
                  std::string &err, const ts_spec *spec);

// One pass (ticks == 1) from power-on with the inputs set, as AIG literals: oa[k] = cell
// el.out_cells_a[k] of a (search/verify) or a's k-th output (equiv), ob[k] = b's matching output.
// Nodes [1, a_end) were created while executing a, [a_end, b_end) while executing b.
void enum_outputs(const ts_program *a, const ts_program *b, const int *in_map, const int *out_map,
                  const EnumLowering &el, Aig &g, std::vector<Lit> &oa, std::vector<Lit> &ob,
                  uint32_t *a_end, uint32_t *b_end);

// ---- tape: an enumeration as data for the precompiled interpreter kernel ----------------------
//
// The hit word of an equiv or search pass (ticks == 1) as a list of AIG AND nodes in groups of up
// to kGroup mutually independent nodes, every value in a slot of a per-lane file in LDS. A slot is
// reused from the group after its last reader's group on, so a group's writes never touch a slot
// any entry of the same group reads (its entries may run in any order). Node order interleaves the two programs of
// an equiv by relative position, so the live values stay those of a window of each program.
// Every group has kGroup entries: a short group is padded with entries that read slot 0 and write
// the last slot (the sink, read by nobody), so the kernel runs each group without branches.
struct Tape {
    static const unsigned kGroup = 16;
    static const unsigned kWords = 5 * kGroup;  // words per group
    static const unsigned kSlotBytes = 256;     // one 32-bit word for each of a wave's 64 lanes
    // per group, field-major: a[16], b[16], d[16] (slot byte offsets), ma[16], mb[16] (0 or ~0:
    // complement masks of a and b); entry j computes d = (a ^ ma) & (b ^ mb)
    std::vector<uint32_t> ops;
    std::vector<uint32_t> count;   // nodes in each group, 1..kGroup
    std::vector<uint32_t> leaves;  // input leaves: slot byte offset | input bit j << 24
    uint32_t root = 0;             // hit word: slot byte offset | complemented << 31
    uint32_t slots = 0;            // slot 0 holds 0 (the constant)
    size_t   and_nodes = 0;
};

// Builds the tape of an enumeration with one tick; false (with a message) if it needs more than
// max_slots slots or has no tape form (ticks > 1; a verify spec with a right shift or a shift by a
// non-constant). A verify's spec is bit-blasted into the graph (spec_bits).
bool build_tape(const ts_program *a, const ts_program *b, const int *in_map, const int *out_map,
                const EnumLowering &el, unsigned max_slots, Tape &t, std::string &err,
                const ts_spec *spec = nullptr);

// The low `width` bits of the spec's value as AIG literals over the input bits (x bit p =
// g.InBit(p)): ripple adders, shift-and-add products, bitwise logic, left shifts by constants, all
// mod 2^width (the low bits of those depend only on the operands' low bits). False for a right
// shift or a shift by a non-constant, whose low bits need the high ones.
bool spec_bits(Aig &g, const ts_spec *sp, int width, std::vector<Lit> &bits, std::string &err);

}  // namespace ts
