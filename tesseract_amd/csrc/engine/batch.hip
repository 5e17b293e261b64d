// Batched execution: thousands of programs of uneven length and cell count in one persistent launch
// per residency class, fed by a device work queue (the reference runs random programs one per
// thread, src/tester.c:182-226, and finds cycle lengths with Brent's algorithm on 512 lanes,
// src/tester.c:96-164).
//
// Programs are data, not compiled code: a hipRTC compile per random program would cost more than
// running it. Each instruction touches one 4-byte word per 32 machines, so the interpreter must
// keep many accesses in flight and spend almost nothing on control. Each program is re-encoded on
// the host into *windows* of slots in which no NAND reads a cell stored earlier in the same
// window, so a window's loads issue together before any of its slots executes. A slot is a NAND
// and the STO that follows it, either of which may be missing (a missing NAND loads a per-wave
// sink cell, a missing STO stores to it). The window runs without a branch: every slot is one load
// and one store, and rr = M ? !(rr & v) : rr is a single v_bitop3_b32 with the slot's
// wave-uniform mask M. A window of W slots is 2W words (scalar
// loads): the byte offsets of the W cells loaded (bit 31 = NAND), then of the W cells stored, so
// the loads issue once the first half is in while the second is still on its way.
// W is 8 or 16 per program, whichever the host's cost model (a fixed wait per window plus a cost
// per slot, padding included) prefers: a window pays one memory latency for all its loads, so long
// programs with few hazards take 16 slots and hazard-dense ones 8.
// A work item is 64 words (2,048 machines) of one program: one wave. A program of one or two
// groups (the reference's tester runs one: 512 machines) gets items of 16 or 32 words, the wave's
// other lanes idle, so that its cells take a quarter or half the room. The hare's cells live in
// LDS (item words x cells x 4 bytes) for small programs, or in a per-wave region of HBM for large
// ones (coalesced: one contiguous row per cell); the tortoise in HBM, or beside the hare in LDS
// when so few items run that the LDS of a CU holds both for all of them. Start states come from the
// reference's t1_start stream, generated in registers. Brent's schedule gives each machine's
// cycle length; a 512-machine group stops once all its machines found theirs or after max_passes
// passes (the reference's t1_cycles stops a group the same way, src/tester.c:131-144), a wave
// once its four groups have; per-program histograms and the groups' pass counts are reduced per
// wave before one atomic each.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "engine_internal.hpp"
#include "tesseract_gpu.h"

namespace {

constexpr unsigned kWinSmall = 8;     // slots per window: 8 or 16
constexpr unsigned kWinLarge = 16;
constexpr unsigned kLanes = 64;       // words per work item (one wave)
constexpr unsigned kStats = 68;       // hist[64], lam_sum, longest, passes (sum over groups),
                                      // items stopped unfinished
constexpr unsigned kSample = 64;      // cells compared before a full comparison (reference: 64)
constexpr int kClasses = 4;           // residency classes: LDS small, LDS, LDS large, HBM
constexpr int kHbm = 3;
struct BatchProg {
    uint32_t rec_off;    // first window record (16-byte units)
    uint32_t nrec;       // windows per pass (2 x win words each)
    uint32_t win;        // slots per window (8 or 16)
    uint32_t ncells;     // cells incl. ZERO and ONE (the sink is cell ncells)
    uint32_t nsample;    // sampled stored cells (<= 64; the list is padded to a multiple of 8)
    uint32_t sample_off; // into the cell lists: the sample
    uint32_t nstored;    // distinct cells a pass stores (live STOs)
    uint32_t cells_off;  // into the cell lists: the stored cells, padded to a multiple of 8 with
    uint32_t nstored8;   //   repeats (nstored8 entries), then the cells read but never stored
    uint32_t ncells_used;//   (ncells_used entries in all)
    uint32_t first_item; // global index of the program's first work item
    uint32_t words;      // words (of 32 machines) per work item: 16, 32 or 64
    uint64_t seed;       // t1_start stream seed
    uint64_t lam_off;    // first of the program's per-machine cycle lengths, or kNoLam
    uint64_t state_off;  // first u64 of the program's start state, or kNoLam (t1_start stream)
    uint32_t lam_w0, lam_w1;  // words [lam_w0, lam_w1) have their cycle lengths written
};

constexpr uint64_t kNoLam = ~0ull;

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Storage of one item's machines: element (cell, lane) at byte cell * row + lane * 4 of the base,
// row = 4 x the item's words.
struct Cells {
    uint8_t *base;
    uint32_t row;
    __device__ __forceinline__ uint32_t ld(uint32_t cell, uint32_t lane) const
    {
        return *(const uint32_t *)(base + cell * row + lane * 4);
    }
    __device__ __forceinline__ void st(uint32_t cell, uint32_t lane, uint32_t v) const
    {
        *(uint32_t *)(base + cell * row + lane * 4) = v;
    }
    __device__ __forceinline__ uint32_t ld_at(uint32_t byte_off, uint32_t lane4) const
    {
        return *(const uint32_t *)(base + byte_off + lane4);
    }
    __device__ __forceinline__ void st_at(uint32_t byte_off, uint32_t lane4, uint32_t v) const
    {
        *(uint32_t *)(base + byte_off + lane4) = v;
    }
};

// One window: its W loads issue together, then the W slots run in order. A load word has bit 31
// set for a NAND; a store word has bit 31 set when its slot stores nothing. Both then name the
// sink, which SKIP = false touches like any cell (no branch at all) and SKIP = true branches
// around on the scalar unit (the words are wave-uniform), saving the memory instructions.
template <unsigned W, bool SKIP>
__device__ __forceinline__ void run_window(const Cells &h, uint32_t lane4, const uint32_t (&w)[2 * W],
                                           uint32_t &rr)
{
    uint32_t v[W];
#pragma unroll
    for (unsigned k = 0; k < W; k++) {
        v[k] = 0u;
        if (!SKIP || (int32_t)w[k] < 0)
            v[k] = h.ld_at(w[k] & 0x7FFFFFFFu, lane4);
    }
#pragma unroll
    for (unsigned k = 0; k < W; k++) {
        const uint32_t m = (uint32_t)((int32_t)w[k] >> 31);  // all ones for a NAND
        rr = __builtin_amdgcn_bitop3_b32(m, rr, v[k], 0x7Cu); // m ? !(rr & v) : rr
        if (!SKIP || (int32_t)w[W + k] >= 0)
            h.st_at(w[W + k] & 0x7FFFFFFFu, lane4, rr);
    }
}

template <unsigned W, bool SKIP>
__device__ __forceinline__ void run_windows(const Cells &h, uint32_t lane4, const uint4 *rec,
                                            uint32_t nwin, uint32_t &rr)
{
    constexpr unsigned Q = 2 * W / 4;  // 16-byte records per window
    for (uint32_t r = 0; r < nwin; r++) {
        uint32_t w[2 * W];
#pragma unroll
        for (unsigned q = 0; q < Q; q++) {
            const uint4 x = rec[Q * r + q];
            w[4 * q] = x.x;
            w[4 * q + 1] = x.y;
            w[4 * q + 2] = x.z;
            w[4 * q + 3] = x.w;
        }
        run_window<W, SKIP>(h, lane4, w, rr);
    }
}

// One pass: each window's record through the scalar cache as the window starts. A lone wave
// (a search of few programs) waits about one memory latency per window (its loads follow the
// previous window's stores) and then runs the window's dependent chain of W slots.
__device__ __forceinline__ void run_pass(const Cells &h, uint32_t lane, const uint4 *rec,
                                         uint32_t nwin, uint32_t win, bool skip, uint32_t &rr)
{
    if (skip) {
        if (win == kWinLarge)
            run_windows<kWinLarge, true>(h, lane * 4, rec, nwin, rr);
        else
            run_windows<kWinSmall, true>(h, lane * 4, rec, nwin, rr);
    } else {
        if (win == kWinLarge)
            run_windows<kWinLarge, false>(h, lane * 4, rec, nwin, rr);
        else
            run_windows<kWinSmall, false>(h, lane * 4, rec, nwin, rr);
    }
}

// sum over the wave's active lanes of popcount(m): one ballot per bit, so lanes that idle (items
// narrower than the wave) take no part
__device__ __forceinline__ uint32_t wave_popcount(uint32_t m)
{
    uint32_t n = 0;
    for (int b = 0; b < 32; b++)
        n += (uint32_t)__builtin_popcountll(__ballot((m >> b) & 1u));
    return n;
}

__device__ __forceinline__ int bucket_of(uint64_t lam)
{
    int b = 0;
    while (b < 62 && (1ull << b) < lam)
        b++;
    return b;
}

// lanes in `found` reached their cycle, of length lam
__device__ __forceinline__ void record(uint64_t *st, uint32_t found, uint64_t lam, uint32_t lane)
{
    uint32_t n = wave_popcount(found);
    if (n && lane == 0) {
        atomicAdd((unsigned long long *)&st[bucket_of(lam)], (unsigned long long)n);
        atomicAdd((unsigned long long *)&st[64], (unsigned long long)(n * lam));
        atomicMax((unsigned long long *)&st[65], (unsigned long long)lam);
    }
}

template <bool LDS, bool TORT_LDS>
__global__ void __launch_bounds__(64) ts_batch(const BatchProg *__restrict__ progs,
                                               const uint32_t *__restrict__ item_prog,
                                               uint32_t item_base, uint32_t nitems,
                                               const uint4 *__restrict__ recs,
                                               const uint32_t *__restrict__ samples,
                                               uint32_t *__restrict__ queue,
                                               uint32_t *__restrict__ scratch, uint32_t region_bytes,
                                               uint64_t *__restrict__ stats, uint32_t max_passes,
                                               uint32_t prog_words, uint32_t skip_sink,
                                               uint64_t *__restrict__ lams,
                                               const uint64_t *__restrict__ states,
                                               const uint32_t *stop)
{
    extern __shared__ uint32_t lds[];
    const uint32_t lane = threadIdx.x;
    // the hare (every instruction touches it) in LDS or HBM; the tortoise (Brent's compare and
    // copy only) in this wave's HBM region or after the hare in LDS
    // a wave's HBM region: hare and tortoise for the HBM class, the tortoise alone for an LDS
    // class that keeps it in HBM, nothing when both states are in LDS; region_bytes holds the
    // class's largest item
    const size_t stride = (size_t)(LDS ? (TORT_LDS ? 0 : 1) : 2) * region_bytes;
    uint8_t *hbm = (uint8_t *)scratch + (size_t)blockIdx.x * stride;
    // the host's stop flag (pinned host memory, read through to the host each time)
    auto stopped = [&]() {
        return stop && __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u;
    };
    for (;;) {
        if (stopped())  // no new item once the run is stopped
            break;
        uint32_t item = 0;
        if (lane == 0)
            item = atomicAdd(queue, 1u);
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0, 64));
        if (item >= nitems)
            break;
        item += item_base;
        const uint32_t pi = __builtin_amdgcn_readfirstlane(item_prog[item]);
        const BatchProg P = progs[pi];
        const uint32_t item_words = P.words, row = item_words * 4;
        if (lane < item_words) {  // an item narrower than the wave: the other lanes idle
            const uint32_t ncells = P.ncells;
            Cells hare{LDS ? (uint8_t *)lds : hbm, row};
            Cells tort{TORT_LDS ? (uint8_t *)lds + region_bytes : LDS ? hbm : hbm + region_bytes, row};
            const uint4 *rec = recs + P.rec_off;
            uint64_t *st = stats + (size_t)pi * kStats;

            // start state: group g = word / 16, u32 word j of its line; the stream continues group
            // after group, cells 2.. then RR, 8 u64 words each (reference src/tester.c:83-91)
            const uint32_t word = (item - P.first_item) * item_words + lane;
            // a program of fewer groups than its items hold fills part of its last item: the other
            // lanes run a throwaway machine and record nothing
            const uint64_t g = word >> 4;
            const uint32_t w64 = (word & 15) >> 1, hi = word & 1;
            const uint64_t s0 = P.seed ^ 0x5DEECE66Dull, per_group = (uint64_t)(ncells - 1) * 8;
            // a caller's start state: group g's lines of cells 0..ncells-1, then its RR line
            const bool given = P.state_off != kNoLam && word < prog_words;
            const uint64_t *gs = states + (given ? P.state_off + g * (ncells + 1) * 8 + w64 : 0);
            // only cells the program reads or stores: the others never change and never
            // matter (equal in hare and tortoise alike)
            const uint32_t *used = samples + P.cells_off;
            for (uint32_t i = 0; i < P.ncells_used; i++) {
                const uint32_t c = used[i];
                uint64_t k = g * per_group + (uint64_t)(c - 2) * 8 + w64;
                uint64_t v = given ? gs[(size_t)c * 8] : mix64(s0 + (k + 1) * 0x9E3779B97F4A7C15ull);
                hare.st(c, lane, hi ? (uint32_t)(v >> 32) : (uint32_t)v);
            }
            hare.st(0, lane, 0u);
            hare.st(1, lane, 0xFFFFFFFFu);
            uint32_t rr;
            {
                uint64_t k = g * per_group + (uint64_t)(ncells - 2) * 8 + w64;
                uint64_t v = given ? gs[(size_t)ncells * 8] : mix64(s0 + (k + 1) * 0x9E3779B97F4A7C15ull);
                rr = hi ? (uint32_t)(v >> 32) : (uint32_t)v;
            }

            // Brent: the tortoise jumps to the hare at powers of two (reference src/tester.c:96-164)
            // Brent's state compares and tortoise copies cover the stored cells and RR: the
            // state is equal exactly when those are
            for (uint32_t i = 0; i < P.nstored8; i++)
                tort.st(used[i], lane, hare.ld(used[i], lane));
            uint32_t trr = rr;
            run_pass(hare, lane, rec, P.nrec, P.win, skip_sink, rr);
            uint64_t lam = 1, power = 1, steps = 1;
            uint32_t active = word < prog_words ? 0xFFFFFFFFu : 0u;
            const uint32_t *smp = samples + P.sample_off;
            // the wave's 4 groups (16 lanes each) that hold machines of the program, and those still
            // searching; a group's passes are the steps at which its last machine found its cycle
            const uint64_t real = __ballot(word < prog_words);
            uint32_t open_groups = 0;
            for (int k = 0; k < 4; k++)
                open_groups |= ((real >> (16 * k)) & 0xFFFFull) ? 1u << k : 0u;
            uint64_t passes = 0;
            bool aborted = false;
            for (;;) {
                // the sample list is padded to a multiple of 8 (repeats of its last cell), so eight
                // loads of each side are in flight at once
                uint32_t d = rr ^ trr;
                for (uint32_t i = 0; i < P.nsample; i += 8) {
                    uint32_t x[8];
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        x[k] = hare.ld(smp[i + k], lane) ^ tort.ld(smp[i + k], lane);
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        d |= x[k];
                }
                if (P.nsample < P.nstored && __any(active & ~d)) {
                    d = rr ^ trr;
                    for (uint32_t i = 0; i < P.nstored8; i += 8) {
                        uint32_t x[8];
#pragma unroll
                        for (int k = 0; k < 8; k++)
                            x[k] = hare.ld(used[i + k], lane) ^ tort.ld(used[i + k], lane);
#pragma unroll
                        for (int k = 0; k < 8; k++)
                            d |= x[k];
                    }
                }
                const uint32_t found = active & ~d;
                if (__any(found)) {
                    record(st, found, lam, lane);
                    if (P.lam_off != kNoLam && word >= P.lam_w0 && word < P.lam_w1)
                        for (uint32_t m = found; m; m &= m - 1)  // machine 32 word + b
                            lams[P.lam_off + (uint64_t)(word - P.lam_w0) * 32 + __builtin_ctz(m)] = lam;
                }
                active &= d;
                const uint64_t searching = __ballot(active != 0);
                for (int k = 0; k < 4; k++)
                    if ((open_groups >> k & 1) && !((searching >> (16 * k)) & 0xFFFFull)) {
                        open_groups &= ~(1u << k);
                        passes += steps;
                    }
                if (!open_groups)
                    break;
                if (steps >= max_passes) {
                    passes += steps * (uint64_t)__builtin_popcount(open_groups);
                    break;
                }
                if ((steps & 63) == 0 && stopped()) {  // the reference's rule, src/tester.c:145
                    passes += steps * (uint64_t)__builtin_popcount(open_groups);
                    aborted = true;
                    break;
                }
                if (power == lam) {
                    for (uint32_t i = 0; i < P.nstored8; i++)
                        tort.st(used[i], lane, hare.ld(used[i], lane));
                    trr = rr;
                    power *= 2;
                    lam = 0;
                }
                run_pass(hare, lane, rec, P.nrec, P.win, skip_sink, rr);
                lam++;
                steps++;
            }
            uint32_t left = wave_popcount(active);
            if (left && lane == 0)
                atomicAdd((unsigned long long *)&st[63], (unsigned long long)left);
            if (lane == 0)
                atomicAdd((unsigned long long *)&st[66], (unsigned long long)passes);
            if (aborted && lane == 0)
                atomicAdd((unsigned long long *)&st[67], 1ull);
        }
        __syncthreads();  // one wave per block: orders this item's LDS use before the next
    }
}

struct Encoded {
    std::vector<uint32_t> words;  // windows of 2 x win words: win load offsets (| 1 << 31 for a
                                  // NAND), then win store offsets (| 1 << 31: nothing stored)
    uint32_t nwin = 0, win = 0;
    std::vector<uint32_t> stored;     // cells a live STO writes, ascending
    std::vector<uint32_t> read_only;  // writable cells a NAND reads that no live STO writes, ascending
};

// STOs whose value no one sees: the next access to the cell later in the same pass is another
// STO. The state after each pass (all the interpreter and Brent's compares ever observe) is the
// same without them. A STO whose cell is not accessed again in the pass stays: its value is the
// cell's end-of-pass value. Random programs lose 42-53 % of their STOs this way.
std::vector<uint8_t> dead_stores(const ts_program *p)
{
    std::vector<uint8_t> dead(p->len, 0);
    std::vector<uint8_t> stored_next(p->ncells, 0);  // the next access later in the pass is a STO
    for (size_t i = p->len; i-- > 0;) {
        const ts_instr in = p->code[i];
        if (in.op == TS_OP_NAND) {
            stored_next[in.addr] = 0;
        } else {
            dead[i] = stored_next[in.addr];
            stored_next[in.addr] = 1;
        }
    }
    return dead;
}

// Windows of up to `win` slots in which no NAND reads a cell stored earlier in the window.
// Offsets are bytes into an item's cells (cell * row). A slot is a NAND (or nothing: a load of the
// sink cell) followed by a STO (or nothing: a store to the sink): a STO right after a NAND shares
// its slot, since the slot stores rr after the NAND, so a random program of NANDs and STOs in
// equal parts takes about three slots per four instructions, and about 0.6 once dead STOs are
// dropped. A STO to a read-only cell (ignored by the machine) takes no slot; padding loads and
// stores the sink.
Encoded encode(const ts_program *p, uint32_t row, unsigned win_slots)
{
    Encoded e;
    e.win = win_slots;
    const uint32_t sink = p->ncells * row;
    const std::vector<uint8_t> dead = dead_stores(p);
    std::vector<uint32_t> lds_, sts, stored;  // the open window's load and store offsets
    bool open_nand = false;                   // the last slot is a NAND that stores nothing yet
    const uint32_t no_store = sink | 0x80000000u;  // a slot that stores nothing
    auto flush = [&]() {
        while (lds_.size() < win_slots) {
            lds_.push_back(sink);
            sts.push_back(no_store);
        }
        e.words.insert(e.words.end(), lds_.begin(), lds_.end());
        e.words.insert(e.words.end(), sts.begin(), sts.end());
        e.nwin++;
        lds_.clear();
        sts.clear();
        stored.clear();
        open_nand = false;
    };
    for (size_t i = 0; i < p->len; i++) {
        const ts_instr in = p->code[i];
        if (in.op == TS_OP_NAND) {
            if (lds_.size() == win_slots ||
                std::find(stored.begin(), stored.end(), in.addr) != stored.end())
                flush();
            lds_.push_back(in.addr * row | 0x80000000u);
            sts.push_back(no_store);
            open_nand = true;
            continue;
        }
        if (in.addr < TS_CELL_FIRST_FREE || (p->special && p->special[in.addr].kind != TS_CK_PLAIN) ||
            dead[i])
            continue;
        if (open_nand) {
            sts.back() = in.addr * row;
            open_nand = false;
        } else {
            if (lds_.size() == win_slots)
                flush();
            lds_.push_back(sink);
            sts.push_back(in.addr * row);
        }
        stored.push_back(in.addr);
    }
    if (!lds_.empty())
        flush();
    std::vector<uint8_t> use(p->ncells, 0);  // 1 read, 2 stored
    for (size_t i = 0; i < p->len; i++) {
        const ts_instr in = p->code[i];
        if (in.op == TS_OP_NAND)
            use[in.addr] |= 1;
        else if (!dead[i] && in.addr >= TS_CELL_FIRST_FREE &&
                 !(p->special && p->special[in.addr].kind != TS_CK_PLAIN))
            use[in.addr] |= 2;
    }
    for (uint32_t c = TS_CELL_FIRST_FREE; c < p->ncells; c++)
        if (use[c] & 2)
            e.stored.push_back(c);
        else if (use[c])
            e.read_only.push_back(c);
    return e;
}

// The window size with the smaller estimated pass time: each window waits `wait` cycles for its
// loads (about a memory latency when nothing else runs on the CU meanwhile, little when other
// waves fill it), then spends kSlotCycles per slot, padding slots included.
Encoded pick_encoding(const ts_program *p, uint32_t row, double wait)
{
    constexpr double kSlotCycles = 8.0;
    Encoded small = encode(p, row, kWinSmall), large = encode(p, row, kWinLarge);
    const double t_small = small.nwin * (wait + kWinSmall * kSlotCycles);
    const double t_large = large.nwin * (wait + kWinLarge * kSlotCycles);
    return t_large < t_small ? large : small;
}

// fn(i) for i in [0, n) on up to 16 host threads
template <class F>
void parallel_for(size_t n, F fn)
{
    const size_t threads = std::min<size_t>({16, std::max(1u, std::thread::hardware_concurrency()),
                                             (n + 63) / 64});
    if (threads <= 1) {
        for (size_t i = 0; i < n; i++)
            fn(i);
        return;
    }
    std::vector<std::thread> pool;
    for (size_t t = 0; t < threads; t++)
        pool.emplace_back([&, t] {
            for (size_t i = t; i < n; i += threads)
                fn(i);
        });
    for (std::thread &th : pool)
        th.join();
}

void say(char *err, size_t errlen, const char *fmt, const char *what)
{
    if (err && errlen)
        snprintf(err, errlen, fmt, what);
}

// A program the interpreter can run: every instruction a NAND or a STO of a cell below ncells,
// and every cell's byte offset (cell x row, the sink cell ncells included) below 2^31, since bit 31
// of a record word marks a NAND. Returns 0 with a message otherwise.
int check_program(const ts_program *p, uint32_t row, char *err, size_t errlen)
{
    char b[128];
    if ((uint64_t)(p->ncells + 1ull) * row >= (1ull << 31)) {
        snprintf(b, sizeof b, "%u cells of %u bytes each are too many for the batch interpreter",
                 p->ncells, row);
        say(err, errlen, "%s", b);
        return 0;
    }
    for (size_t i = 0; i < p->len; i++) {
        const ts_instr in = p->code[i];
        if (in.addr >= p->ncells || in.op > TS_OP_STO) {
            if (in.op > TS_OP_STO)
                snprintf(b, sizeof b, "instruction %zu has opcode %u", i, (unsigned)in.op);
            else
                snprintf(b, sizeof b, "instruction %zu reads cell %u of %u", i, in.addr, p->ncells);
            say(err, errlen, "%s", b);
            return 0;
        }
    }
    return 1;
}

template <class T>
struct DevBuf {
    T *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    bool alloc(size_t n) { return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess; }
};

}  // namespace

#define BCHK(call)                                                                           \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            say(err, errlen, "%s", hipGetErrorString(e_));                                   \
            return 0;                                                                        \
        }                                                                                    \
    } while (0)

extern "C" int gpu_batch_records(const ts_program *p, unsigned item_words, unsigned win,
                                 uint32_t *out, size_t cap, size_t *nwords, char *err,
                                 size_t errlen)
{
    if (!p || !nwords || (win != kWinSmall && win != kWinLarge) ||
        (item_words != 16 && item_words != 32 && item_words != kLanes)) {
        say(err, errlen, "%s", "gpu_batch_records: a program, 8 or 16 slots and 16, 32 or 64 item words");
        return 0;
    }
    if (!check_program(p, item_words * 4, err, errlen))
        return 0;
    Encoded enc = encode(p, item_words * 4, win);
    *nwords = enc.words.size();
    if (out)
        memcpy(out, enc.words.data(), std::min(cap, enc.words.size()) * sizeof(uint32_t));
    return 1;
}

extern "C" int gpu_batch_run(gpu_engine *e, const gpu_batch_item *items, size_t n,
                             const gpu_batch_tune *tune_in, gpu_batch_stats *stats_out,
                             gpu_batch_info *info, char *err, size_t errlen)
{
    gpu_batch_tune tune{};
    if (tune_in)
        tune = *tune_in;
    if (!tune.groups)
        tune.groups = 2048;
    if (!tune.max_passes)
        tune.max_passes = 64;
    if (!tune.lds_cells)
        tune.lds_cells = 130;
    if (!tune.waves_per_cu)
        tune.waves_per_cu = 16;
    if (info)
        memset(info, 0, sizeof *info);
    if (!n) {
        return 1;
    }
    for (size_t i = 0; i < n; i++) {
        const ts_program *p = items[i].prog;
        if (!p) {
            say(err, errlen, "%s", "a batch item without a program");
            return 0;
        }
        if (items[i].residency > 5) {
            say(err, errlen, "%s", "a batch item's residency is 0-5");
            return 0;
        }
        if (!check_program(p, kLanes * 4, err, errlen))
            return 0;
        if (items[i].lam_groups &&
            (uint64_t)items[i].lam_group + items[i].lam_groups > tune.groups) {
            say(err, errlen, "%s", "the cycle lengths asked for are of groups the run does not have");
            return 0;
        }
        if (p->ncells < 3) {
            say(err, errlen, "%s", "a batched program needs at least one writable cell");
            return 0;
        }
        for (uint32_t c = 0; p->special && c < p->ncells; c++)
            if (p->special[c].kind == TS_CK_LINK || p->special[c].kind == TS_CK_ID) {
                say(err, errlen, "%s", "batched programs have no id or link cells");
                return 0;
            }
    }
    BCHK(hipSetDevice(e->device));
    const uint32_t prog_words = tune.groups * 16;
    // items as wide as the program needs, up to the wave (16 words for one group, 32 for two), or
    // as wide as the caller asks per program
    const uint32_t item_words = tune.groups == 1 ? 16 : tune.groups == 2 ? 32 : kLanes;
    const uint32_t row = item_words * 4;  // bytes of one cell of a default-width item
    std::vector<uint32_t> words(n), per_prog(n);  // per program: words per item, items
    uint64_t items_total = 0;
    for (size_t i = 0; i < n; i++) {
        const unsigned w = items[i].words;
        if (w && w != 16 && w != 32 && w != kLanes) {
            say(err, errlen, "%s", "a batch item's words per work item must be 16, 32 or 64");
            return 0;
        }
        words[i] = w ? w : item_words;
        per_prog[i] = (prog_words + words[i] - 1) / words[i];
        items_total += per_prog[i];
    }
    const size_t lds_max = e->lds_per_block;
    if (!tune_in || !tune_in->lds_cells) {
        // as many bytes per item as 130 cells of a full wave; more when so few items run that
        // each CU holds only a few (a search of few large programs keeps its hare in LDS)
        const uint64_t per_cu_needed = std::max<uint64_t>(1, (items_total + e->cus - 1) / e->cus);
        const uint64_t fit = lds_max / per_cu_needed / row;  // cells, the sink included
        tune.lds_cells = (unsigned)std::max<uint64_t>(130ull * kLanes / item_words,
                                                      fit ? fit - 1 : 0);
    }
    tune.lds_cells = (unsigned)std::min<uint64_t>(tune.lds_cells, lds_max / row - 1);
    // the large LDS class: programs of which two items still fit a CU's LDS
    const uint32_t two_fit = (uint32_t)(lds_max / (2 * (size_t)row));
    if (!tune.lds_large_cells)
        tune.lds_large_cells = two_fit;
    tune.lds_large_cells = std::min<uint32_t>(tune.lds_large_cells, two_fit);

    // residency classes: 0 = LDS (<= 32 cells with the sink), 1 = LDS (<= lds_cells), 2 = LDS
    // (<= lds_large_cells, as few as two waves per CU), 3 = HBM; each class is one persistent
    // launch sized by what a CU can hold. The thresholds count cells of a default-width item; a
    // program's need is the bytes of one of its items (cells and sink x 4 x its words)
    const size_t limit[kClasses] = {std::min<uint32_t>(32, tune.lds_cells) * (size_t)row,
                                    tune.lds_cells * (size_t)row, tune.lds_large_cells * (size_t)row,
                                    ~(size_t)0};
    auto need_of = [&](size_t i) { return (size_t)(items[i].prog->ncells + 1) * 4 * words[i]; };
    // a caller's residency: 1 = the smallest LDS class with room (up to what two items per CU
    // allow), 2 = HBM, 3-5 = LDS class 0-2 whatever the thresholds say (the class's launch then
    // sizes its LDS for its largest item)
    auto class_of_item = [&](size_t i) {
        const size_t need = need_of(i);
        const unsigned r = items[i].residency;
        if (r == 2)
            return kHbm;
        if (r >= 3)
            return (int)(r - 3);
        if (r == 1)
            return need <= limit[0] ? 0 : need <= limit[1] ? 1 : need <= lds_max / 2 ? 2 : kHbm;
        return need <= limit[0] ? 0 : need <= limit[1] ? 1 : need <= limit[2] ? 2 : kHbm;
    };
    std::vector<BatchProg> progs(n);
    std::vector<uint32_t> recs, samples;
    std::vector<uint64_t> states;  // callers' start states
    std::vector<uint32_t> item_of[kClasses];
    size_t maxbytes[kClasses] = {0, 0, 0, 0};  // the largest item of each class, in bytes
    uint64_t class_items[kClasses] = {0, 0, 0, 0};
    uint64_t ops = 0, lam_total = 0;
    std::vector<int> cls_of(n);
    for (size_t i = 0; i < n; i++) {
        const int c = cls_of[i] = class_of_item(i);
        if (c != kHbm && need_of(i) > lds_max) {
            say(err, errlen, "%s", "a batch item asked for LDS needs more than a CU has");
            return 0;
        }
        class_items[c] += per_prog[i];
        maxbytes[c] = std::max(maxbytes[c], need_of(i));
    }
    // per class: persistent waves per CU (LDS bounds the LDS classes; the HBM regions of all
    // waves stay within 16 GiB) and its HBM regions (hare and tortoise per wave)
    unsigned per_cu[kClasses];
    bool tort_lds[kClasses] = {false, false, false, false};
    size_t scratch_off[kClasses + 1] = {0, 0, 0, 0, 0};
    constexpr size_t kRegionBudget = (size_t)16 << 30;
    for (int c = 0; c < kClasses; c++) {
        per_cu[c] = tune.waves_per_cu;
        if (c != kHbm && maxbytes[c]) {
            per_cu[c] = (unsigned)std::max<size_t>(
                1, std::min<size_t>(per_cu[c], lds_max / maxbytes[c]));
            // the tortoise joins the hare in LDS when every item of the class still finds a
            // wave slot (Brent compares the whole state each pass once the sample matches)
            const size_t both = std::min<size_t>(tune.waves_per_cu, lds_max / (2 * maxbytes[c]));
            if (both && class_items[c] <= (size_t)e->cus * both) {
                tort_lds[c] = true;
                per_cu[c] = (unsigned)both;
            }
        }
        const size_t region = (size_t)(c == kHbm ? 2 : tort_lds[c] ? 0 : 1) * maxbytes[c];
        if (region)
            per_cu[c] = (unsigned)std::max<size_t>(
                1, std::min<size_t>(per_cu[c], kRegionBudget / ((size_t)e->cus * region)));
        scratch_off[c + 1] = scratch_off[c] + (size_t)e->cus * per_cu[c] * region / 4;
    }
    // a window's wait for its loads: HBM latency in the HBM class; in LDS, hidden when the class
    // runs more than two waves per CU, otherwise the lone wave's LDS latency
    double wait[kClasses];
    for (int c = 0; c < kClasses; c++)
        wait[c] = c == kHbm || per_cu[c] <= 2 || class_items[c] <= 2ull * e->cus ? 150.0 : 20.0;
    std::vector<Encoded> encs(n);  // the records of each program, encoded on host threads
    parallel_for(n, [&](size_t i) {
        const unsigned w = items[i].win;
        const uint32_t r = words[i] * 4;
        encs[i] = w == kWinSmall || w == kWinLarge ? encode(items[i].prog, r, w)
                                                   : pick_encoding(items[i].prog, r, wait[cls_of[i]]);
    });
    for (size_t i = 0; i < n; i++) {
        const ts_program *p = items[i].prog;
        const int cls = cls_of[i];
        Encoded enc = std::move(encs[i]);
        BatchProg &bp = progs[i];
        bp.rec_off = (uint32_t)(recs.size() / 4);
        bp.nrec = enc.nwin;
        bp.win = enc.win;
        recs.insert(recs.end(), enc.words.begin(), enc.words.end());
        bp.ncells = p->ncells;
        // Brent's sample: up to 64 of the stored cells, evenly spaced; the kernel compares 8 at a
        // time, so both this list and the stored cells are padded with repeats of their last cell
        const std::vector<uint32_t> &sc = enc.stored;
        bp.nstored = (uint32_t)sc.size();
        bp.nsample = std::min<uint32_t>(bp.nstored, kSample);
        bp.sample_off = (uint32_t)samples.size();
        for (uint32_t k = 0; k < bp.nsample; k++)
            samples.push_back(sc[(uint64_t)k * bp.nstored / bp.nsample]);
        while ((samples.size() - bp.sample_off) % 8)
            samples.push_back(samples.back());
        bp.cells_off = (uint32_t)samples.size();
        samples.insert(samples.end(), sc.begin(), sc.end());
        while ((samples.size() - bp.cells_off) % 8)
            samples.push_back(samples.back());
        bp.nstored8 = (uint32_t)(samples.size() - bp.cells_off);
        samples.insert(samples.end(), enc.read_only.begin(), enc.read_only.end());
        bp.ncells_used = (uint32_t)(samples.size() - bp.cells_off);
        bp.seed = items[i].seed;
        bp.lam_off = kNoLam;
        bp.state_off = kNoLam;
        if (items[i].mem && items[i].rr) {  // groups x (ncells + 1) lines: the cells, then RR
            bp.state_off = states.size();
            for (uint32_t g = 0; g < tune.groups; g++) {
                const ts_line *m = items[i].mem + (size_t)g * p->ncells;
                for (uint32_t c = 0; c < p->ncells; c++)
                    states.insert(states.end(), m[c].w, m[c].w + 8);
                states.insert(states.end(), items[i].rr[g].w, items[i].rr[g].w + 8);
            }
        }
        if (items[i].lam) {
            const uint32_t g1 = items[i].lam_groups ? items[i].lam_group + items[i].lam_groups
                                                    : tune.groups;
            bp.lam_w0 = items[i].lam_groups ? items[i].lam_group * 16 : 0;
            bp.lam_w1 = g1 * 16;
            bp.lam_off = lam_total;
            lam_total += (uint64_t)(bp.lam_w1 - bp.lam_w0) * 32;
        }
        bp.first_item = (uint32_t)item_of[cls].size();
        bp.words = words[i];
        for (uint32_t k = 0; k < per_prog[i]; k++)
            item_of[cls].push_back((uint32_t)i);
        ops += (uint64_t)p->len * tune.max_passes * tune.groups * 512ull;
    }
    DevBuf<BatchProg> d_progs;
    DevBuf<uint4> d_recs;
    DevBuf<uint32_t> d_samples, d_items[kClasses], d_queue;
    DevBuf<uint64_t> d_stats, d_lams, d_states;
    bool ok = d_progs.alloc(n) && d_recs.alloc(recs.size() / 4) && d_samples.alloc(samples.size()) &&
              d_queue.alloc(kClasses + n) && d_stats.alloc(n * kStats) && d_lams.alloc(lam_total) &&
              d_states.alloc(states.size());
    for (int c = 0; c < kClasses; c++)
        ok = ok && d_items[c].alloc(item_of[c].size());
    if (!ok) {
        say(err, errlen, "%s", "out of device memory for the batch");
        return 0;
    }
    const size_t scratch_bytes = std::max<size_t>(scratch_off[kClasses], 1) * sizeof(uint32_t);
    if (e->batch_scratch_bytes < scratch_bytes) {
        if (e->batch_scratch)
            (void)hipFree(e->batch_scratch);
        e->batch_scratch = nullptr;
        e->batch_scratch_bytes = 0;
        if (hipMalloc(&e->batch_scratch, scratch_bytes) != hipSuccess) {
            e->batch_scratch = nullptr;
            say(err, errlen, "%s", "out of device memory for the machines' HBM regions");
            return 0;
        }
        e->batch_scratch_bytes = scratch_bytes;
    }
    uint32_t *const scratch = (uint32_t *)e->batch_scratch;
    hipStream_t s = e->stream;
    // LDS beyond the default 64 KiB per workgroup for the LDS classes
    BCHK(hipFuncSetAttribute((const void *)ts_batch<true, false>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max));
    BCHK(hipFuncSetAttribute((const void *)ts_batch<true, true>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max));
    BCHK(hipMemcpyAsync(d_progs.p, progs.data(), n * sizeof(BatchProg), hipMemcpyHostToDevice, s));
    BCHK(hipMemcpyAsync(d_recs.p, recs.data(), recs.size() * 4, hipMemcpyHostToDevice, s));
    BCHK(hipMemcpyAsync(d_samples.p, samples.data(), samples.size() * 4, hipMemcpyHostToDevice, s));
    if (!states.empty())
        BCHK(hipMemcpyAsync(d_states.p, states.data(), states.size() * 8, hipMemcpyHostToDevice, s));
    for (int c = 0; c < kClasses; c++)
        if (!item_of[c].empty())
            BCHK(hipMemcpyAsync(d_items[c].p, item_of[c].data(), item_of[c].size() * 4,
                                hipMemcpyHostToDevice, s));
    BCHK(hipMemsetAsync(d_queue.p, 0, (kClasses + n) * 4, s));
    BCHK(hipMemsetAsync(d_stats.p, 0, n * kStats * 8, s));
    if (lam_total)
        BCHK(hipMemsetAsync(d_lams.p, 0, lam_total * 8, s));

    // a deadline: the kernels poll a pinned host word the host sets when time is up
    uint32_t *stop_dev = nullptr;
    if (tune.seconds > 0 && !tune.plain) {
        if (!e->batch_stop)
            BCHK(hipHostMalloc((void **)&e->batch_stop, sizeof(uint32_t),
                               hipHostMallocCoherent | hipHostMallocMapped));
        __atomic_store_n(e->batch_stop, 0u, __ATOMIC_SEQ_CST);
        BCHK(hipHostGetDevicePointer((void **)&stop_dev, e->batch_stop, 0));
    }

    uint64_t launches = 0;
    auto sink_branch = [&](int cls) -> uint32_t { return (tune.sink_branch >> cls) & 1u; };
    auto launch_class = [&](int cls, hipStream_t st, uint32_t *queue, uint32_t first,
                            uint32_t count) -> int {
        const uint32_t *ip = d_items[cls].p;
        unsigned grid = std::min<unsigned>((unsigned)e->cus * per_cu[cls], count);
        uint32_t *sc = scratch + scratch_off[cls];
        if (cls != kHbm && tort_lds[cls]) {
            const size_t lds_bytes = 2 * maxbytes[cls];
            hipLaunchKernelGGL((ts_batch<true, true>), dim3(grid), dim3(kLanes), lds_bytes, st,
                               d_progs.p, ip, first, count, d_recs.p, d_samples.p, queue, sc,
                               (uint32_t)maxbytes[cls], d_stats.p, tune.max_passes, prog_words, sink_branch(cls), d_lams.p, d_states.p,
                               stop_dev);
        } else if (cls != kHbm) {
            const size_t lds_bytes = maxbytes[cls];
            hipLaunchKernelGGL((ts_batch<true, false>), dim3(grid), dim3(kLanes), lds_bytes, st,
                               d_progs.p, ip, first, count, d_recs.p, d_samples.p, queue, sc,
                               (uint32_t)maxbytes[cls], d_stats.p, tune.max_passes, prog_words, sink_branch(cls), d_lams.p, d_states.p,
                               stop_dev);
        } else {
            hipLaunchKernelGGL((ts_batch<false, false>), dim3(grid), dim3(kLanes), 0, st, d_progs.p,
                               ip, first, count, d_recs.p, d_samples.p, queue, sc,
                               (uint32_t)maxbytes[cls], d_stats.p, tune.max_passes, prog_words,
                               sink_branch(cls), d_lams.p, d_states.p,
                               stop_dev);
        }
        launches++;
        return hipGetLastError() == hipSuccess;
    };
    // batched: the classes run concurrently, LDS-bound next to HBM-bound work. The HBM class runs
    // on the engine's stream, the large LDS class on a side stream and the two small LDS classes
    // (short) one after the other on another, so that no two long classes share one of the
    // process's four hardware queues (a fifth stream would share one, and its kernels would wait
    // behind that queue's work); the engine's stream joins the side streams after every launch.
    // Plain: one launch per program on the engine's stream.
    // the side streams are the engine's, created once (a stream's creation costs more than a
    // small batch's whole run)
    bool used[2] = {false, false};
    hipEvent_t start_ev = nullptr, done[kClasses] = {nullptr, nullptr, nullptr, nullptr};
    struct Cleanup {
        hipEvent_t *st, *d;
        ~Cleanup()
        {
            for (int k = 0; k < kClasses; k++)
                if (d[k]) (void)hipEventDestroy(d[k]);
            if (*st) (void)hipEventDestroy(*st);
        }
    } cleanup{&start_ev, done};
    BCHK(hipEventRecord(e->ev0, s));
    const auto t_launch = std::chrono::steady_clock::now();
    if (!tune.plain) {
        BCHK(hipEventCreateWithFlags(&start_ev, hipEventDisableTiming));
        BCHK(hipEventRecord(start_ev, s));
        const int order[kClasses] = {kHbm, 2, 1, 0};
        for (int cls : order) {
            if (item_of[cls].empty())
                continue;
            hipStream_t st = s;
            if (cls != kHbm) {
                const int k = cls == 2 ? 0 : 1;
                if (!e->batch_side[k])
                    BCHK(hipStreamCreateWithFlags(&e->batch_side[k], hipStreamNonBlocking));
                if (!used[k]) {
                    BCHK(hipStreamWaitEvent(e->batch_side[k], start_ev, 0));
                    used[k] = true;
                }
                st = e->batch_side[k];
            }
            BCHK(hipEventCreate(&done[cls]));
            if (!launch_class(cls, st, d_queue.p + cls, 0, (uint32_t)item_of[cls].size())) {
                say(err, errlen, "%s", "batch launch failed");
                return 0;
            }
            BCHK(hipEventRecord(done[cls], st));
        }
        for (int k = 0; k < 2; k++)
            if (used[k]) {
                hipEvent_t last = k == 0 ? done[2] : (done[0] ? done[0] : done[1]);
                BCHK(hipStreamWaitEvent(s, last, 0));
            }
    } else {
        // one launch per program, in program order (the baseline the batch is compared with)
        for (size_t i = 0; i < n; i++) {
            int cls = cls_of[i];
            if (!launch_class(cls, s, d_queue.p + kClasses + i, progs[i].first_item, per_prog[i])) {
                say(err, errlen, "%s", "batch launch failed");
                return 0;
            }
        }
    }
    BCHK(hipEventRecord(e->ev1, s));
    if (stop_dev) {
        const auto t_end = t_launch + std::chrono::duration<double>(tune.seconds);
        for (;;) {
            const hipError_t q = hipEventQuery(e->ev1);
            if (q == hipSuccess)
                break;
            if (q != hipErrorNotReady)
                BCHK(q);
            if (std::chrono::steady_clock::now() >= t_end) {
                __atomic_store_n(e->batch_stop, 1u, __ATOMIC_SEQ_CST);
                break;
            }
            std::this_thread::sleep_for(std::chrono::microseconds(500));
        }
    }
    BCHK(hipEventSynchronize(e->ev1));
    float ms = 0;
    BCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    e->last_ms = ms;
    std::vector<uint64_t> st(n * kStats);
    BCHK(hipMemcpy(st.data(), d_stats.p, st.size() * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++)
        if (items[i].lam)
            BCHK(hipMemcpy(items[i].lam, d_lams.p + progs[i].lam_off,
                           (size_t)(progs[i].lam_w1 - progs[i].lam_w0) * 32 * sizeof(uint64_t),
                           hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
        memcpy(stats_out[i].hist, &st[i * kStats], 64 * 8);
        stats_out[i].lam_sum = st[i * kStats + 64];
        stats_out[i].longest = st[i * kStats + 65];
        stats_out[i].passes = st[i * kStats + 66];
        stats_out[i].unfinished = st[i * kStats + 67] ? 1 : 0;
    }
    // executed work, counted as the reference counts it (src/tester.c:305): passes x len x 512
    ops = 0;
    for (size_t i = 0; i < n; i++)
        ops += st[i * kStats + 66] * items[i].prog->len * 512ull;
    if (info && !tune.plain)
        for (int k = 0; k < kClasses; k++) {
            float cms = 0;
            if (done[k])
                BCHK(hipEventElapsedTime(&cms, e->ev0, done[k]));
            info->class_ms[k] = cms;  // from the batch's start to the class's end (they overlap)
        }
    if (info) {
        for (size_t i = 0; i < n; i++)
            info->class_ops[cls_of[i]] += st[i * kStats + 66] * items[i].prog->len * 512ull;
        info->ms = ms;
        info->ops = ops;
        info->launches = launches;
        info->lds_cells = tune.lds_cells;
    }
    return 1;
}
