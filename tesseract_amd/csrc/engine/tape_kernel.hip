// The tape interpreter: one precompiled kernel that runs any equiv/search enumeration given as a
// tape (groups of independent AND nodes over a per-lane file of slots in LDS). It is the no-compile
// tier of the engine: a request starts as soon as its tape is built on the host.
//
// Layout: a workgroup is four waves over the same 64 words: lane l of any of them owns the 32-bit
// word at byte slot * 256 + 4 l of the workgroup's LDS file, so every slot access of a wave is one
// conflict-free 256-byte row, and the waves split each group's 16 entries (4 each) with one
// barrier per group: four times the waves per CU that one wave per file allows. A group's
// entries are described by 5 x kGroup words (the tape format in lower.hpp), each wave loading its share
// one group ahead. Each lane enumerates one word (32 assignments) at a time exactly like the
// compiled ts_enum kernel — same word order, same share rule, same valid mask, same hit list of
// (word, mask) pairs — so the host's collection of hits is shared by both tiers.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine_internal.hpp"

namespace {

constexpr unsigned kGroup = ts::Tape::kGroup;

// bit j < 5 of x = 32 w + k for the 32 lanes k of a word
__device__ inline uint32_t lane_pattern(uint32_t j)
{
    return j == 0u ? 0xAAAAAAAAu
         : j == 1u ? 0xCCCCCCCCu
         : j == 2u ? 0xF0F0F0F0u
         : j == 3u ? 0xFF00FF00u
                   : 0xFFFF0000u;
}

constexpr unsigned kWaves = 4;               // waves per workgroup, over one slot file
constexpr unsigned kShare = kGroup / kWaves;  // entries per wave
static_assert(kShare == 2 || kShare % 4 == 0, "a wave's share loads as 8- or 16-byte vectors");

// A wave's share of a group's descriptors: fields a, b, d, ma, mb, kShare words each (entries
// kShare h .. kShare h + kShare - 1), loaded as 8- or 16-byte vectors. `zero` is 0 at run time;
// adding (lane & zero) makes the address per-lane for the compiler, so the words arrive through
// vector memory (counted by vmcnt) rather than scalar loads, whose counter is shared with LDS
// reads and would make every LDS wait also wait for the next group's descriptors.
struct Share {
    uint32_t w[5][kShare];
};

__device__ inline void load_share(Share &d, const uint32_t *__restrict__ ops, uint32_t g,
                                 uint32_t wave, uint32_t lane_zero)
{
    const uint32_t *p = ops + (size_t)g * ts::Tape::kWords + wave * kShare + lane_zero;
#pragma unroll
    for (unsigned f = 0; f < 5; f++) {
        if constexpr (kShare == 2) {
            const uint2 x = *(const uint2 *)(p + kGroup * f);
            d.w[f][0] = x.x;
            d.w[f][1] = x.y;
        } else {
#pragma unroll
            for (unsigned q = 0; q < kShare / 4; q++) {
                const uint4 x = *(const uint4 *)(p + kGroup * f + 4 * q);
                d.w[f][4 * q] = x.x;
                d.w[f][4 * q + 1] = x.y;
                d.w[f][4 * q + 2] = x.z;
                d.w[f][4 * q + 3] = x.w;
            }
        }
    }
}

// Hits, reported a wave at a time (every lane of the wave takes part). Each lane counts its hits in
// a register (*acc) and the wave adds its total to out[0] with one atomic at the end (count_hits);
// a word with hits takes a list entry (word index, 32-bit hit mask), one atomic on out[1]
// reserving the wave's entries. A wave that finds the list full (out[1] at or past the capacity)
// makes sure out[1] is past it and stops reserving (*full, the same in every lane). A wave
// without hits only ballots.
__device__ inline void report_hits(uint64_t *out, uint64_t *list, uint64_t cap, uint64_t gw,
                                   uint32_t hit, uint32_t *acc, uint32_t *full)
{
    const uint64_t who = __ballot(hit != 0u);
    if (!who)
        return;
    *acc += __builtin_popcount(hit);
    if (*full)
        return;
    const uint64_t n = (uint64_t)__builtin_popcountll(who);
    const int leader = __builtin_ctzll(who);
    uint64_t base = 0;
    if ((int)(threadIdx.x & 63u) == leader) {
        const uint64_t seen = __atomic_load_n(&out[1], __ATOMIC_RELAXED);
        if (seen < cap) {
            base = atomicAdd((unsigned long long *)&out[1], (unsigned long long)n);
        } else {
            base = cap;
            if (seen == cap)
                atomicAdd((unsigned long long *)&out[1], 1ull);
        }
    }
    base = __shfl(base, leader);
    if (base + n > cap)
        *full = 1u;
    if (hit) {
        const uint64_t k =
            base + __builtin_amdgcn_mbcnt_hi((uint32_t)(who >> 32),
                                             __builtin_amdgcn_mbcnt_lo((uint32_t)who, 0u));
        if (k < cap) {
            list[2 * k] = gw;
            list[2 * k + 1] = hit;
        }
    }
}

__device__ inline void count_hits(uint64_t *out, uint32_t acc)
{
    uint64_t sum = acc;
    for (int off = 32; off > 0; off >>= 1)
        sum += __shfl_xor(sum, off);
    if ((threadIdx.x & 63u) == 0u && sum)
        atomicAdd((unsigned long long *)&out[0], (unsigned long long)sum);
}

// a wave's entries of one group: every read, then every write
__device__ inline void run_share(const Share &d, char *my)
{
    uint32_t v[kShare];
#pragma unroll
    for (unsigned j = 0; j < kShare; j++) {
        const uint32_t x = *(const uint32_t *)(my + d.w[0][j]) ^ d.w[3][j];
        const uint32_t y = *(const uint32_t *)(my + d.w[1][j]) ^ d.w[4][j];
        v[j] = x & y;
    }
#pragma unroll
    for (unsigned j = 0; j < kShare; j++)
        *(uint32_t *)(my + d.w[2][j]) = v[j];
}

__global__ void __launch_bounds__(64 * kWaves)
    ts_enum_tape(const uint32_t *__restrict__ ops, uint32_t ngroups, const uint32_t *__restrict__ leaves,
                 uint32_t nleaves, uint32_t root, uint32_t zero, uint64_t *__restrict__ out,
                 uint64_t *__restrict__ list, uint64_t cap, uint64_t i0, uint64_t i1, uint64_t space,
                 uint32_t parts, uint32_t part, uint32_t nshare)
{
    extern __shared__ uint32_t file[];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    char *const my = (char *)file + lane * 4u;
    const uint32_t lane_zero = lane & zero;
    uint32_t hits = 0, full = 0;  // this lane's hits (wave 0 only); the hit list is full
    if (wave == 0)
        *(uint32_t *)my = 0u;  // slot 0: the constant
    // The waves run the same rounds (the bounds depend on the workgroup only; a lane past i1
    // computes a word with no valid assignment), so every barrier is reached by all of them.
    for (uint64_t base = i0 + (uint64_t)blockIdx.x * 64u; base < i1;
         base += (uint64_t)gridDim.x * 64u) {
        const uint64_t i = base + lane;
        uint64_t gw = i;
        if (parts > 1u) {  // active chunk (64 passes of 512 = 1024 words) -> chunk index
            const uint64_t ci = i >> 10, q = ci / nshare, r = part + ci % nshare;
            gw = ((q * parts + r) << 10) | (i & 1023u);
        }
        const uint64_t xb = gw << 5;
        const uint32_t valid = i >= i1 || xb >= space    ? 0u
                               : space - xb >= 32ull     ? 0xffffffffu
                                                         : ((1u << (uint32_t)(space - xb)) - 1u);
        if (wave == 0)
            for (uint32_t k = 0; k < nleaves; k++) {
                const uint32_t lf = leaves[k], j = lf >> 24;
                *(uint32_t *)(my + (lf & 0xffffffu)) =
                    j < 5u ? lane_pattern(j) : 0u - (uint32_t)((gw >> (j - 5u)) & 1ull);
            }
        __syncthreads();
        // two descriptor buffers: the next group's loads are in flight while this one runs
        Share da, db;
        if (ngroups)
            load_share(da, ops, 0, wave, lane_zero);
        for (uint32_t g = 0; g < ngroups; g += 2) {
            if (g + 1 < ngroups)
                load_share(db, ops, g + 1, wave, lane_zero);
            run_share(da, my);
            __syncthreads();
            if (g + 1 < ngroups) {
                if (g + 2 < ngroups)
                    load_share(da, ops, g + 2, wave, lane_zero);
                run_share(db, my);
                __syncthreads();
            }
        }
        if (wave == 0) {
            const uint32_t hit =
                (*(const uint32_t *)(my + (root & 0x7fffffffu)) ^ (0u - (root >> 31))) & valid;
            report_hits(out, list, cap, gw, hit, &hits, &full);
        }
        __syncthreads();  // the root is read before the next round's leaves are written
    }
    if (wave == 0)
        count_hits(out, hits);
}

}  // namespace

namespace ts_internal {

size_t tape_lds_bytes(const TapeDev &t) { return (size_t)t.slots * ts::Tape::kSlotBytes; }

bool tape_prepare(int device, const TapeDev &t, unsigned *blocks_per_cu)
{
    const size_t lds = tape_lds_bytes(t);
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
            hipSuccess ||
        lds > (size_t)max_lds)
        return false;
    // the kernel's limit is the device's (the same value from every thread), so launches of
    // tapes of any size may run concurrently
    if (hipFuncSetAttribute((const void *)ts_enum_tape, hipFuncAttributeMaxDynamicSharedMemorySize,
                            max_lds) != hipSuccess)
        return false;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, ts_enum_tape, 64 * kWaves, lds) !=
            hipSuccess ||
        blocks < 1)
        return false;
    *blocks_per_cu = (unsigned)blocks;
    return true;
}

bool tape_launch(const TapeDev &t, hipStream_t stream, unsigned grid, uint64_t *out,
                 uint64_t *list, uint64_t cap, uint64_t i0, uint64_t i1, uint64_t space,
                 uint32_t parts, uint32_t part, uint32_t nshare)
{
    ts_enum_tape<<<grid, 64 * kWaves, tape_lds_bytes(t), stream>>>(t.ops, t.ngroups,
                                                           t.leaves, t.nleaves, t.root, 0u, out,
                                                           list, cap, i0, i1, space, parts, part,
                                                           nshare);
    return hipGetLastError() == hipSuccess;
}

size_t tape_bytes(const ts::Tape &h) { return (h.ops.size() + h.leaves.size() + 1) * 4; }

bool tape_upload(const ts::Tape &h, TapeDev &d, uint32_t *into)
{
    d = TapeDev();
    const size_t nops = h.ops.size(), nleaves = h.leaves.size();
    const size_t words = nops + nleaves + 1;
    uint32_t *buf = into;
    if (!buf && hipMalloc(&buf, words * sizeof(uint32_t)) != hipSuccess)
        return false;
    std::vector<uint32_t> all;
    all.reserve(words);
    all.insert(all.end(), h.ops.begin(), h.ops.end());
    all.insert(all.end(), h.leaves.begin(), h.leaves.end());
    all.push_back(0);
    if (hipMemcpy(buf, all.data(), words * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
        if (!into)
            (void)hipFree(buf);
        return false;
    }
    d.base = into ? nullptr : buf;  // a caller's buffer stays the caller's
    d.ops = buf;
    d.leaves = buf + nops;
    d.ngroups = (uint32_t)h.count.size();
    d.nleaves = (uint32_t)nleaves;
    d.root = h.root;
    d.slots = h.slots;
    d.bytes = words * sizeof(uint32_t);
    return true;
}

void tape_free(TapeDev &d)
{
    if (d.base)
        (void)hipFree(d.base);
    d = TapeDev();
}

}  // namespace ts_internal
