// Engine state shared by the engine's translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "tesseract_gpu.h"

struct gpu_engine {
    int         device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t  ev0 = nullptr, ev1 = nullptr;
    int         cus = 0;
    size_t      lds_per_block = 160 * 1024;  // bytes of LDS one workgroup may allocate
    std::string arch, desc;
    double      last_ms = 0;
    // the batch interpreter's per-wave HBM regions, kept between gpu_batch_run calls (freeing a
    // large buffer waits for the device) and released by gpu_engine_close
    void       *batch_scratch = nullptr;
    size_t      batch_scratch_bytes = 0;
    uint32_t   *batch_stop = nullptr;  // pinned host word a batch with a deadline polls
    hipStream_t batch_side[2] = {nullptr, nullptr};  // the LDS classes' streams of a batch
    std::vector<gpu_t1_choice> t1_choices;  // the last tester run's residency per size
};

#include <cstdint>
#include <vector>

#include "lower.hpp"
#include "tesseract_gpu.h"

namespace ts_internal {

// An enumeration's share of the assignment space (reference src/enum.c:215-242): passes of 512 in
// chunks of 64; chunk c belongs to share c % parts; words are enumerated in increasing x.
struct ShareInfo {
    uint64_t space, passes, nchunks, tried, passes_run, nact_words;
    unsigned nshare;
};
ShareInfo shares(unsigned n, int part, int part_end, int parts);

// hit words a launch can list before the list overflows
constexpr uint64_t kListCap = 1u << 16;

// n words from p set to 0 by a kernel on `stream` (clear.hip; no memset node under graph capture)
bool clear_device_words(uint64_t *p, uint32_t n, hipStream_t stream);

// Where an enumeration's kernels count and list their hits: d_out[0] = hits, d_out[1] = hit words
// listed, d_list = `cap` hit words of two words each (word index, 32-bit hit mask).
struct HitBuffers {
    uint64_t *d_out, *d_list;
    uint64_t  cap;
};
// the whole enumeration of c into r, hits counted and listed in b (gpu_enum_exec with its own
// buffers); allocates nothing
int enum_exec(gpu_enum_code *c, const HitBuffers &b, ts_enum_result *r, char *err, size_t errlen);
// the device an enumeration runs on
int enum_device(const gpu_enum_code *c);

// gpu_enum_exec_noalloc's file-scope scratch (noalloc.hip): its address on `device`, which must be
// current, taken once per device when an engine opens
int static_scratch_init(int device, char *err, size_t errlen);

// HIP source -> code object for arch (no device needed). sched picks the compiler's instruction
// scheduling strategy: 0 or 1 its default, 3 max-ilp, 4 iterative-minreg; 2 (iterative-ilp) is
// refused: the compiler itself faults on some kernels under it (the reference's PERM8, Life's
// segment kernel, the mul16 verify against a * b + (a & 1)) and takes the process with it.
int rtc_compile(const std::string &src, const std::string &arch, std::vector<char> &code,
                double *seconds, char *err, size_t errlen, unsigned sched = 0);
const char *sched_strategy(unsigned sched);  // the strategy's name, nullptr for the default
// The strategy enumeration kernels are compiled with when none is timed (short enumerations and
// the server): the compiler's default.
constexpr unsigned kEnumSched = 1;

// validation with the reference's messages, spec parsing and code generation of an enumeration
int enum_prepare(const gpu_engine *e, const gpu_enum *en, gpu_tune &tune, ts::KernelSource &ks,
                 char *err, size_t errlen);

// the same validation, then the enumeration's tape for the interpreter kernel (equiv and search
// with one tick); max_slots bounds the per-lane LDS file
int enum_tape(const gpu_engine *e, const gpu_enum *en, gpu_tune &tune, unsigned max_slots,
              ts::Tape &t, char *err, size_t errlen);

// slots of the tape interpreter's file that fit one workgroup's LDS
inline unsigned tape_max_slots(const gpu_engine *e)
{
    return (unsigned)(e->lds_per_block / ts::Tape::kSlotBytes);
}

// A tape on the device (one allocation) and the interpreter kernel's launch (tape_kernel.hip).
struct TapeDev {
    uint32_t *base = nullptr, *ops = nullptr, *leaves = nullptr;
    uint32_t  ngroups = 0, nleaves = 0, root = 0, slots = 0;
    size_t    bytes = 0;
};
// device bytes of a tape; upload into `into` (at least tape_bytes, kept by the caller) or into a
// fresh allocation that tape_free releases
size_t tape_bytes(const ts::Tape &h);
bool   tape_upload(const ts::Tape &h, TapeDev &d, uint32_t *into = nullptr);
void   tape_free(TapeDev &d);
size_t tape_lds_bytes(const TapeDev &t);
// sets the kernel's LDS limit for this tape and reports resident workgroups per CU
bool   tape_prepare(int device, const TapeDev &t, unsigned *blocks_per_cu);
// one launch over words [i0, i1) with the compiled kernel's arguments and results
bool   tape_launch(const TapeDev &t, hipStream_t stream, unsigned grid, uint64_t *out,
                   uint64_t *list, uint64_t cap, uint64_t i0, uint64_t i1, uint64_t space,
                   uint32_t parts, uint32_t part, uint32_t nshare);

}  // namespace ts_internal
