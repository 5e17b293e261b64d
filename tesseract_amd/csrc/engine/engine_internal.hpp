// Engine state shared by the engine's translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

struct gpu_engine {
    int         device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t  ev0 = nullptr, ev1 = nullptr;
    int         cus = 0;
    std::string arch, desc;
    double      last_ms = 0;
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

// HIP source -> code object for arch (no device needed)
int rtc_compile(const std::string &src, const std::string &arch, std::vector<char> &code,
                double *seconds, char *err, size_t errlen);

// validation with the reference's messages, spec parsing and code generation of an enumeration
int enum_prepare(const gpu_engine *e, const gpu_enum *en, gpu_tune &tune, ts::KernelSource &ks,
                 char *err, size_t errlen);

}  // namespace ts_internal
