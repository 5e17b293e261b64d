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
