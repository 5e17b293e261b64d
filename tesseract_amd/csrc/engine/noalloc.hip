// gpu_enum_exec_noalloc: gpu_enum_exec for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Its hit counters and
// hit list live in one file-scope device array, `scratch`, of a fixed size; each device the library
// runs on holds its own copy until the process ends.
//
// In host code the array's name is a host-side stand-in, not a device address. Its address on a
// device is taken once, with that device current, when the engine of that device opens
// (hipGetSymbolAddress), and only that address reaches kernels and copies afterwards. Every call
// on one device shares the array and leaves the last call's contents in it: calls that overlap on
// one device race. The counters are cleared by a kernel at the start of each launch.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "engine_internal.hpp"

namespace {

// the list holds at most this many hit words per launch; a launch that finds more re-runs its
// words by halves until the 16 smallest hits are known, as gpu_enum_exec does with its own list
constexpr uint32_t kMaxListWords = 4096;

// words 0 and 1: the hit count and the number of hit words listed; then the list, two words per
// hit word (its word index, its 32-bit hit mask)
__device__ __align__(256) uint64_t scratch[2 + 2 * kMaxListWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint64_t  *g_ptr[kMaxDevices];  // scratch's address on each device, null until taken

}  // namespace

int ts_internal::static_scratch_init(int device, char *err, size_t errlen)
{
    if (device < 0 || device >= kMaxDevices) {
        snprintf(err, errlen, "device %d is beyond the %d devices the static scratch serves",
                 device, kMaxDevices);
        return 0;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device])
        return 1;
    void *p = nullptr;
    const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (r != hipSuccess || !p) {
        snprintf(err, errlen, "hipGetSymbolAddress of the static scratch failed: %s",
                 hipGetErrorString(r));
        return 0;
    }
    g_ptr[device] = (uint64_t *)p;
    return 1;
}

extern "C" int gpu_enum_exec_noalloc(gpu_enum_code *c, unsigned list_words, ts_enum_result *r,
                                     char *err, size_t errlen)
{
    memset(r, 0, sizeof *r);
    if (list_words > kMaxListWords) {
        snprintf(err, errlen, "a hit list of %u words exceeds the static scratch (at most %u)",
                 list_words, kMaxListWords);
        return 0;
    }
    const int device = ts_internal::enum_device(c);
    uint64_t *p = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (device >= 0 && device < kMaxDevices)
            p = g_ptr[device];
    }
    if (!p) {
        snprintf(err, errlen, "the static scratch of device %d was not set up", device);
        return 0;
    }
    const uint64_t cap = list_words ? list_words : kMaxListWords;
    return ts_internal::enum_exec(c, {p, p + 2, cap}, r, err, errlen);
}
