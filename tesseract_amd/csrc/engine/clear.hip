// A fill kernel for device counters: used where a caller may capture the stream into a HIP graph,
// in which a memset node is not wanted.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine_internal.hpp"

namespace {

__global__ void clear_words(uint64_t *p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

}  // namespace

bool ts_internal::clear_device_words(uint64_t *p, uint32_t n, hipStream_t stream)
{
    const uint32_t block = 64, grid = (n + block - 1) / block;
    hipLaunchKernelGGL(clear_words, dim3(grid), dim3(block), 0, stream, p, n);
    return hipGetLastError() == hipSuccess;
}
