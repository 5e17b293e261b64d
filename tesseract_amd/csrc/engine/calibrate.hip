// Measured 32-bit integer VALU rate of the device: many waves, each lane running 16 independent
// chains of v_bitop3_b32. The engine reports 1-bit ops against this rate as well as against the
// published 256 CU x 4 SIMD x 16 lane x 2.4 GHz figure.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

__global__ void __launch_bounds__(256) valu_probe(uint32_t *out, uint32_t seed, int iters)
{
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    uint32_t x[16];
#pragma unroll
    for (int k = 0; k < 16; k++)
        x[k] = seed * (uint32_t)(k + 1) ^ gid * 0x9E3779B9u;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 16; k++)
            x[k] = __builtin_amdgcn_bitop3_b32(x[k], x[(k + 1) & 15], x[(k + 5) & 15], 0x6Au);
    }
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 16; k++)
        r ^= x[k];
    out[gid] = r;
}

}  // namespace

// lane-ops per second of v_bitop3_b32 over a launch that fills every SIMD several times
extern "C" int ts_valu_calibrate(hipStream_t stream, int cus, double *lane_ops_per_s)
{
    const unsigned blocks = (unsigned)cus * 16u;
    const int iters = 8192;
    uint32_t *out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint32_t)) != hipSuccess)
        return 0;
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    valu_probe<<<blocks, 256, 0, stream>>>(out, 1u, iters);  // warm-up
    hipEventRecord(a, stream);
    valu_probe<<<blocks, 256, 0, stream>>>(out, 2u, iters);
    hipEventRecord(b, stream);
    hipError_t e = hipEventSynchronize(b);
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    hipFree(out);
    if (e != hipSuccess || ms <= 0)
        return 0;
    *lane_ops_per_s = (double)blocks * 256.0 * iters * 16.0 / (ms * 1e-3);
    return 1;
}
