// Measured rates of the device.
// VALU: many waves, each lane running 16 independent chains of v_bitop3_b32; the engine reports
// 1-bit ops against this rate as well as against the published 256 CU x 4 SIMD x 16 lane x 2.4 GHz
// figure.
// HBM: a plain 16-byte-per-lane copy between two device buffers; its byte count is known exactly,
// so its FETCH_SIZE / WRITE_SIZE counter readings also scale those of the engine's kernels.
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

__global__ void __launch_bounds__(256) copy_probe(const uint4 *__restrict__ src,
                                                   uint4 *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (size_t)gridDim.x * 256u)
        dst[i] = src[i];
}

}  // namespace

namespace {

// device buffers and a pair of events, released on every exit path; `ok` collects the HIP status
// of every call made through it
struct Probe {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool ok = true;

    bool check(hipError_t e)
    {
        ok = ok && e == hipSuccess;
        return ok;
    }
    bool setup(size_t bytes0, size_t bytes1)
    {
        return check(hipMalloc(&buf[0], bytes0)) && (!bytes1 || check(hipMalloc(&buf[1], bytes1))) &&
               check(hipEventCreate(&ev[0])) && check(hipEventCreate(&ev[1]));
    }
    // milliseconds between the two events once the second has completed (0 on failure)
    float elapsed()
    {
        float ms = 0;
        if (check(hipEventSynchronize(ev[1])) && check(hipEventElapsedTime(&ms, ev[0], ev[1])))
            return ms;
        return 0;
    }
    ~Probe()
    {
        for (hipEvent_t e : ev)
            if (e)
                (void)hipEventDestroy(e);
        for (void *b : buf)
            if (b)
                (void)hipFree(b);
    }
};

}  // namespace

// lane-ops per second of v_bitop3_b32 over a launch that fills every SIMD several times
extern "C" int ts_valu_calibrate(hipStream_t stream, int cus, double *lane_ops_per_s)
{
    const unsigned blocks = (unsigned)cus * 16u;
    const int iters = 8192;
    Probe p;
    if (!p.setup((size_t)blocks * 256 * sizeof(uint32_t), 0))
        return 0;
    uint32_t *out = (uint32_t *)p.buf[0];
    valu_probe<<<blocks, 256, 0, stream>>>(out, 1u, iters);  // warm-up
    p.check(hipGetLastError());
    p.check(hipEventRecord(p.ev[0], stream));
    valu_probe<<<blocks, 256, 0, stream>>>(out, 2u, iters);
    p.check(hipGetLastError());
    p.check(hipEventRecord(p.ev[1], stream));
    const float ms = p.ok ? p.elapsed() : 0;
    if (!p.ok || ms <= 0)
        return 0;
    *lane_ops_per_s = (double)blocks * 256.0 * iters * 16.0 / (ms * 1e-3);
    return 1;
}

// bytes per second read + written by copy_probe over `bytes` (a multiple of 16), best of 3 launches
extern "C" int ts_hbm_calibrate(hipStream_t stream, int cus, size_t bytes, double *bytes_per_s)
{
    if (bytes < 16 || bytes % 16)
        return 0;
    Probe p;
    if (!p.setup(bytes, bytes))
        return 0;
    uint4 *a = (uint4 *)p.buf[0], *b = (uint4 *)p.buf[1];
    const size_t n = bytes / 16;
    const unsigned blocks = (unsigned)cus * 8u;
    p.check(hipMemsetAsync(a, 0x5A, bytes, stream));
    copy_probe<<<blocks, 256, 0, stream>>>(a, b, n);  // warm-up
    p.check(hipGetLastError());
    float best = 0;
    for (int k = 0; k < 3 && p.ok; k++) {
        p.check(hipEventRecord(p.ev[0], stream));
        copy_probe<<<blocks, 256, 0, stream>>>(k & 1 ? b : a, k & 1 ? a : b, n);
        p.check(hipGetLastError());
        p.check(hipEventRecord(p.ev[1], stream));
        const float ms = p.ok ? p.elapsed() : 0;
        if (ms > 0 && (best == 0 || ms < best))
            best = ms;
    }
    if (!p.ok || best <= 0)
        return 0;
    *bytes_per_s = 2.0 * (double)bytes / (best * 1e-3);
    return 1;
}
