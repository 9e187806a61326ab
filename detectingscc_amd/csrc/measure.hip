// measure.hip -- a plain device copy of known size, used to calibrate the
// HBM byte counters (FETCH_SIZE / WRITE_SIZE) that profile the SCC kernels:
// the counters read for this kernel, divided by the bytes it moves, give the
// scale factor applied to the counters of the other kernels.
#include <cstring>

#include "engine.h"
#include "host_util.h"

namespace sccamd {

__global__ __launch_bounds__(256) void copy_u4_kernel(const uint4 *__restrict__ src,
                                                      uint4 *__restrict__ dst, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

} // namespace sccamd

using namespace sccamd;

extern "C" scc_status scc_copy_bench(scc_ctx *ctx, uint64_t bytes, int iters, double *ms_per_copy)
{
    clear_error();
    if (!ctx || bytes < 16 || iters < 1) return fail(SCC_ERR_INVALID, "bad copy benchmark arguments");
    if (hipSetDevice(ctx->eng.device) != hipSuccess) return fail(SCC_ERR_DEVICE, "hipSetDevice failed");
    const uint64_t n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, n * 16) != hipSuccess || hipMalloc((void **)&b, n * 16) != hipSuccess) {
        if (a) (void)hipFree(a);
        (void)hipGetLastError();
        return fail(SCC_ERR_NOMEM, "copy benchmark allocation of %llu bytes failed",
                    (unsigned long long)bytes);
    }
    hipStream_t s = ctx->eng.stream;
    (void)hipMemsetAsync(a, 1, n * 16, s);
    const uint32_t grid = 1024 * 8;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    copy_u4_kernel<<<grid, 256, 0, s>>>(a, b, n);   // warm-up
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < iters; i++) copy_u4_kernel<<<grid, 256, 0, s>>>(a, b, n);
    (void)hipEventRecord(e1, s);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return fail(SCC_ERR_DEVICE, "copy benchmark failed: %s", hipGetErrorString(err));
    if (ms_per_copy) *ms_per_copy = ms / iters;
    return SCC_OK;
}
