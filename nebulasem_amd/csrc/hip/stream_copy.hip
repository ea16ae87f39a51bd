// Plain device copy of a known size: the achievable-HBM-bandwidth yardstick
// and the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters that
// the residual and PCG kernels are measured with.  No reference counterpart.
#include <hip/hip_runtime.h>

#include "nebulasem_amd.h"
#include "status.hpp"

namespace {

__global__ void __launch_bounds__(256) stream_copy_kernel(const float4* __restrict__ src,
                                                          float4* __restrict__ dst, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) stream_fill_kernel(float4* __restrict__ dst, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = make_float4((float)(i & 1023), 1.0f, 2.0f, 3.0f);
}

}  // namespace

extern "C" int nsem_stream_copy(uint64_t bytes, int32_t reps, double* seconds_per_copy) {
    if (bytes < 16 || bytes % 16 != 0 || reps < 1 || !seconds_per_copy)
        return nsem_fail(NSEM_EINVAL, "nsem_stream_copy: bytes must be a positive multiple of 16, reps >= 1");
    const uint64_t n = bytes / 16;
    const uint64_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffull) return nsem_fail(NSEM_EINVAL, "nsem_stream_copy: size too large");
    float4 *src = nullptr, *dst = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = NSEM_OK;
    float ms = 0;
    if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess ||
        hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = nsem_fail(NSEM_EDEVICE, "nsem_stream_copy: allocation failed");
        goto done;
    }
    hipLaunchKernelGGL(stream_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, n);
    hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, dst, n);
    hipEventRecord(e0, st);
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, dst, n);
    hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = nsem_fail(NSEM_EDEVICE, "nsem_stream_copy: kernel failed");
        goto done;
    }
    *seconds_per_copy = ms * 1e-3 / reps;
done:
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    if (src) hipFree(src);
    if (dst) hipFree(dst);
    return rc;
}
