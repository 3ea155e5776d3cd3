// Device diagnostics: a register-only MFMA loop whose measured rate is the
// attainable dense bf16 MFMA throughput of the device at its current clock,
// and the reference point for normalising MFMA busy-cycle counters of the
// library's kernels. No counterpart in the reference.
#include <hip/hip_runtime.h>

#include "mlk_device.h"
#include "mlk_internal.h"

namespace mlk {

// Four independent accumulator chains per wave keep the MFMA pipe full; the
// operands never change, so the loop reads no memory.
__global__ __launch_bounds__(256) void k_mfma_probe(float* __restrict__ out, int iterations) {
    bf16x8 a, b;
    for (int i = 0; i < 8; ++i) {
        a[i] = (__bf16)(0.001f * (float)(threadIdx.x + i));
        b[i] = (__bf16)(0.002f * (float)(i + 1));
    }
    f32x16 acc[4];
    for (int c = 0; c < 4; ++c)
        for (int i = 0; i < 16; ++i) acc[c][i] = 0.0f;
    for (int it = 0; it < iterations; ++it) {
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[c], 0, 0, 0);
    }
    float s = 0.0f;
    for (int c = 0; c < 4; ++c)
        for (int i = 0; i < 16; ++i) s += acc[c][i];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace mlk

using namespace mlk;

extern "C" int mlk_mfma_probe(int32_t blocks, int32_t iterations, double* tflops, void* stream_) {
    if (blocks < 1 || blocks > (1 << 20) || iterations < 1 || !tflops) {
        set_error("mfma_probe: blocks in [1, 2^20], iterations >= 1 and an output are required");
        return MLK_EINVAL;
    }
    hipStream_t stream = (hipStream_t)stream_;
    DevBuf out;
    MLK_TRY(out.reserve((size_t)blocks * 256 * sizeof(float)));
    Events events;
    hipEvent_t e0, e1;
    MLK_CHECK_HIP(events.make(&e0));
    MLK_CHECK_HIP(events.make(&e1));
    hipLaunchKernelGGL(k_mfma_probe, dim3((unsigned)blocks), dim3(256), 0, stream, out.as<float>(), iterations);
    MLK_CHECK_HIP(hipGetLastError());
    MLK_CHECK_HIP(hipEventRecord(e0, stream));
    hipLaunchKernelGGL(k_mfma_probe, dim3((unsigned)blocks), dim3(256), 0, stream, out.as<float>(), iterations);
    MLK_CHECK_HIP(hipGetLastError());
    MLK_CHECK_HIP(hipEventRecord(e1, stream));
    MLK_CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    MLK_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    // 4 waves per block, 4 MFMAs of 32 x 32 x 16 per wave and iteration
    const double flops = 2.0 * 32 * 32 * 16 * 4.0 * 4.0 * (double)iterations * (double)blocks;
    *tflops = ms > 0.0f ? flops / (ms * 1e-3) / 1e12 : 0.0;
    return MLK_OK;
}
