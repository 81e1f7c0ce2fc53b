// Calibration kernel for the bench's roofline: the sustained rate of
// v_mfma_f32_32x32x16_f16 on operands held in registers, with no memory
// traffic in the timed loop.  Each wave keeps 16 independent 32x32
// accumulators (4 A and 4 B fragments) and issues 16 MFMAs per iteration.
// Fed random operands it gives the rate the matrix cores sustain on real
// data; fed zeros, the rate at minimum switching activity.
#include "common.h"
#include "dcvc_ops.h"

namespace dcvc {

__global__ __launch_bounds__(256) void mfma_rate_kernel(const f16x8* __restrict__ ops, float* __restrict__ sink,
                                                        int iters) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const f16x8* src = ops + ((size_t)blockIdx.x * 4 + wave) * 8 * 64;
    f16x8 a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        a[i] = src[i * 64 + lane];
        b[i] = src[(4 + i) * 64 + lane];
    }
    f32x16 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) s += acc[i][j][v];
    sink[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace dcvc

extern "C" int dcvc_mfma_calibrate(const void* operands, float* sink, int blocks, int iters, void* stream) {
    if (!operands || !sink || blocks <= 0 || blocks > 65535 || iters <= 0 || ((uintptr_t)operands) % 16)
        return DCVC_ERR_ARG;
    hipLaunchKernelGGL(dcvc::mfma_rate_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       static_cast<const f16x8*>(operands), sink, iters);
    return hipGetLastError() == hipSuccess ? DCVC_OK : DCVC_ERR_HIP;
}
