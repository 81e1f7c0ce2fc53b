// Calibration kernel for the bench's roofline: the sustained rate of the
// fp16, bf16 and fp8 MFMAs the GEMMs use, on operands held in registers, with
// no memory traffic in the timed loop.  Each wave keeps 16 independent 32x32
// accumulators (4 A and 4 B fragments) and issues 16 MFMAs per iteration.
// Fed random operands it gives the rate the matrix cores sustain on real
// data; fed zeros, the rate at minimum switching activity.
#include "common.h"
#include "dcvc_ops.h"

namespace dcvc {

typedef __bf16 bf16x8c __attribute__((ext_vector_type(8)));
typedef int i32x8c __attribute__((ext_vector_type(8)));

// PREC 0: v_mfma_f32_32x32x16_f16, 1: v_mfma_f32_32x32x16_bf16 (the same
// operand bits read as bf16), 2: v_mfma_scale_f32_32x32x64_f8f6f4 on OCP e4m3
// (scale 2^0).  Each wave reads 8 fragments of 32 bytes per lane.
template <int PREC>
__global__ __launch_bounds__(256) void mfma_rate_kernel(const i32x8c* __restrict__ ops, float* __restrict__ sink,
                                                        int iters) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const i32x8c* src = ops + ((size_t)blockIdx.x * 4 + wave) * 8 * 64;
    i32x8c a[4], b[4];
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
            for (int j = 0; j < 4; ++j) {
                if constexpr (PREC == 0) {
                    const f16x8 x = __builtin_bit_cast(f16x8, __builtin_shufflevector(a[i], a[i], 0, 1, 2, 3));
                    const f16x8 y = __builtin_bit_cast(f16x8, __builtin_shufflevector(b[j], b[j], 0, 1, 2, 3));
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(x, y, acc[i][j], 0, 0, 0);
                } else if constexpr (PREC == 1) {
                    const bf16x8c x = __builtin_bit_cast(bf16x8c, __builtin_shufflevector(a[i], a[i], 0, 1, 2, 3));
                    const bf16x8c y = __builtin_bit_cast(bf16x8c, __builtin_shufflevector(b[j], b[j], 0, 1, 2, 3));
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, acc[i][j], 0, 0, 0);
                } else {
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[i], b[j], acc[i][j], 0, 0, 0, 127,
                                                                                0, 127);
                }
            }
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

extern "C" int dcvc_mfma_calibrate(const void* operands, float* sink, int blocks, int iters, int prec, void* stream) {
    if (!operands || !sink || blocks <= 0 || blocks > 65535 || iters <= 0 || prec < 0 || prec > 2 ||
        ((uintptr_t)operands) % 32)
        return DCVC_ERR_ARG;
    const dim3 grid(blocks), block(256);
    const auto* ops = static_cast<const dcvc::i32x8c*>(operands);
    const hipStream_t s = (hipStream_t)stream;
    if (prec == 0)
        hipLaunchKernelGGL(dcvc::mfma_rate_kernel<0>, grid, block, 0, s, ops, sink, iters);
    else if (prec == 1)
        hipLaunchKernelGGL(dcvc::mfma_rate_kernel<1>, grid, block, 0, s, ops, sink, iters);
    else
        hipLaunchKernelGGL(dcvc::mfma_rate_kernel<2>, grid, block, 0, s, ops, sink, iters);
    return hipGetLastError() == hipSuccess ? DCVC_OK : DCVC_ERR_HIP;
}
