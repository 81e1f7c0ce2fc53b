// Shared device helpers for the DCVC MI355X kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef _Float16 f16;
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

#define DCVC_DEV __device__ __forceinline__

// ---------------------------------------------------------------------------
// status codes returned across the C ABI
// ---------------------------------------------------------------------------
enum DcvcStatus {
    DCVC_OK = 0,
    DCVC_ERR_ARG = 1,        // invalid shape / pointer / unsupported configuration
    DCVC_ERR_HIP = 2,        // HIP runtime failure
    DCVC_ERR_STREAM = 3,     // malformed or truncated bitstream
    DCVC_ERR_OVERFLOW = 4,   // entropy coder buffer overflow
    DCVC_ERR_STATE = 5,      // call order violated (e.g. decode before set_param)
    DCVC_ERR_BUDGET = 6,     // admission rejected: byte budget exceeded
};

// ---------------------------------------------------------------------------
// numerics shared by every kernel that feeds the entropy coder.  Encoder and
// decoder call the same functions, so both sides round identically.
// ---------------------------------------------------------------------------

// C roundf semantics (half away from zero): the rounding the reference's
// quantisation kernels apply (elementwise/stream.cu:587, :873).
// Exact: x - trunc(x) is representable, so no x + 0.5 rounding error.
DCVC_DEV float round_half_away(float x) {
    const float t = __builtin_truncf(x);
    return (__builtin_fabsf(x - t) >= 0.5f) ? t + __builtin_copysignf(1.0f, x) : t;
}

// x * sigmoid(4x) with a deterministic exp (layers.py:106-111).
DCVC_DEV float wsilu(float x) {
    // sigmoid(4x) = 1 / (1 + 2^(-4x*log2(e))); v_exp_f32 and v_rcp_f32 (1 ulp)
    // instead of an IEEE divide, whose scale/fmas/fixup sequence costs ~4x
    // the VALU time in the GEMM epilogues.  x * (1/inf) = 0 for large -x.
    const float e = __builtin_amdgcn_exp2f(-5.770780163555854f * x);
    return x * __builtin_amdgcn_rcpf(1.0f + e);
}

// Micro-mask of the 4-step checkerboard (common_model.py:174-195,
// dmci_proxy.cpp:678-699).  Step k applies, to channel group g, the 2x2
// position perm[k][g]:  k=0: 0 1 2 3 | k=1: 3 2 1 0 | k=2: 2 3 0 1 |
// k=3: 1 0 3 2, which is g ^ {0,3,2,1}[k].  Position p = (h & 1) * 2 + (w & 1).
DCVC_DEV int mask4_pos(int step, int group) { return group ^ ((4 - step) & 3); }
// The one channel group active at position p in step k (inverse of the above).
DCVC_DEV int mask4_group(int step, int pos) { return pos ^ ((4 - step) & 3); }

// Scale -> CDF index via the float32 threshold table (127 entries; entry 0
// unused).  index = #{k >= 1 : s >= th[k]} for the clamped scale, i.e. the
// FP32 semantics of stream.cu:77-87 with def_const.h constants.
DCVC_DEV int scale_index(float s, const float* __restrict__ th) {
    // binary search over the monotone table (th[k] = +inf beyond the clamp)
    int lo = 0;
#pragma unroll
    for (int step = 64; step >= 1; step >>= 1) {
        int k = lo + step;
        if (k < 128 && s >= th[k]) lo = k;
    }
    return lo;
}

static inline int ceil_div(int a, int b) { return (a + b - 1) / b; }
