// The fp16 -> OCP fp8 e4m3 conversion of activations, shared by every kernel
// that produces or consumes them (gemm.hip, dwconv.hip), so an fp8 copy
// stored by a producer holds exactly the bytes a GEMM would get converting
// the fp16 value itself.
#pragma once

#include "common.h"

namespace dcvc {

typedef short fp8_s16x2 __attribute__((ext_vector_type(2)));

// 8 fp16 -> 8 e4m3 bytes in the same order (x: elements 0-3, y: 4-7),
// saturated at +-448 (the largest finite e4m3), then rounded to nearest even
// at scale 1.
DCVC_DEV int2 fp8x8_from_f16x8(const f16x8& v) {
    const f16x2 lim = {(f16)448.f, (f16)448.f};
    int w[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        f16x2 lo = {v[4 * h], v[4 * h + 1]}, hi = {v[4 * h + 2], v[4 * h + 3]};
        lo = __builtin_elementwise_max(__builtin_elementwise_min(lo, lim), -lim);
        hi = __builtin_elementwise_max(__builtin_elementwise_min(hi, lim), -lim);
        fp8_s16x2 t = {0, 0};
        t = __builtin_amdgcn_cvt_scalef32_pk_fp8_f16(t, lo, 1.0f, false);
        t = __builtin_amdgcn_cvt_scalef32_pk_fp8_f16(t, hi, 1.0f, true);
        w[h] = __builtin_bit_cast(int, t);
    }
    return make_int2(w[0], w[1]);
}

}  // namespace dcvc
