#pragma once
#include <hip/hip_runtime.h>
#include "common.h"

namespace dcvc {

struct DwArgs {
    const f16* x = nullptr;  // [n_img][H][W] pixels, ldx halves apart, C channels
    int ldx = 0;
    const f16* wt = nullptr; // [9][C] taps (ky * 3 + kx), channel contiguous
    f16* y = nullptr;        // nullptr: no fp16 output (y8 only)
    int ldy = 0;
    uint8_t* y8 = nullptr;   // optional e4m3 copy of the output (fp8x8_from_f16x8), ldy8 bytes apart
    int ldy8 = 0;
    int n_img = 1, H = 0, W = 0, C = 0;
};

int dw3x3_validate(const DwArgs& a);
hipError_t dw3x3_launch(const DwArgs& a, hipStream_t stream);
// Replaces the strip-length table with n entries {n_img, H, W, C, rows (4 or
// 8)}; n = 0 clears it.
int dw3x3_tune_set(const long long* entries, int n);
// Strip length a launch of this shape uses now (table entry or heuristic).
int dw3x3_strip_for(const DwArgs& a);

}  // namespace dcvc
