// Depthwise 3x3 stride-1 pad-1 convolution, NHWC fp16, fp32 accumulate.
//
// Replaces d3x3 (cutlass/d3x3.cu:443-446, d3x3_kernel.h:104-200).  The bias
// of the depthwise conv is folded into the following 1x1 conv at load time,
// as the reference does (layers_proxy.cpp:175-178), so there is none here.
//
// Mapping: one thread owns 8 consecutive channels (one 16-byte vector) of one
// image column w and walks TH output rows, keeping the 3x3 input window in
// registers (3 new 16-byte loads per output row).  Consecutive threads take
// consecutive channel vectors, then consecutive columns, so a wavefront
// reads contiguous NHWC rows.  The tap order (ky, kx) is fixed, so the
// result does not depend on the launch shape.
#include "common.h"
#include <mutex>
#include <shared_mutex>
#include <unordered_map>

#include "dwconv.h"
#include "fp8_dev.h"

namespace dcvc {

// OUT8: the output also (or, with y == nullptr, only) as e4m3 bytes
template <int TH, bool OUT8>
DCVC_DEV void dw3x3_body(const DwArgs& a) {
    const int CG = a.C >> 3;
    const long item = (long)blockIdx.x * 256 + threadIdx.x;
    const int strips = (a.H + TH - 1) / TH;
    const long total = (long)CG * a.W * strips * a.n_img;
    if (item >= total) return;
    const int cg = (int)(item % CG);
    long rest = item / CG;
    const int w = (int)(rest % a.W);
    rest /= a.W;
    const int hs = (int)(rest % strips);
    const int img = (int)(rest / strips);
    const int c0 = cg * 8;

    float wt[9][8];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const f16x8 v = *reinterpret_cast<const f16x8*>(a.wt + (size_t)t * a.C + c0);
#pragma unroll
        for (int e = 0; e < 8; ++e) wt[t][e] = (float)v[e];
    }

    const f16* base = a.x + (size_t)img * a.H * a.W * a.ldx + c0;
    const f16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
    auto load_row = [&](int h, f16x8 (&row)[3]) {
        if (h < 0 || h >= a.H) {
            row[0] = row[1] = row[2] = zero8;
            return;
        }
        const f16* p = base + (size_t)h * a.W * a.ldx;
        row[0] = (w > 0) ? *reinterpret_cast<const f16x8*>(p + (size_t)(w - 1) * a.ldx) : zero8;
        row[1] = *reinterpret_cast<const f16x8*>(p + (size_t)w * a.ldx);
        row[2] = (w + 1 < a.W) ? *reinterpret_cast<const f16x8*>(p + (size_t)(w + 1) * a.ldx) : zero8;
    };

    // rows[i] holds input row h0 - 1 + i.  The strip is unrolled, and the
    // loads run two input rows ahead of the output row being computed, so
    // each wave keeps a row of loads in flight while it computes (2-7 %
    // faster than loading each row just before its use, same bits).
    const int h0 = hs * TH;
    f16x8 rows[TH + 2][3];
    load_row(h0 - 1, rows[0]);
    load_row(h0, rows[1]);
    load_row(h0 + 1, rows[2]);
    const size_t pix0 = (size_t)img * a.H * a.W;
#pragma unroll
    for (int r = 0; r < TH; ++r) {
        if (r + 3 < TH + 2) load_row(h0 + r + 2, rows[r + 3]);
        const int h = h0 + r;
        if (h >= a.H) break;
        const f16x8(&r0)[3] = rows[r];
        const f16x8(&r1)[3] = rows[r + 1];
        const f16x8(&r2)[3] = rows[r + 2];
        f16x8 o;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float s = 0.f;
            s += wt[0][e] * (float)r0[0][e];
            s += wt[1][e] * (float)r0[1][e];
            s += wt[2][e] * (float)r0[2][e];
            s += wt[3][e] * (float)r1[0][e];
            s += wt[4][e] * (float)r1[1][e];
            s += wt[5][e] * (float)r1[2][e];
            s += wt[6][e] * (float)r2[0][e];
            s += wt[7][e] * (float)r2[1][e];
            s += wt[8][e] * (float)r2[2][e];
            // the fp32 sum is rounded to fp32 first: without this pin the
            // compiler fuses the last fma with the fp16 conversion
            // (v_fma_mix*_f16, one rounding), which flips rare outputs
            asm volatile("" : "+v"(s));
            o[e] = (f16)s;
        }
        const size_t pix = pix0 + (size_t)h * a.W + w;
        if constexpr (OUT8) {
            if (a.y) *reinterpret_cast<f16x8*>(a.y + pix * a.ldy + c0) = o;
            *reinterpret_cast<int2*>(a.y8 + pix * a.ldy8 + c0) = fp8x8_from_f16x8(o);
        } else {
            *reinterpret_cast<f16x8*>(a.y + pix * a.ldy + c0) = o;
        }
    }
}

template <int TH>
__global__ __launch_bounds__(256) void dw3x3_kernel(DwArgs a) {
    dw3x3_body<TH, false>(a);
}

template <int TH>
__global__ __launch_bounds__(256) void dw3x3_e4m3_kernel(DwArgs a) {
    dw3x3_body<TH, true>(a);
}

int dw3x3_validate(const DwArgs& a) {
    if (!a.x || !a.wt || (!a.y && !a.y8)) return DCVC_ERR_ARG;
    if (a.C <= 0 || a.C % 8 || a.H <= 0 || a.W <= 0 || a.n_img <= 0) return DCVC_ERR_ARG;
    if (a.ldx < a.C || a.ldx % 8) return DCVC_ERR_ARG;
    if (((uintptr_t)a.x) % 16 || ((uintptr_t)a.wt) % 16) return DCVC_ERR_ARG;
    if (a.y && (a.ldy < a.C || a.ldy % 8 || ((uintptr_t)a.y) % 16)) return DCVC_ERR_ARG;
    if (a.y8 && (a.ldy8 < a.C || a.ldy8 % 8 || ((uintptr_t)a.y8) % 8)) return DCVC_ERR_ARG;
    return DCVC_OK;
}

namespace {
// Tuned strip lengths, (n_img, H, W, C) -> 4 or 8 rows, from the autotuner's
// database; shapes with no entry use the heuristic below.
std::shared_mutex g_dw_mu;
std::unordered_map<uint64_t, int> g_dw_tune;

bool dw_key(long long n_img, long long H, long long W, long long C, uint64_t* key) {
    if (n_img <= 0 || n_img >= (1 << 12) || H <= 0 || H >= (1 << 16) || W <= 0 || W >= (1 << 16) || C <= 0 ||
        C >= (1 << 16))
        return false;
    *key = ((uint64_t)n_img << 48) | ((uint64_t)H << 32) | ((uint64_t)W << 16) | (uint64_t)C;
    return true;
}
}  // namespace

int dw3x3_tune_set(const long long* entries, int n) {
    if (n < 0 || (n > 0 && !entries)) return DCVC_ERR_ARG;
    std::unordered_map<uint64_t, int> t;
    for (int i = 0; i < n; ++i) {
        const long long* e = entries + 5LL * i;
        uint64_t key;
        if (!dw_key(e[0], e[1], e[2], e[3], &key) || (e[4] != 4 && e[4] != 8)) return DCVC_ERR_ARG;
        t[key] = (int)e[4];
    }
    std::unique_lock<std::shared_mutex> lk(g_dw_mu);
    g_dw_tune.swap(t);
    return DCVC_OK;
}

// Strips of 8 rows, or of 4 when 8-row strips give fewer than 480 blocks
// (under two per CU): measured on MI355X, 4-row strips take 10-35 % less
// time on the 68x120 and 34x60 grids and are slower on 136x240.  A tuning
// table entry overrides the rule.  Each output's arithmetic is the same for
// either strip length.
int dw3x3_strip_for(const DwArgs& a) {
    uint64_t key;
    if (dw_key(a.n_img, a.H, a.W, a.C, &key)) {
        std::shared_lock<std::shared_mutex> lk(g_dw_mu);
        const auto it = g_dw_tune.find(key);
        if (it != g_dw_tune.end()) return it->second;
    }
    const long strips = (a.H + 7) / 8;
    const long b8 = ((long)(a.C / 8) * a.W * strips * a.n_img + 255) / 256;
    return b8 < 480 ? 4 : 8;
}

hipError_t dw3x3_launch(const DwArgs& a, hipStream_t stream) {
    auto blocks_for = [&](int th) {
        const long strips = (a.H + th - 1) / th;
        const long total = (long)(a.C / 8) * a.W * strips * a.n_img;
        return (long)((total + 255) / 256);
    };
    const int th = dw3x3_strip_for(a);
    const dim3 grid((unsigned)blocks_for(th));
    if (a.y8) {
        if (th == 4)
            hipLaunchKernelGGL(dw3x3_e4m3_kernel<4>, grid, dim3(256), 0, stream, a);
        else
            hipLaunchKernelGGL(dw3x3_e4m3_kernel<8>, grid, dim3(256), 0, stream, a);
    } else if (th == 4) {
        hipLaunchKernelGGL(dw3x3_kernel<4>, grid, dim3(256), 0, stream, a);
    } else {
        hipLaunchKernelGGL(dw3x3_kernel<8>, grid, dim3(256), 0, stream, a);
    }
    return hipGetLastError();
}

}  // namespace dcvc
