// Bandwidth-bound kernels: layout shuffles, per-channel scaling, the masked
// quantisation / scale-index boundary and stable stream compaction.
// Reference kernels replaced: elementwise/stream.cu, shuffle.cu, cat_and_pad.cu.
#include <cmath>
#include <cstring>

#include "common.h"
#include "elementwise.h"

// The quantisation formulas below are specified as separate IEEE float32
// operations (no fused multiply-add), so that host oracles can reproduce them
// bit-for-bit.
#pragma clang fp contract(off)

namespace dcvc {

// ---------------------------------------------------------------------------
// scale -> index thresholds (def_const.h:6-11 constants, FP32 semantics)
// ---------------------------------------------------------------------------
static const float kScaleMin = 0.11f, kScaleMax = 16.f;
static const float kLogScaleMin = -2.2073f, kLogScaleMax = 2.7726f;

static long scale_index_exact(float s) {
    const float step = (kLogScaleMax - kLogScaleMin) / (float)(128 - 1);
    const float recip = 1.f / step;
    double v = (double)s;
    v = v < (double)kScaleMin ? (double)kScaleMin : (v > (double)kScaleMax ? (double)kScaleMax : v);
    return (long)std::floor((std::log(v) - (double)kLogScaleMin) * (double)recip);
}

void host_scale_index_thresholds(float th[128]) {
    th[0] = -INFINITY;
    int32_t lo_bits, hi_bits;
    const float smin = kScaleMin, smax = kScaleMax;
    for (int k = 1; k < 128; ++k) {
        th[k] = INFINITY;
        if (scale_index_exact(smax) < k) continue;
        std::memcpy(&lo_bits, &smin, 4);
        std::memcpy(&hi_bits, &smax, 4);
        while (lo_bits < hi_bits) {
            int32_t mid = lo_bits + (hi_bits - lo_bits) / 2;
            float v;
            std::memcpy(&v, &mid, 4);
            if (scale_index_exact(v) >= k)
                hi_bits = mid;
            else
                lo_bits = mid + 1;
        }
        std::memcpy(&th[k], &lo_bits, 4);
    }
}

// ---------------------------------------------------------------------------
// element formulas shared by encoder and decoder
// ---------------------------------------------------------------------------
DCVC_DEV int quant_elem(float y, float mean, float scale, float thr) {
    float r = round_half_away(y - mean);
    if (!(scale > thr)) r = 0.f;
    r = fminf(fmaxf(r, -128.f), 127.f);
    return (int)r;
}
DCVC_DEV f16 yhat_elem(int yq, float mean) { return (f16)((float)yq + mean); }
DCVC_DEV float qstep(float q) { return fmaxf(q, 0.5f); }

static inline dim3 grid1d(long n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

// ---------------------------------------------------------------------------
// layout kernels
// ---------------------------------------------------------------------------
// pad_and_unshuffle_8 / pixel_unshuffle (cat_and_pad.cu:7-51, shuffle.cu:134-185):
// out channel co = ci*f*f + dy*f + dx; replicate padding by clamping.
__global__ void pad_unshuffle_kernel(const f16* __restrict__ x, int ldx, int n_img, int H, int W, int Cin,
                                     int f, int Ho, int Wo, f16* __restrict__ y, int ldy) {
    const int Co = Cin * f * f;
    const int CG = Co >> 3;
    const long item = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n_img * Ho * Wo * CG;
    if (item >= total) return;
    const int cg = (int)(item % CG);
    const long pix = item / CG;
    const int wo = (int)(pix % Wo);
    const int ho = (int)((pix / Wo) % Ho);
    const int img = (int)(pix / ((long)Wo * Ho));
    const f16* xb = x + (size_t)img * H * W * ldx;
    f16x8 v;
    const int ff = f * f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int co = cg * 8 + e;
        const int ci = co / ff;
        const int r = co - ci * ff;
        const int h = min(ho * f + r / f, H - 1);
        const int w = min(wo * f + r % f, W - 1);
        v[e] = xb[((size_t)h * W + w) * ldx + ci];
    }
    *reinterpret_cast<f16x8*>(y + pix * ldy + cg * 8) = v;
}

// f = 8 with CIN channels packed (ldx == CIN): one thread per (image, output
// pixel, dy) reads the 8 input pixels of its row segment as CIN 16-byte
// loads (8 * CIN halves, contiguous) and writes the CIN output groups
// co = ci * 64 + dy * 8 + [0, 8) as 16-byte stores.  Segments that reach
// the right padding take the per-element path of the kernel above.
template <int CIN>
__global__ __launch_bounds__(256) void unshuffle8_rows_kernel(const f16* __restrict__ x, int n_img, int H, int W,
                                                              int Ho, int Wo, f16* __restrict__ y, int ldy) {
    const long item = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n_img * Ho * Wo * 8;
    if (item >= total) return;
    const int dy = (int)(item & 7);
    const long pix = item >> 3;
    const int wo = (int)(pix % Wo);
    const int ho = (int)((pix / Wo) % Ho);
    const int img = (int)(pix / ((long)Wo * Ho));
    const int h = min(ho * 8 + dy, H - 1);
    const f16* xr = x + ((size_t)img * H + h) * W * CIN;
    f16* yp = y + pix * ldy + dy * 8;
    if (wo * 8 + 7 < W) {
        f16x8 v[CIN];
        const f16x8* src = reinterpret_cast<const f16x8*>(xr + (size_t)wo * 8 * CIN);
#pragma unroll
        for (int j = 0; j < CIN; ++j) v[j] = src[j];
#pragma unroll
        for (int ci = 0; ci < CIN; ++ci) {
            f16x8 o;
#pragma unroll
            for (int p = 0; p < 8; ++p) o[p] = v[(p * CIN + ci) >> 3][(p * CIN + ci) & 7];
            *reinterpret_cast<f16x8*>(yp + ci * 64) = o;
        }
    } else {
#pragma unroll
        for (int ci = 0; ci < CIN; ++ci) {
            f16x8 o;
#pragma unroll
            for (int p = 0; p < 8; ++p) o[p] = xr[(size_t)min(wo * 8 + p, W - 1) * CIN + ci];
            *reinterpret_cast<f16x8*>(yp + ci * 64) = o;
        }
    }
}

hipError_t pad_unshuffle(const f16* x, int ldx, int n_img, int H, int W, int Cin, int f, int padB,
                         int padR, f16* y, int ldy, hipStream_t s) {
    const int Ho = (H + padB) / f, Wo = (W + padR) / f;
    const bool aligned = ((uintptr_t)x % 16 == 0) && ((uintptr_t)y % 16 == 0) && ldy % 8 == 0 && (W * Cin) % 8 == 0;
    if (f == 8 && ldx == Cin && aligned && (Cin == 3 || Cin == 24)) {
        const long total = (long)n_img * Ho * Wo * 8;
        if (Cin == 3)
            hipLaunchKernelGGL(unshuffle8_rows_kernel<3>, grid1d(total), dim3(256), 0, s, x, n_img, H, W, Ho, Wo, y,
                               ldy);
        else
            hipLaunchKernelGGL(unshuffle8_rows_kernel<24>, grid1d(total), dim3(256), 0, s, x, n_img, H, W, Ho, Wo,
                               y, ldy);
        return hipGetLastError();
    }
    const long total = (long)n_img * Ho * Wo * (Cin * f * f / 8);
    hipLaunchKernelGGL(pad_unshuffle_kernel, grid1d(total), dim3(256), 0, s, x, ldx, n_img, H, W, Cin, f,
                       Ho, Wo, y, ldy);
    return hipGetLastError();
}

// pixel_shuffle_8 with clamp to [-0.5, 0.5] (shuffle.cu:224-316)
__global__ void pixel_shuffle8_kernel(const f16* __restrict__ x, int ldx, int n_img, int H8, int W8, int C,
                                      f16* __restrict__ y, int ldy) {
    const int H = H8 * 8, W = W8 * 8;
    const long pix = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= (long)n_img * H * W) return;
    const int w = (int)(pix % W);
    const int h = (int)((pix / W) % H);
    const int img = (int)(pix / ((long)W * H));
    const f16* src = x + (((size_t)img * H8 + (h >> 3)) * W8 + (w >> 3)) * ldx + (h & 7) * 8 + (w & 7);
    for (int c = 0; c < C; ++c) {
        float v = (float)src[c * 64];
        v = fminf(fmaxf(v, -0.5f), 0.5f);
        y[pix * ldy + c] = (f16)v;
    }
}

// C = 3, packed output (ldy == 3): a block takes up to 32 consecutive source
// pixels of one 1/8-scale row. It loads their 192 channels contiguously into
// LDS. It then writes the 8 output rows they expand to, 256 pixels x 3
// channels each, as contiguous 16-byte stores. Output position q of a row
// is pixel q / 3, channel q % 3: source pixel (q / 3) >> 3, channel
// (q % 3) * 64 + dy * 8 + ((q / 3) & 7).
constexpr int kShufPix = 32;
__global__ __launch_bounds__(256) void pixel_shuffle8_c3_kernel(const f16* __restrict__ x, int ldx, int n_img,
                                                                int H8, int W8, f16* __restrict__ y) {
    __shared__ f16x8 tile[kShufPix * 24];
    const int groups = (W8 + kShufPix - 1) / kShufPix;
    const int g = blockIdx.x % groups;
    const int row8 = blockIdx.x / groups;  // img * H8 + h8
    const int sp0 = g * kShufPix;
    const int n_sp = min(kShufPix, W8 - sp0);
    const f16* src = x + ((size_t)row8 * W8 + sp0) * ldx;
    for (int c = threadIdx.x; c < n_sp * 24; c += 256) {
        const int sp = c / 24, j = c - sp * 24;
        tile[c] = *reinterpret_cast<const f16x8*>(src + (size_t)sp * ldx + j * 8);
    }
    __syncthreads();
    const f16* t = reinterpret_cast<const f16*>(tile);
    const int img = row8 / H8, h8 = row8 - img * H8;
    const int W = W8 * 8;
    const int per_row = n_sp * 3;  // 16-byte chunks per output row segment
    for (int c = threadIdx.x; c < 8 * per_row; c += 256) {
        const int dy = c / per_row, k = c - dy * per_row;
        f16x8 o;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int q = k * 8 + i;
            const int px = q / 3, ch = q - px * 3;
            const float v = (float)t[(px >> 3) * 192 + ch * 64 + dy * 8 + (px & 7)];
            o[i] = (f16)fminf(fmaxf(v, -0.5f), 0.5f);
        }
        f16* dst = y + (((size_t)img * H8 * 8 + h8 * 8 + dy) * W + sp0 * 8) * 3;
        reinterpret_cast<f16x8*>(dst)[k] = o;
    }
}

hipError_t pixel_shuffle8_clamp(const f16* x, int ldx, int n_img, int H8, int W8, int C, f16* y, int ldy,
                                hipStream_t s) {
    if (C == 3 && ldy == 3 && ldx == 192 && (uintptr_t)x % 16 == 0 && (uintptr_t)y % 16 == 0) {
        const int blocks = n_img * H8 * ((W8 + kShufPix - 1) / kShufPix);
        hipLaunchKernelGGL(pixel_shuffle8_c3_kernel, dim3(blocks), dim3(256), 0, s, x, ldx, n_img, H8, W8, y);
        return hipGetLastError();
    }
    const long total = (long)n_img * H8 * 8 * W8 * 8;
    hipLaunchKernelGGL(pixel_shuffle8_kernel, grid1d(total), dim3(256), 0, s, x, ldx, n_img, H8, W8, C, y, ldy);
    return hipGetLastError();
}

// multiply_with_broadcast_cuda (stream.cu:484-546): y = x * q[qp][c]
// x and y may be the same buffer (in place): each lane reads its 8 elements before writing them
__global__ void mul_channel_kernel(const f16* x, int ldx, long M, int C, const f16* __restrict__ q, int qstride,
                                   const int* __restrict__ qp, long hw, f16* y, int ldy) {
    const int CG = C >> 3;
    const long item = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= M * CG) return;
    const long m = item / CG;
    const int c0 = (int)(item % CG) * 8;
    const f16* qr = q + (size_t)(qp ? qp[hw < M ? (int)(m / hw) : 0] : 0) * qstride + c0;
    f16x8 a = *reinterpret_cast<const f16x8*>(x + m * ldx + c0);
    f16x8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o[e] = (f16)((float)a[e] * (float)qr[e]);
    *reinterpret_cast<f16x8*>(y + m * ldy + c0) = o;
}

hipError_t mul_channel(const f16* x, int ldx, long M, int C, const f16* q, int qstride, const int* qp, f16* y,
                       int ldy, hipStream_t s, long hw) {
    if (hw <= 0) hw = M;
    hipLaunchKernelGGL(mul_channel_kernel, grid1d(M * (C / 8)), dim3(256), 0, s, x, ldx, M, C, q, qstride, qp, hw, y,
                       ldy);
    return hipGetLastError();
}

// round_z_cuda (stream.cu:862-884): clamp(round(z), -64, 63) as fp16 and int8 (NHWC-linear)
__global__ void round_z_kernel(const f16* __restrict__ z, int ldz, long M, int C, f16* __restrict__ zhat, int ldh,
                               int8_t* __restrict__ zi8) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * C) return;
    const long m = i / C;
    const int c = (int)(i % C);
    float v = round_half_away((float)z[m * ldz + c]);
    v = fminf(fmaxf(v, -64.f), 63.f);
    zhat[m * ldh + c] = (f16)v;
    zi8[i] = (int8_t)v;
}

hipError_t round_z(const f16* z, int ldz, long M, int C, f16* zhat, int ldh, int8_t* zi8, hipStream_t s) {
    hipLaunchKernelGGL(round_z_kernel, grid1d(M * C), dim3(256), 0, s, z, ldz, M, C, zhat, ldh, zi8);
    return hipGetLastError();
}

// int8_to_dtype_cuda (stream.cu:454-482)
__global__ void i8_to_f16_kernel(const int8_t* __restrict__ zi8, long M, int C, f16* __restrict__ zhat, int ldh) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * C) return;
    zhat[(i / C) * ldh + (i % C)] = (f16)(float)zi8[i];
}

hipError_t i8_to_f16(const int8_t* zi8, long M, int C, f16* zhat, int ldh, hipStream_t s) {
    hipLaunchKernelGGL(i8_to_f16_kernel, grid1d(M * C), dim3(256), 0, s, zi8, M, C, zhat, ldh);
    return hipGetLastError();
}

// replicate_pad_cuda / slice_cuda (cat_and_pad.cu:53-110): y[h][w] = x[min(h,H-1)][min(w,W-1)]
__global__ void replicate_pad_kernel(const f16* __restrict__ x, int ldx, int n_img, int H, int W, int C, int Ho,
                                     int Wo, f16* __restrict__ y, int ldy) {
    const int CG = C >> 3;
    const long item = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= (long)n_img * Ho * Wo * CG) return;
    const int cg = (int)(item % CG);
    const long pix = item / CG;
    const int wo = (int)(pix % Wo);
    const int ho = (int)((pix / Wo) % Ho);
    const int img = (int)(pix / ((long)Wo * Ho));
    const int h = min(ho, H - 1), w = min(wo, W - 1);
    *reinterpret_cast<f16x8*>(y + pix * ldy + cg * 8) =
        *reinterpret_cast<const f16x8*>(x + (((size_t)img * H + h) * W + w) * ldx + cg * 8);
}

hipError_t replicate_pad(const f16* x, int ldx, int n_img, int H, int W, int C, int Ho, int Wo, f16* y, int ldy,
                         hipStream_t s) {
    const long total = (long)n_img * Ho * Wo * (C / 8);
    hipLaunchKernelGGL(replicate_pad_kernel, grid1d(total), dim3(256), 0, s, x, ldx, n_img, H, W, C, Ho, Wo, y, ldy);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// HT-S quantisation boundary (dmc_hts_proxy.cpp:529-556, 676-696)
// ---------------------------------------------------------------------------
// Encoder step k over all (pixel, channel):
//   active = channel group (c / (C/4)) is the one step k uses at (h&1, w&1)
//   ys     = y / max(q_dec, 0.5)                          (stream.cu:422-443)
//   y_q    = clamp(round(ys - mean) * [scale > thr], -128, 127)
//   y_hat  = y_q + mean (active), unchanged (inactive), 0 for inactive at k=0;
//   k=3 multiplies every y_hat by max(q_dec, 0.5)          (stream.cu:548-630)
//   sym    = (y_q << 8) + index(scale), cond = scale > thr   (stream.cu:130-161)
// Active test of checkerboard step `step` for element (pixel h, w; channel c
// of C): kSteps = 4 is the 4x mask (channel quarters, get_mask_4x), kSteps = 2
// the 2x mask (channel halves, get_mask_2x: half g active where
// (h + w + g + step) is even).
template <int kSteps>
DCVC_DEV bool step_active(int step, int c, int C, int h, int w) {
    if constexpr (kSteps == 4) {
        return mask4_pos(step, c / (C >> 2)) == (((h & 1) << 1) | (w & 1));
    } else {
        return ((h + w + c / (C >> 1) + step) & 1) == 0;
    }
}

// One step of the full-tensor boundary (HT-S: 4 steps, LD: 2): y / max(q, .5)
// quantised where active; y_hat accumulates; the last step scales by
// max(q, .5).  Symbols and conditions are written at the active elements.
template <int kSteps>
__global__ void full_enc_step_kernel(Act y, const f16* __restrict__ scales, int lds, const f16* __restrict__ means,
                                     int ldm, const f16* __restrict__ qdec, int ldq, f16* __restrict__ yhat, int ldh,
                                     int16_t* __restrict__ sym, uint8_t* __restrict__ cond,
                                     const float* __restrict__ th, float thr, int step) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int C = y.C;
    if (i >= y.pixels() * C) return;
    const long m = i / C;
    const int c = (int)(i % C);
    const int w = (int)(m % y.W);
    const int h = (int)((m / y.W) % y.H);
    const bool active = step_active<kSteps>(step, c, C, h, w);
    const bool last = step == kSteps - 1;
    const float q = qstep((float)qdec[m * ldq + c]);
    float out;
    if (active) {
        const float ys = (float)y.p[m * y.ld + c] * (1.0f / q);
        const float mean = (float)means[m * ldm + c];
        const float sc = (float)scales[m * lds + c];
        const int yq = quant_elem(ys, mean, sc, thr);
        sym[i] = (int16_t)((yq << 8) + scale_index(sc, th));
        cond[i] = sc > thr ? 1 : 0;
        out = (float)yhat_elem(yq, mean);
    } else {
        out = step == 0 ? 0.f : (float)yhat[m * ldh + c];
    }
    if (last) out = out * q;
    if (active || step == 0 || last) yhat[m * ldh + c] = (f16)out;
}

hipError_t hts_enc_step(const Act& y, const f16* scales, int lds, const f16* means, int ldm, const f16* qdec,
                        int ldq, f16* yhat, int ldh, int16_t* sym, uint8_t* cond, const float* th, float thr,
                        int step, hipStream_t s) {
    hipLaunchKernelGGL(full_enc_step_kernel<4>, grid1d(y.pixels() * y.C), dim3(256), 0, s, y, scales, lds, means,
                       ldm, qdec, ldq, yhat, ldh, sym, cond, th, thr, step);
    return hipGetLastError();
}

hipError_t ld_enc_step(const Act& y, const f16* scales, int lds, const f16* means, int ldm, const f16* qdec, int ldq,
                       f16* yhat, int ldh, int16_t* sym, uint8_t* cond, const float* th, float thr, int step,
                       hipStream_t s) {
    hipLaunchKernelGGL(full_enc_step_kernel<2>, grid1d(y.pixels() * y.C), dim3(256), 0, s, y, scales, lds, means,
                       ldm, qdec, ldq, yhat, ldh, sym, cond, th, thr, step);
    return hipGetLastError();
}

// build_index_dec_cuda (stream.cu:89-115): dense NHWC-linear index + condition
__global__ void hts_dec_index_kernel(Act sc, const float* __restrict__ th, float thr, uint8_t* __restrict__ idx,
                                     uint8_t* __restrict__ cond) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sc.pixels() * sc.C) return;
    const float s = (float)sc.p[(i / sc.C) * sc.ld + (i % sc.C)];
    idx[i] = (uint8_t)scale_index(s, th);
    cond[i] = s > thr ? 1 : 0;
}

hipError_t hts_dec_index(const Act& scales, const float* th, float thr, uint8_t* idx, uint8_t* cond, hipStream_t s) {
    hipLaunchKernelGGL(hts_dec_index_kernel, grid1d(scales.pixels() * scales.C), dim3(256), 0, s, scales, th, thr,
                       idx, cond);
    return hipGetLastError();
}

// restore_y* (stream.cu:685-754) in the encoder's rounding order
template <int kSteps>
__global__ void full_dec_step_kernel(Act g, const int8_t* __restrict__ yq, const f16* __restrict__ means, int ldm,
                                     const f16* __restrict__ qdec, int ldq, f16* __restrict__ yhat, int ldh,
                                     int step) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int C = g.C;
    if (i >= g.pixels() * C) return;
    const long m = i / C;
    const int c = (int)(i % C);
    const int w = (int)(m % g.W);
    const int h = (int)((m / g.W) % g.H);
    const bool active = step_active<kSteps>(step, c, C, h, w);
    const bool last = step == kSteps - 1;
    float out;
    if (active) {
        out = (float)yhat_elem((int)yq[i], (float)means[m * ldm + c]);
    } else {
        out = step == 0 ? 0.f : (float)yhat[m * ldh + c];
    }
    if (last) out = out * qstep((float)qdec[m * ldq + c]);
    if (active || step == 0 || last) yhat[m * ldh + c] = (f16)out;
}

hipError_t hts_dec_step(const Act& g, const int8_t* yq, const f16* means, int ldm, const f16* qdec, int ldq,
                        f16* yhat, int ldh, int step, hipStream_t s) {
    hipLaunchKernelGGL(full_dec_step_kernel<4>, grid1d(g.pixels() * g.C), dim3(256), 0, s, g, yq, means, ldm, qdec,
                       ldq, yhat, ldh, step);
    return hipGetLastError();
}

hipError_t ld_dec_step(const Act& g, const int8_t* yq, const f16* means, int ldm, const f16* qdec, int ldq, f16* yhat,
                       int ldh, int step, hipStream_t s) {
    hipLaunchKernelGGL(full_dec_step_kernel<2>, grid1d(g.pixels() * g.C), dim3(256), 0, s, g, yq, means, ldm, qdec,
                       ldq, yhat, ldh, step);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// DMCI and HT-L quantisation boundary over quarter tensors
// (dmci_proxy.cpp:333-393, 497-597; dmc_htl_proxy.cpp:614-688, 744-903)
// Quarter element (pixel m, channel c' < C/4) maps to channel g*C/4 + c' of the
// one group g active at the pixel's 2x2 position in step k
// (single_part_for_writing/reading_4x, stream.cu:896-962).
// ---------------------------------------------------------------------------
// kElemQ = false: DMCI, y enters as is and y_hat is scaled by the per-QP
// channel vector q[qp] at the end.  kElemQ = true: HT-L, the quantisation
// step is elementwise, qs = max(q_dec, 0.5) from the common params (q at
// ldq): y / qs enters the step and y_hat * qs leaves it
// (divide_with_clamp_min / add_and_multiply_with_clamp_min, stream.cu:40-66,
// 422-443).
template <bool kElemQ>
__global__ void quarter_enc_step_kernel(Act y, const f16* __restrict__ scales, int lds, const f16* __restrict__ means,
                                        int ldm, f16* __restrict__ ysf, int ldsf, const f16* __restrict__ q, int ldq,
                                        const int* __restrict__ qp, f16* __restrict__ yout, int ldo,
                                        int16_t* __restrict__ sym, uint8_t* __restrict__ cond,
                                        const float* __restrict__ th, float thr, int step) {
    const int C4 = y.C >> 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= y.pixels() * C4) return;
    const long m = i / C4;
    const int cq = (int)(i % C4);
    const int w = (int)(m % y.W);
    const int h = (int)((m / y.W) % y.H);
    const int pos = ((h & 1) << 1) | (w & 1);
    const int ga = mask4_group(step, pos);
    const int ca = ga * C4 + cq;
    const float mean = (float)means[m * ldm + ca];
    const float sc = (float)scales[m * lds + ca];
    float yv = (float)y.p[m * y.ld + ca];
    if constexpr (kElemQ) yv = yv * (1.0f / qstep((float)q[m * ldq + ca]));
    const int yq = quant_elem(yv, mean, sc, thr);
    sym[i] = (int16_t)((yq << 8) + scale_index(sc, th));
    cond[i] = sc > thr ? 1 : 0;
    const f16 yh = yhat_elem(yq, mean);
    if (step < 3) {
        if (step == 0) {
#pragma unroll
            for (int g = 0; g < 4; ++g) ysf[m * ldsf + g * C4 + cq] = (g == ga) ? yh : (f16)0.f;
        } else {
            ysf[m * ldsf + ca] = yh;
        }
    } else {
        const f16* qr = kElemQ ? q + m * ldq : q + (size_t)(qp ? *qp : 0) * ldq;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int c = g * C4 + cq;
            const float v = (g == ga) ? (float)yh : (float)ysf[m * ldsf + c];
            const float qc = kElemQ ? qstep((float)qr[c]) : (float)qr[c];
            yout[m * ldo + c] = (f16)(v * qc);
        }
    }
}

hipError_t intra_enc_step(const Act& y, const f16* scales, int lds, const f16* means, int ldm, f16* ysf, int ldsf,
                          const f16* q, int qstride, const int* qp, f16* yout, int ldo, int16_t* sym, uint8_t* cond,
                          const float* th, float thr, int step, hipStream_t s) {
    hipLaunchKernelGGL(quarter_enc_step_kernel<false>, grid1d(y.pixels() * (y.C / 4)), dim3(256), 0, s, y, scales, lds,
                       means, ldm, ysf, ldsf, q, qstride, qp, yout, ldo, sym, cond, th, thr, step);
    return hipGetLastError();
}

hipError_t htl_enc_step(const Act& y, const f16* scales, int lds, const f16* means, int ldm, const f16* qdec, int ldq,
                        f16* ysf, int ldsf, f16* yout, int ldo, int16_t* sym, uint8_t* cond, const float* th, float thr,
                        int step, hipStream_t s) {
    hipLaunchKernelGGL(quarter_enc_step_kernel<true>, grid1d(y.pixels() * (y.C / 4)), dim3(256), 0, s, y, scales, lds,
                       means, ldm, ysf, ldsf, qdec, ldq, nullptr, yout, ldo, sym, cond, th, thr, step);
    return hipGetLastError();
}

__global__ void intra_dec_index_kernel(Act g, const f16* __restrict__ scales, int lds, const float* __restrict__ th,
                                       float thr, int step, uint8_t* __restrict__ idx, uint8_t* __restrict__ cond) {
    const int C4 = g.C >> 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.pixels() * C4) return;
    const long m = i / C4;
    const int cq = (int)(i % C4);
    const int w = (int)(m % g.W);
    const int h = (int)((m / g.W) % g.H);
    const int ga = mask4_group(step, ((h & 1) << 1) | (w & 1));
    const float sc = (float)scales[m * lds + ga * C4 + cq];
    idx[i] = (uint8_t)scale_index(sc, th);
    cond[i] = sc > thr ? 1 : 0;
}

hipError_t intra_dec_index(const Act& g, const f16* scales, int lds, const float* th, float thr, int step,
                           uint8_t* idx, uint8_t* cond, hipStream_t s) {
    hipLaunchKernelGGL(intra_dec_index_kernel, grid1d(g.pixels() * (g.C / 4)), dim3(256), 0, s, g, scales, lds, th,
                       thr, step, idx, cond);
    return hipGetLastError();
}

template <bool kElemQ>
__global__ void quarter_dec_step_kernel(Act g, const int8_t* __restrict__ yq, const f16* __restrict__ means, int ldm,
                                        f16* __restrict__ ysf, int ldsf, const f16* __restrict__ q, int ldq,
                                        const int* __restrict__ qp, f16* __restrict__ yout, int ldo, int step) {
    const int C4 = g.C >> 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.pixels() * C4) return;
    const long m = i / C4;
    const int cq = (int)(i % C4);
    const int w = (int)(m % g.W);
    const int h = (int)((m / g.W) % g.H);
    const int ga = mask4_group(step, ((h & 1) << 1) | (w & 1));
    const int ca = ga * C4 + cq;
    const f16 yh = yhat_elem((int)yq[i], (float)means[m * ldm + ca]);
    if (step < 3) {
        if (step == 0) {
#pragma unroll
            for (int gg = 0; gg < 4; ++gg) ysf[m * ldsf + gg * C4 + cq] = (gg == ga) ? yh : (f16)0.f;
        } else {
            ysf[m * ldsf + ca] = yh;
        }
    } else {
        const f16* qr = kElemQ ? q + m * ldq : q + (size_t)(qp ? *qp : 0) * ldq;
#pragma unroll
        for (int gg = 0; gg < 4; ++gg) {
            const int c = gg * C4 + cq;
            const float v = (gg == ga) ? (float)yh : (float)ysf[m * ldsf + c];
            const float qc = kElemQ ? qstep((float)qr[c]) : (float)qr[c];
            yout[m * ldo + c] = (f16)(v * qc);
        }
    }
}

hipError_t intra_dec_step(const Act& g, const int8_t* yq, const f16* means, int ldm, f16* ysf, int ldsf,
                          const f16* q, int qstride, const int* qp, f16* yout, int ldo, int step, hipStream_t s) {
    hipLaunchKernelGGL(quarter_dec_step_kernel<false>, grid1d(g.pixels() * (g.C / 4)), dim3(256), 0, s, g, yq, means,
                       ldm, ysf, ldsf, q, qstride, qp, yout, ldo, step);
    return hipGetLastError();
}

hipError_t htl_dec_step(const Act& g, const int8_t* yq, const f16* means, int ldm, const f16* qdec, int ldq, f16* ysf,
                        int ldsf, f16* yout, int ldo, int step, hipStream_t s) {
    hipLaunchKernelGGL(quarter_dec_step_kernel<true>, grid1d(g.pixels() * (g.C / 4)), dim3(256), 0, s, g, yq, means,
                       ldm, ysf, ldsf, qdec, ldq, nullptr, yout, ldo, step);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// stable stream compaction (conditional_index_part1/2, stream.cu:175-357) and
// expansion (conditional_recover_with_type_conversion, stream.cu:359-420).
// Tile = 256 threads x 16 consecutive elements; block offsets are an exclusive
// scan of per-tile counts, so the output is in NHWC-linear order.
// ---------------------------------------------------------------------------
constexpr int CPT = 16;  // elements per thread
static_assert(COMPACT_TILE == 256 * CPT, "tile");

DCVC_DEV int block_exclusive_scan(int v, int* sh, int* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int add = (t >= off) ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const int incl = sh[t];
    if (total) *total = sh[255];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(256) void compact_count_kernel(const uint8_t* __restrict__ cond, long N,
                                                            int* __restrict__ block_counts) {
    __shared__ int sh[256];
    const long base = (long)blockIdx.x * COMPACT_TILE + threadIdx.x * CPT;
    int c = 0;
    for (int e = 0; e < CPT; ++e) {
        const long i = base + e;
        if (i < N) c += cond[i] ? 1 : 0;
    }
    int total;
    block_exclusive_scan(c, sh, &total);
    if (threadIdx.x == 0) block_counts[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void compact_scan_kernel(int* __restrict__ offs, int nb, int* __restrict__ count) {
    __shared__ int sh[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    int local = 0;
    for (int e = 0; e < per; ++e) {
        const int b = t * per + e;
        if (b < nb) local += offs[b];
    }
    sh[t] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int add = (t >= off) ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    int run = sh[t] - local;
    if (t == 1023) *count = sh[1023];
    for (int e = 0; e < per; ++e) {
        const int b = t * per + e;
        if (b < nb) {
            const int v = offs[b];
            offs[b] = run;
            run += v;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void compact_scatter_kernel(const T* __restrict__ vals,
                                                              const uint8_t* __restrict__ cond, long N,
                                                              const int* __restrict__ offs, T* __restrict__ out) {
    __shared__ int sh[256];
    const long base = (long)blockIdx.x * COMPACT_TILE + threadIdx.x * CPT;
    int c = 0;
    for (int e = 0; e < CPT; ++e) {
        const long i = base + e;
        if (i < N) c += cond[i] ? 1 : 0;
    }
    int dst = offs[blockIdx.x] + block_exclusive_scan(c, sh, nullptr);
    for (int e = 0; e < CPT; ++e) {
        const long i = base + e;
        if (i < N && cond[i]) out[dst++] = vals[i];
    }
}

__global__ __launch_bounds__(256) void recover_kernel(const int8_t* __restrict__ dec, const uint8_t* __restrict__ cond,
                                                      long N, const int* __restrict__ offs,
                                                      int8_t* __restrict__ out) {
    __shared__ int sh[256];
    const long base = (long)blockIdx.x * COMPACT_TILE + threadIdx.x * CPT;
    int c = 0;
    for (int e = 0; e < CPT; ++e) {
        const long i = base + e;
        if (i < N) c += cond[i] ? 1 : 0;
    }
    int src = offs[blockIdx.x] + block_exclusive_scan(c, sh, nullptr);
    for (int e = 0; e < CPT; ++e) {
        const long i = base + e;
        if (i < N) out[i] = cond[i] ? dec[src++] : (int8_t)0;
    }
}

hipError_t compact(const void* vals, int esize, const uint8_t* cond, long N, void* out, int* offs, int* count,
                   hipStream_t s) {
    const int nb = compact_blocks(N);
    if (nb > 0) hipLaunchKernelGGL(compact_count_kernel, dim3(nb), dim3(256), 0, s, cond, N, offs);
    hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(1024), 0, s, offs, nb, count);
    if (nb == 0) return hipGetLastError();
    if (esize == 2)
        hipLaunchKernelGGL(compact_scatter_kernel<int16_t>, dim3(nb), dim3(256), 0, s, (const int16_t*)vals, cond, N,
                           offs, (int16_t*)out);
    else
        hipLaunchKernelGGL(compact_scatter_kernel<uint8_t>, dim3(nb), dim3(256), 0, s, (const uint8_t*)vals, cond, N,
                           offs, (uint8_t*)out);
    return hipGetLastError();
}

hipError_t recover(const int8_t* dec, const uint8_t* cond, long N, const int* offs, int8_t* out, hipStream_t s) {
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(recover_kernel, dim3(compact_blocks(N)), dim3(256), 0, s, dec, cond, N, offs, out);
    return hipGetLastError();
}

}  // namespace dcvc
