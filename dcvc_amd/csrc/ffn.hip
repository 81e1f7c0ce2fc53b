// Fused FFN half of a DepthConvBlock (layers.py:128-159), which the
// reference's proxy runs as two GEMM launches (layers_proxy.cpp:84-97):
//
//   t1[m, n] = sum_j wsilu(o[m, :] . W0[4n + j, :] + b0[4n + j])   (ffn.0 + WSiLU chunk add)
//   D[m, c]  = tail(t1[m, :] . W2[c, :] + b2[c] + o[m, c])          (ffn.2 + shortcut)
//
// tail: nothing, + S2 (second shortcut) or * q[c] (quant).  One workgroup owns
// 128 pixels and runs both products for them; t1 (128 x CM fp16) stays in LDS
// instead of making an HBM round trip, and the second product starts without
// a launch boundary.
//
// Results are bit-identical to the two-launch path (gemm.hip EPI_QUAD, then
// the shortcut epilogue): every output element is accumulated by one lane on
// v_mfma_f32_32x32x16_f16 in increasing k, 16 at a time, with the same
// fragment k order, and both epilogues apply the same operations in the same
// order in fp32 (t1 rounded to fp16 exactly where the quad GEMM stores it).
//
// Layout (4 waves in a 2 x 2 grid, one workgroup per CU, 160 KB LDS):
//   stage 1, quad GEMM over K = C: passes of 256 ffn.0 rows (64 hidden
//     channels); wave tile 64 pixels x 128 rows (the 4 rows of each hidden
//     channel in one wave).  The K tiles of all passes form one DMA stream
//     through a 4-slot ring (A and B tiles, direct global->LDS loads, 3 tiles
//     ahead), so a pass boundary does not drain it.  Each pass ends with the
//     quad sum in registers, written as fp16 into t1 in the swizzled K-tile
//     layout the second product reads its A fragments from;
//   stage 2, ffn.2 over K = CM: all C outputs in one pass, wave tile
//     64 pixels x C/2, A fragments from t1, B (W2) K tiles through the ring;
//   epilogue: fp32 (acc + b2) staged through LDS 64 rows at a time, each lane
//     finishing 8 consecutive channels of one pixel with 16-byte accesses.
#include "common.h"
#include "ffn.h"
#include "gemm.h"
#include "gemm_dev.h"

namespace dcvc {

namespace {

int g_ffn_mode = 0;

template <int C_, int CM_, int EPI_>
struct FfnCfg {
    static constexpr int C = C_, CM = CM_, EPI = EPI_;
    static constexpr int BM = 128;
    static constexpr int kThreads = 256;
    static constexpr int SM = 2;                      // 32-row sub-tiles per wave (64 rows)
    // stage 1
    static constexpr int SN1 = 4;
    static constexpr int BN1 = 256;                   // ffn.0 rows per pass
    static constexpr int P1 = CM / 64;
    static constexpr int K1 = C / kBK;
    static constexpr int G1 = P1 * K1;                // tiles of the stage-1 stream
    static constexpr int AI = BM / 4 / 16;            // A DMAs per lane per tile
    static constexpr int BI1 = BN1 / 4 / 16;          // B DMAs per lane per tile
    static constexpr int kDma1 = AI + BI1;
    static constexpr int kSlot1 = (BM + BN1) * kRowBytes;
    static constexpr int kStages1 = 4, kAhead1 = 3;   // refills the slot of tile t - 1
    // stage 2
    static constexpr int SN2 = C / 64;
    static constexpr int K2 = CM / kBK;
    static constexpr int BI2 = C / 4 / 16;
    static constexpr int kSlot2 = C * kRowBytes;
    static constexpr int kRing = kStages1 * kSlot1;
    static constexpr int kStages2 = (kRing / kSlot2 >= 5) ? 5 : kRing / kSlot2;
    static constexpr int kAhead2 = (kStages2 >= 5) ? kStages2 - 2 : kStages2 - 1;
    static constexpr bool kWaitLds2 = (kAhead2 == kStages2 - 1);
    // t1 after the ring: CM / 32 K tiles of BM 64-byte rows
    static constexpr int kT1 = kRing;
    static constexpr int kLds = kRing + BM * CM * 2;
    static constexpr int kEPitch = C + 4;
    static constexpr int kEChunks = 64 * C / 8 / kThreads;  // per lane per epilogue pass
    static_assert(C % 256 == 0 && CM % 64 == 0, "shape");
    static_assert(kLds <= 160 * 1024, "LDS");
    static_assert(64 * kEPitch * 4 <= kLds, "epilogue staging");
    static_assert(kStages2 >= 3 && kStages2 * kSlot2 <= kRing, "stage-2 ring");
    static_assert((kAhead1 - 1) * kDma1 < 64 && (kAhead2 - 1) * BI2 < 64, "vmcnt range");
};

template <class F>
__global__ __launch_bounds__(256, 1) void ffn_kernel(FfnArgs g, uint32_t a_bytes) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    constexpr int C = F::C, CM = F::CM, BM = F::BM, SM = F::SM, SN1 = F::SN1, SN2 = F::SN2, NT = F::kThreads;
    constexpr int EPI = F::EPI;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 1;
    const int wn = wave >> 1;
    const int m0 = (int)blockIdx.x * BM;
    const int lr = lane & 31;
    // DMA lane geometry: row dr of a 16-row group, K chunk kc (swizzled)
    const int dr = lane >> 2;
    const int kc = (lane & 3) ^ swz(dr);
    const int pos0 = ((lane >> 5) ^ swz(lr)) * 16;
    const int pos1 = ((2 + (lane >> 5)) ^ swz(lr)) * 16;

    const __amdgpu_buffer_rsrc_t rA = make_rsrc(g.A, a_bytes);
    const __amdgpu_buffer_rsrc_t rW0 = make_rsrc(g.W0, (uint32_t)(4 * CM * C * 2));
    const __amdgpu_buffer_rsrc_t rW2 = make_rsrc(g.W2, (uint32_t)(C * CM * 2));

    // ---------------- stage 1: quad GEMM into t1 ----------------
    uint32_t a_off[F::AI];
#pragma unroll
    for (int q = 0; q < F::AI; ++q) {
        const int m = m0 + wave * (BM / 4) + q * 16 + dr;
        a_off[q] = (m < g.M) ? (uint32_t)(((size_t)m * g.lda + kc * 8) * 2) : kOOB;
    }
    // LDS B row -> wave column row / 128, slab (row / 32) % 4, hidden channel
    // (row / 128) * 32 + row % 32 of the pass: ffn.0 row 4 n + slab
    uint32_t w0_off[F::BI1];
#pragma unroll
    for (int q = 0; q < F::BI1; ++q) {
        const int row = wave * (F::BN1 / 4) + q * 16 + dr;
        const int n = (row >> 7) * 32 + (row & 31);
        w0_off[q] = (uint32_t)(((size_t)(4 * n + ((row >> 5) & 3)) * C + kc * 8) * 2);
    }
    constexpr uint32_t kPassBytes = 256u * C * 2;
    // DMAs of stream tile gt (pass gt / K1, K tile gt % K1); tiles past the
    // stream read zeros into a slot that is never read
    auto dma1 = [&](int gt) {
        const bool live = gt < F::G1;
        const uint32_t past = live ? 0u : kOOB;
        const int p = live ? gt / F::K1 : 0;
        const uint32_t kb = live ? (uint32_t)(gt - p * F::K1) * kRowBytes : 0u;
        char* slot = smem + (gt % F::kStages1) * F::kSlot1;
#pragma unroll
        for (int q = 0; q < F::AI; ++q)
            dma16(rA, slot + (wave * (BM / 4) + q * 16) * kRowBytes, (a_off[q] + kb) | past);
#pragma unroll
        for (int q = 0; q < F::BI1; ++q)
            dma16(rW0, slot + (BM + wave * (F::BN1 / 4) + q * 16) * kRowBytes,
                  (w0_off[q] + (uint32_t)p * kPassBytes + kb) | past);
    };

    // ffn.0 biases of every pass (loaded before the first DMA)
    float b1r[F::P1][4];
#pragma unroll
    for (int p = 0; p < F::P1; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j) b1r[p][j] = g.b0[4 * (p * 64 + wn * 32 + lr) + j];

    f32x16 acc[SM][SN1];
    f16x8 f0a[SM], f0b[SN1], f1a[SM], f1b[SN1];
    const int fa1 = (wm * 64 + lr) * kRowBytes;
    const int fb1 = (BM + wn * 128 + lr) * kRowBytes;
    auto read1 = [&](int gt, int pos, f16x8* fa, f16x8* fb) {
        const char* base = smem + (gt % F::kStages1) * F::kSlot1 + pos;
#pragma unroll
        for (int i = 0; i < SM; ++i) fa[i] = *reinterpret_cast<const f16x8*>(base + fa1 + i * 32 * kRowBytes);
#pragma unroll
        for (int j = 0; j < SN1; ++j) fb[j] = *reinterpret_cast<const f16x8*>(base + fb1 + j * 32 * kRowBytes);
    };
    auto mfma1 = [&](const f16x8* fa, const f16x8* fb) {
#pragma unroll
        for (int i = 0; i < SM; ++i)
#pragma unroll
            for (int j = 0; j < SN1; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    };
    constexpr int WAIT1 = (F::kAhead1 - 1) * F::kDma1;

#pragma unroll
    for (int st = 0; st < F::kAhead1; ++st) dma1(st);
#pragma unroll
    for (int p = 0; p < F::P1; ++p) {
        const int g0 = p * F::K1;
#pragma unroll
        for (int i = 0; i < SM; ++i)
#pragma unroll
            for (int j = 0; j < SN1; ++j)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;
        // tile g0 landed for every wave; every read of the slot refilled next is done
        wait_tile_and_barrier<WAIT1, true>();
        dma1(g0 + F::kAhead1);
        read1(g0, pos0, f0a, f0b);
        read1(g0, pos1, f1a, f1b);
        mfma1(f0a, f0b);
#pragma unroll 1
        for (int kt = 1; kt < F::K1; ++kt) {
            wait_tile_and_barrier<WAIT1, true>();
            dma1(g0 + kt + F::kAhead1);
            read1(g0 + kt, pos0, f0a, f0b);
            mfma1(f1a, f1b);
            interleave<SM + SN1, SM * SN1, F::kDma1>();
            read1(g0 + kt, pos1, f1a, f1b);
            mfma1(f0a, f0b);
            interleave<SM + SN1, SM * SN1, 0>();
        }
        mfma1(f1a, f1b);
        // quad sum of hidden channel n = p * 64 + wn * 32 + lr for the wave's
        // 64 pixels -> t1 K tile n / 32, 16-byte chunk (n % 32) / 8 of each row
        char* t1 = smem + F::kT1 + (2 * p + wn) * BM * kRowBytes + (lr & 7) * 2;
        const int c8 = lr >> 3;
#pragma unroll
        for (int i = 0; i < SM; ++i)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int r = wm * 64 + i * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                float x = 0.f;
#pragma unroll
                for (int j = 0; j < 4; ++j) x += wsilu(acc[i][j][v] + b1r[p][j]);
                // the quad GEMM rounds the sum to fp32 (its LDS staging) before
                // fp16; keep the compiler from merging the last fma and the
                // conversion into one fp16-rounded v_fma_mix
                asm volatile("" : "+v"(x));
                *reinterpret_cast<f16*>(t1 + r * kRowBytes + ((c8 ^ swz(r)) << 4)) = (f16)x;
            }
    }
    // every DMA of the stage-1 stream (the zero tiles past its end too) has
    // landed and t1 is complete before the ring changes geometry
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    // ---------------- stage 2: ffn.2 over t1 ----------------
    float b2r[SN2];
#pragma unroll
    for (int j = 0; j < SN2; ++j) b2r[j] = g.b2[wn * (C / 2) + j * 32 + lr];
    // epilogue chunk (pass i, t): lane finishes channels cc .. cc + 7 of pixel m
    constexpr int EC = F::kEChunks;
    auto chunk_mn = [&](int i, int t, int& rl, int& cc, int& m) {
        const int c = tid + t * NT;
        rl = c / (C / 8);
        cc = (c - rl * (C / 8)) * 8;
        m = m0 + (rl >> 5) * 64 + i * 32 + (rl & 31);
    };
    // first shortcut (the block input itself): the chunks of epilogue pass 0
    // are loaded before the K loop when they fit in 32 registers, the others
    // at the start of their pass
    constexpr bool kS1Early = EC <= 8;
    f16x8 s1r[EC];
    auto load_s1 = [&](int i) {
#pragma unroll
        for (int t = 0; t < EC; ++t) {
            int rl, cc, m;
            chunk_mn(i, t, rl, cc, m);
            const f16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
            s1r[t] = (m < g.M) ? *reinterpret_cast<const f16x8*>(g.A + (size_t)m * g.lda + cc) : zero8;
        }
    };
    if constexpr (kS1Early) load_s1(0);

    uint32_t w2_off[F::BI2];
#pragma unroll
    for (int q = 0; q < F::BI2; ++q) {
        const int row = wave * (C / 4) + q * 16 + dr;
        w2_off[q] = (uint32_t)(((size_t)row * CM + kc * 8) * 2);
    }
    auto dma2 = [&](int kt) {
        const uint32_t past = (kt < F::K2) ? 0u : kOOB;
        const uint32_t kb = (kt < F::K2) ? (uint32_t)kt * kRowBytes : 0u;
        char* slot = smem + (kt % F::kStages2) * F::kSlot2 + wave * (C / 4) * kRowBytes;
#pragma unroll
        for (int q = 0; q < F::BI2; ++q) dma16(rW2, slot + q * 16 * kRowBytes, (w2_off[q] + kb) | past);
    };
    f32x16 acc2[SM][SN2];
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
        for (int j = 0; j < SN2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc2[i][j][v] = 0.f;
    f16x8 h0a[SM], h0b[SN2], h1a[SM], h1b[SN2];
    const int fa2 = (wm * 64 + lr) * kRowBytes;
    const int fb2 = (wn * (C / 2) + lr) * kRowBytes;
    auto read2 = [&](int kt, int pos, f16x8* fa, f16x8* fb) {
        const char* ta = smem + F::kT1 + kt * BM * kRowBytes + pos;
        const char* tb = smem + (kt % F::kStages2) * F::kSlot2 + pos;
#pragma unroll
        for (int i = 0; i < SM; ++i) fa[i] = *reinterpret_cast<const f16x8*>(ta + fa2 + i * 32 * kRowBytes);
#pragma unroll
        for (int j = 0; j < SN2; ++j) fb[j] = *reinterpret_cast<const f16x8*>(tb + fb2 + j * 32 * kRowBytes);
    };
    auto mfma2 = [&](const f16x8* fa, const f16x8* fb) {
#pragma unroll
        for (int i = 0; i < SM; ++i)
#pragma unroll
            for (int j = 0; j < SN2; ++j)
                acc2[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[i], fb[j], acc2[i][j], 0, 0, 0);
    };
    constexpr int WAIT2 = (F::kAhead2 - 1) * F::BI2;
#pragma unroll
    for (int st = 0; st < F::kAhead2; ++st) dma2(st);
    wait_tile_and_barrier<WAIT2, F::kWaitLds2>();
    dma2(F::kAhead2);
    read2(0, pos0, h0a, h0b);
    read2(0, pos1, h1a, h1b);
    mfma2(h0a, h0b);
#pragma unroll 1
    for (int kt = 1; kt < F::K2; ++kt) {
        wait_tile_and_barrier<WAIT2, F::kWaitLds2>();
        dma2(kt + F::kAhead2);
        read2(kt, pos0, h0a, h0b);
        mfma2(h1a, h1b);
        interleave<SM + SN2, SM * SN2, F::BI2>();
        read2(kt, pos1, h1a, h1b);
        mfma2(h0a, h0b);
        interleave<SM + SN2, SM * SN2, 0>();
    }
    mfma2(h1a, h1b);
    __syncthreads();  // ring and t1 are free: the epilogue staging reuses them

    // ---------------- epilogue ----------------
    const f16* q = nullptr;
    if constexpr (EPI == EPI_SHORTCUT_QUANT) {
        const int qp = g.qp ? *g.qp : 0;
        q = g.q + (size_t)qp * g.qstride;
    }
    float* E = reinterpret_cast<float*>(smem);
    constexpr int EP = F::kEPitch;
#pragma unroll
    for (int i = 0; i < SM; ++i) {
        if (i > 0) __syncthreads();  // previous pass fully read back
        if (!kS1Early || i > 0) load_s1(i);
#pragma unroll
        for (int j = 0; j < SN2; ++j) {
            const int col = wn * (C / 2) + j * 32 + lr;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = wm * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                E[row * EP + col] = acc2[i][j][v] + b2r[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < EC; ++t) {
            int rl, cc, m;
            chunk_mn(i, t, rl, cc, m);
            if (m >= g.M) continue;
            const f32x4 lo = *reinterpret_cast<const f32x4*>(E + rl * EP + cc);
            const f32x4 hi = *reinterpret_cast<const f32x4*>(E + rl * EP + cc + 4);
            float x[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] += (float)s1r[t][e];
            if constexpr (EPI == EPI_SHORTCUT2) {
                const f16x8 s = *reinterpret_cast<const f16x8*>(g.S2 + (size_t)m * g.lds2 + cc);
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] += (float)s[e];
            }
            if constexpr (EPI == EPI_SHORTCUT_QUANT) {
                if (g.q_round) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) x[e] = (float)(f16)x[e];
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] *= (float)q[cc + e];
            }
            f16x8 o;
#pragma unroll
            for (int e = 0; e < 8; ++e) o[e] = (f16)x[e];
            *reinterpret_cast<f16x8*>(g.D + (size_t)m * g.ldd + cc) = o;
        }
    }
}

template <int C, int CM, int EPI>
hipError_t launch_ffn(const FfnArgs& f, hipStream_t stream) {
    using F = FfnCfg<C, CM, EPI>;
    auto kern = ffn_kernel<F>;
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, F::kLds);
    if (attr != hipSuccess) return attr;
    const uint32_t a_bytes = (uint32_t)(((size_t)(f.M - 1) * f.lda + C) * 2);
    hipLaunchKernelGGL(kern, dim3(ffn_blocks(f.M)), dim3(F::kThreads), F::kLds, stream, f, a_bytes);
    return hipGetLastError();
}

template <int C, int CM>
hipError_t dispatch_epi(const FfnArgs& f, int epi, hipStream_t stream) {
    switch (epi) {
    case EPI_SHORTCUT: return launch_ffn<C, CM, EPI_SHORTCUT>(f, stream);
    case EPI_SHORTCUT2: return launch_ffn<C, CM, EPI_SHORTCUT2>(f, stream);
    case EPI_SHORTCUT_QUANT: return launch_ffn<C, CM, EPI_SHORTCUT_QUANT>(f, stream);
    default: return hipErrorInvalidValue;
    }
}

bool aligned16(const void* p) { return ((uintptr_t)p) % 16 == 0; }

}  // namespace

bool ffn_supported(int C, int CM) { return CM == 256 && (C == 256 || C == 512); }

int ffn_validate(const FfnArgs& f, int epi) {
    if (!f.A || !f.W0 || !f.b0 || !f.W2 || !f.b2 || !f.D) return DCVC_ERR_ARG;
    if (!ffn_supported(f.C, f.CM) || f.M <= 0) return DCVC_ERR_ARG;
    if (f.lda < f.C || f.lda % 8 != 0 || f.ldd < f.C || f.ldd % 8 != 0) return DCVC_ERR_ARG;
    if (!aligned16(f.A) || !aligned16(f.D) || !aligned16(f.W0) || !aligned16(f.W2)) return DCVC_ERR_ARG;
    // buffer addressing uses 32-bit byte offsets below 2^31
    if (((size_t)(f.M - 1) * f.lda + f.C) * 2 >= kOOB) return DCVC_ERR_ARG;
    if (epi == EPI_SHORTCUT2) {
        if (!f.S2 || f.lds2 < f.C || f.lds2 % 8 != 0 || !aligned16(f.S2)) return DCVC_ERR_ARG;
        if (f.D == f.S2 && f.ldd != f.lds2) return DCVC_ERR_ARG;
    } else if (epi == EPI_SHORTCUT_QUANT) {
        if (!f.q) return DCVC_ERR_ARG;
    } else if (epi != EPI_SHORTCUT) {
        return DCVC_ERR_ARG;
    }
    // in place on the block input: each workgroup reads only its own rows
    if (f.D == f.A && f.ldd != f.lda) return DCVC_ERR_ARG;
    return DCVC_OK;
}

hipError_t ffn_launch(const FfnArgs& f, int epi, hipStream_t stream) {
    if (f.CM != 256) return hipErrorInvalidValue;
    if (f.C == 512) return dispatch_epi<512, 256>(f, epi, stream);
    if (f.C == 256) return dispatch_epi<256, 256>(f, epi, stream);
    return hipErrorInvalidValue;
}

int ffn_set_mode(int mode) {
    if (mode < 0 || mode > 2) return DCVC_ERR_ARG;
    g_ffn_mode = mode;
    return DCVC_OK;
}

int ffn_mode() { return g_ffn_mode; }

}  // namespace dcvc
