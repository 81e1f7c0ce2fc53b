// MFMA GEMM family for the 1x1-conv / implicit-GEMM conv layers of DCVC-UF.
//
//   D[m, n] = epilogue( sum_k A[m, k] * B[n, k] + bias[n] )
//
// A: NHWC fp16 activations (row m = one output pixel, row stride lda), or an
//    implicit im2col view of them for the strided convs.
// B: fp16 weights, row n = output channel, K contiguous (PyTorch conv layout
//    [N, K, 1, 1]; the strided convs use tap-major K = tap * Cin + c).
// Accumulation: fp32 on v_mfma_f32_32x32x16_f16.  Each output element is
// accumulated by one lane in increasing k, 16 at a time, whatever the tile
// configuration, so results do not depend on the tile shape, the grid or the
// batch composition (every configuration below is bit-identical).
//
// Structure (gfx950, wave64):
//   * block tile BM x BN (B rows), 4 waves in a WM x WN grid, wave tile up to
//     128 x 128 (16 MFMA accumulators of 32x32 in the accumulation registers),
//     so every k16 step issues 8 LDS fragment reads for 16 MFMAs;
//   * operands stream into a 4-stage LDS ring (K tile 32) through
//     buffer_load_dwordx4 ... lds (direct global->LDS, no register staging),
//     three K tiles in flight; out-of-range rows and conv padding taps read as
//     zero through the buffer range check.  Each 64-byte LDS row holds its
//     four 16-byte K chunks XOR-swizzled by bits 1-3 of the row, so the
//     fragment reads of 8 (or 16) consecutive rows hit distinct banks;
//   * one barrier per K tile, preceded by an explicit vmcnt wait for exactly
//     the tile about to be read; fragment reads run one k16 step ahead of the
//     MFMAs, across the barrier, so the matrix pipe stays fed while waves
//     synchronise;
//   * epilogue through LDS: bias / WSiLU / quad-sum in fp32 registers, the
//     fp32 tile staged in LDS 32 rows per wave at a time, then each lane
//     finishes 8 consecutive channels of one pixel (shortcut, second
//     shortcut, per-channel quant) with 16-byte loads and stores, so every
//     output row segment is written whole;
//   * configurations 8-11 take 64-deep K tiles (128-byte LDS rows, one cache
//     line per row and tile): half the line requests per byte, for the
//     same per-element accumulation order;
//   * XCD-aware tile order: hardware hands block b to XCD b % 8, so
//     consecutive tiles (the N tiles of one M tile, which share A) are
//     assigned to blocks of the same XCD and meet in its L2.
//
// Reference operators replaced (def_cutlass.h:10-39):
//   conv1x1_bias, conv1x1_bias_wsilu, conv1x1_bias_shortcut(2),
//   conv1x1_bias_shortcut_with_quant, conv1x1_bias_with_quant,
//   conv1x1_bias_wsilu_chunk_add (quad GEMM), conv_bias (k=2/3, s=2),
//   transposed_conv (2x2 s2).  Epilogue op order follows
//   cutlass/cutlass_epilogue.h:79-131,175-224: (acc + b) [-> wsilu] [+ S1]
//   [+ S2 | * q].
#include <atomic>
#include <mutex>
#include <shared_mutex>
#include <unordered_map>
#include <utility>

#include "common.h"
#include "gemm.h"
#include "gemm_dev.h"
#include "fp8_dev.h"

namespace dcvc {

namespace {

int g_tile_policy = 0;  // 0 automatic, otherwise a forced configuration id

// Tuned tile choices, (prec, epi, amode, dmode, M, N, K) -> configuration id,
// installed from the autotuner's database.  Read at launch (under graph
// capture: once, when the graph is recorded); shapes with no entry use the
// built-in heuristic.
std::shared_mutex g_tune_mu;
std::unordered_map<uint64_t, int> g_tune;

bool tune_key(int prec, int epi, int amode, int dmode, long long M, long long N, long long K, uint64_t* key) {
    if (prec < 0 || prec > 3 || epi < 0 || epi > 15 || amode < 0 || amode > 3 || dmode < 0 || dmode > 1) return false;
    if (M <= 0 || M >= (1LL << 24) || N <= 0 || N >= (1LL << 15) || K <= 0 || K >= (1LL << 16)) return false;
    *key = ((uint64_t)prec << 62) | ((uint64_t)epi << 58) | ((uint64_t)amode << 56) | ((uint64_t)dmode << 55) |
           ((uint64_t)M << 31) | ((uint64_t)N << 16) | (uint64_t)K;
    return true;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

// 8 fp16 -> 8 bf16 (through fp32, round to nearest even)
DCVC_DEV bf16x8 to_bf16x8(f16x8 v) {
    bf16x8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o[e] = (__bf16)(float)v[e];
    return o;
}

// 32 fp16 (k order) -> 32 OCP e4m3 bytes in the same order (fp8x8_from_f16x8)
DCVC_DEV i32x8 to_fp8x32(const f16x8 (&v)[4]) {
    i32x8 o;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int2 p = fp8x8_from_f16x8(v[c]);
        o[2 * c] = p.x;
        o[2 * c + 1] = p.y;
    }
    return o;
}

}  // namespace

template <int BM_, int BN_, int WM_, int WN_, int STAGES_, int EPI_, int AMODE_, int DMODE_, int PREC_ = PREC_F16,
          int KB_ = 32>
struct Tile {
    static constexpr int BM = BM_, BN = BN_, WM = WM_, WN = WN_;
    static constexpr int PREC = PREC_;
    static constexpr bool kFp8 = (PREC_ == PREC_FP8 || PREC_ == PREC_FP8A);
    static constexpr bool kA8 = (PREC_ == PREC_FP8A);  // A is e4m3 bytes in memory
    // K per ring tile: 32 (one 64-byte fp16 row), or 64 for fp8, whose B row
    // is 64 fp8 bytes and whose A tile is two fp16 half tiles (one 64-byte
    // row of e4m3 bytes under kA8)
    // fp16 / bf16 rings take K tiles of KB_ = 32 (64-byte LDS rows, 16 rows
    // per DMA wave instruction) or 64 (128-byte rows, 8 rows per
    // instruction).  A 128-byte row is one whole cache line per row and K
    // tile, so each DMA instruction requests half as many lines; the plain
    // load stream of the quad GEMM runs about 1.45x faster that way.
    static constexpr int kKT = kFp8 ? 64 : KB_;
    static constexpr int kRB = kFp8 ? 64 : 2 * KB_;    // LDS row bytes
    static constexpr int kLPR = kRB / 16;              // DMA lanes per row
    static constexpr int kRPI = 64 / kLPR;             // rows per DMA instruction
    static constexpr int kSteps = kFp8 ? 1 : KB_ / 16;  // k16 MFMA steps per K tile
    static constexpr int kAH = (kFp8 && !kA8) ? 2 : 1;
    // LDS ring of STAGES K tiles, STAGES - 2 in flight ahead of the one being
    // read: the slot refilled at iteration kt is that of tile kt - 2, whose
    // fragment reads every wave consumed before this iteration's barrier.
    // STAGES == 3 keeps one tile ahead and refills the slot of tile kt - 1,
    // so its barrier also waits for the wave's own fragment reads (lgkmcnt).
    // (fp8 reads a whole tile's fragments one iteration ahead of their
    // MFMAs; its barrier waits for them, and STAGES - 1 tiles are in flight)
    static constexpr int kStages = STAGES_;
    static constexpr int kAhead = (STAGES_ >= 5 && !kFp8) ? STAGES_ - 2 : STAGES_ - 1;
    static constexpr bool kWaitLds = kFp8 || (kAhead == STAGES_ - 1);
    // configurations meant to run 2 blocks per CU keep within 256 registers
    // per lane (architectural + accumulation) so two waves share each SIMD
    // (and 8-wave configurations run two waves on every SIMD of their CU)
    static constexpr int kWavesPerSimd =
        (WM_ * WN_ >= 8 || (STAGES_ == 3 && !kFp8) || (BM_ * BN_ <= 128 * 128)) ? 2 : 1;
    static constexpr int EPI = EPI_, AMODE = AMODE_, DMODE = DMODE_;
    static constexpr int kWaves = WM * WN;
    static constexpr int kThreads = kWaves * 64;
    static constexpr int TM = BM / WM, TN = BN / WN;
    static constexpr int SM = TM / 32, SN = TN / 32;
    static constexpr bool kQuad = (EPI == EPI_QUAD);
    static constexpr int BNO = kQuad ? BN / 4 : BN;  // output columns per block
    static constexpr int TNO = kQuad ? TN / 4 : TN;  // output columns per wave
    static constexpr int ARW = BM / kWaves;           // A rows each wave loads per K tile
    static constexpr int BRW = BN / kWaves;
    static constexpr int AI = ARW / kRPI;             // DMA instructions (kRPI rows each)
    static constexpr int BI = BRW / kRPI;
    static constexpr int kDma = AI * kAH + BI;        // per lane per K tile
    static constexpr int kStageBytes = (BM * kAH + BN) * kRB;
    static constexpr int kEPitch = BNO + 4;           // epilogue fp32 row pitch
    static constexpr int kERows = WM * 32;            // rows staged per epilogue pass
    static constexpr int kEChunks = kERows * BNO / 8;
    static constexpr size_t kLdsBytes = (kStages * kStageBytes > kERows * kEPitch * 4)
                                            ? (size_t)kStages * kStageBytes
                                            : (size_t)kERows * kEPitch * 4;
    static_assert(TM % 32 == 0 && TN % 32 == 0, "wave tile");
    static_assert(!kQuad || SN % 4 == 0, "quad GEMM needs the 4 B slabs of an output in one wave");
    static_assert(ARW % kRPI == 0 && BRW % kRPI == 0 && AI > 0 && BI > 0, "DMA split");
    static_assert(KB_ == 32 || (KB_ == 64 && !kFp8 && AMODE_ == A_PLAIN), "64-deep K tiles: plain fp16 / bf16 A");
    static_assert(kEChunks % kThreads == 0, "epilogue split");
    static_assert(kAhead * kDma < 64, "vmcnt range");
    // two slots (64-deep K tiles only): the slot refilled at iteration kt is
    // that of tile kt - 1, whose fragment reads the barrier's lgkmcnt(0) retired
    static_assert((kStages >= 3 && kAhead >= 2) || (KB_ == 64 && kStages == 2 && kAhead == 1), "ring depth");
    static_assert(kLdsBytes <= 160 * 1024, "LDS");
};

template <class T>
__global__ __launch_bounds__(T::kThreads, T::kWavesPerSimd) void gemm_kernel(GemmArgs g, int tiles_n) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    if (g.batch > 1) {
        // problem blockIdx.y of a batched launch (uniform over the block)
        const long pb = (long)blockIdx.y;
        g.A += (pb / g.a_div) * g.sA;
        g.B = reinterpret_cast<const f16*>(reinterpret_cast<const char*>(g.B) + pb * g.sB * (T::kFp8 ? 1 : 2));
        g.bias += pb * g.sBias;
        if constexpr (T::kFp8) g.wscale += pb * g.sBias;
        if (g.S1) g.S1 += pb * g.sS1;
        if (g.S2) g.S2 += pb * g.sS2;
        g.D += pb * g.sD;
    }
    constexpr int BM = T::BM, BN = T::BN, SM = T::SM, SN = T::SN, NT = T::kThreads;
    constexpr int AMODE = T::AMODE, EPI = T::EPI;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % T::WM;
    const int wn = wave / T::WM;

    // ---- XCD-aware tile order: XCD x gets a contiguous run of tiles ----
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int xq = nb >> 3, xr = nb & 7, xcd = b & 7;
    const int tile = xcd * xq + (xcd < xr ? xcd : xr) + (b >> 3);
    const int mt = tile / tiles_n;
    const int nt = tile - mt * tiles_n;
    const int m0 = mt * BM;

    const __amdgpu_buffer_rsrc_t rA = make_rsrc(g.A, g.a_bytes);
    const __amdgpu_buffer_rsrc_t rB = make_rsrc(g.B, g.b_bytes);
    // DMA lane geometry: one wave instruction fills kRPI consecutive LDS rows
    // (1 KB); lane -> row dr of the group and 16-byte position dp, which
    // holds K chunk dp ^ swizzle(row of the operand's block)
    constexpr int RB = T::kRB;
    const int dr = lane / T::kLPR;
    const int dp = lane % T::kLPR;

    // ---- A rows of this lane's DMAs (fixed over the K loop) ----
    uint32_t a_off[T::AI];
    int a_kc[T::AI], a_pix[T::AI], a_h[T::AI], a_w[T::AI];
#pragma unroll
    for (int q = 0; q < T::AI; ++q) {
        const int r = wave * T::ARW + q * T::kRPI + dr;
        const int m = m0 + r;
        const int kc = dp ^ swz_rows<RB>(r);
        a_kc[q] = kc;
        if constexpr (AMODE == A_PLAIN && T::kA8) {
            a_off[q] = (m < g.M) ? (uint32_t)((size_t)m * g.lda + kc * 16) : kOOB;
        } else if constexpr (AMODE == A_PLAIN) {
            a_off[q] = (m < g.M) ? (uint32_t)(((size_t)m * g.lda + kc * 8) * 2) : kOOB;
        } else {
            const int hw = g.outH * g.outW;
            const int img = m / hw;
            const int rr = m - img * hw;
            const int ho = rr / g.outW;
            const int wo = rr - ho * g.outW;
            a_pix[q] = img * g.inH * g.inW;
            if constexpr (AMODE == A_S2K2) {
                a_h[q] = 2 * ho;
                a_w[q] = 2 * wo;
            } else if constexpr (AMODE == A_S2K3) {
                a_h[q] = 2 * ho - 1;
                a_w[q] = 2 * wo - 1;
            } else {
                a_h[q] = ho - 1;
                a_w[q] = wo - 1;
            }
            if (m >= g.M) a_h[q] = -(1 << 20);  // every tap out of range
            a_off[q] = 0;
        }
    }
    // ---- B rows of this lane's DMAs ----
    uint32_t b_off[T::BI];
#pragma unroll
    for (int q = 0; q < T::BI; ++q) {
        const int row = wave * T::BRW + q * T::kRPI + dr;
        const int kc = dp ^ swz_rows<RB>(row);
        int grow;
        bool valid;
        if constexpr (T::kQuad) {
            // LDS row -> wave column, sub-tile s (output group s / 4, slab s % 4), row r:
            // B row 4 * n + slab (the reference ffn.0 channel order, layers.py:114-125)
            const int w_n = row / T::TN;
            const int s = (row % T::TN) / 32;
            const int n = nt * T::BNO + w_n * T::TNO + (s >> 2) * 32 + (row & 31);
            valid = n < g.N;
            grow = 4 * n + (s & 3);
        } else {
            grow = nt * T::BN + row;
            valid = grow < g.N;
        }
        if constexpr (T::kFp8)
            b_off[q] = valid ? (uint32_t)((size_t)grow * g.K + kc * 16) : kOOB;
        else
            b_off[q] = valid ? (uint32_t)(((size_t)grow * g.K + kc * 8) * 2) : kOOB;
    }

    // issue the DMAs of K tile kt into ring stage st; tiles past the end read
    // zeros into a slot that is never read (uniform instruction counts keep
    // the vmcnt bookkeeping static)
    const int KT = g.K / T::kKT;
    auto dma_tile = [&](int kt, int st) {
        const uint32_t past = (kt < KT) ? 0u : kOOB;
        const int kbase = (kt < KT) ? kt * T::kKT : 0;
        char* sb = smem + st * T::kStageBytes + (BM * T::kAH + wave * T::BRW) * RB;
#pragma unroll
        for (int h = 0; h < T::kAH; ++h) {
            const int k0 = kbase + h * kBK;
            char* sa = smem + st * T::kStageBytes + (h * BM + wave * T::ARW) * RB;
            if constexpr (AMODE == A_PLAIN) {
#pragma unroll
                for (int q = 0; q < T::AI; ++q)
                    dma16(rA, sa + q * T::kRPI * RB, (a_off[q] + (uint32_t)k0 * (T::kA8 ? 1 : 2)) | past);
            } else {
                // cin % 64 == 0, so one K tile never straddles two taps
                const int tap = k0 / g.cin;
                const int c0 = k0 - tap * g.cin;
                int dy, dx;
                if constexpr (AMODE == A_S2K2) {
                    dy = tap >> 1;
                    dx = tap & 1;
                } else {
                    dy = tap / 3;
                    dx = tap - dy * 3;
                }
#pragma unroll
                for (int q = 0; q < T::AI; ++q) {
                    const int hi = a_h[q] + dy, wi = a_w[q] + dx;
                    const bool ok = (unsigned)hi < (unsigned)g.inH && (unsigned)wi < (unsigned)g.inW;
                    const uint32_t off =
                        ok ? (uint32_t)(((size_t)(a_pix[q] + hi * g.inW + wi) * g.lda + c0 + a_kc[q] * 8) * 2) : kOOB;
                    dma16(rA, sa + q * T::kRPI * RB, off | past);
                }
            }
        }
        const uint32_t kb = T::kFp8 ? (uint32_t)kbase : (uint32_t)kbase * 2;
#pragma unroll
        for (int q = 0; q < T::BI; ++q) dma16(rB, sb + q * T::kRPI * RB, (b_off[q] + kb) | past);
    };

    f32x16 acc[SM][SN];
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
        for (int j = 0; j < SN; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

    // fragment addresses: lane reads row lr, K chunk 2 * ks + (lane >> 5)
    const int lr = lane & 31;
    // (fragment rows sit at multiples of 32 in their operand block, so the
    // swizzle of the row is that of lr)
    const int fa_row = (wm * T::TM + lr) * RB;
    const int fb_row = (BM * T::kAH + wn * T::TN + lr) * RB;
    int pos[T::kSteps];
#pragma unroll
    for (int s = 0; s < T::kSteps; ++s) pos[s] = ((2 * s + (lane >> 5)) ^ swz_rows<RB>(lr)) * 16;
    // two fragment sets, alternating k16 steps
    f16x8 fra[2][SM], frb[2][SN];
    auto read_frags = [&](int st, int pos, f16x8* fa, f16x8* fb) {
        const char* base = smem + st * T::kStageBytes + pos;
#pragma unroll
        for (int i = 0; i < SM; ++i)
            fa[i] = *reinterpret_cast<const f16x8*>(base + fa_row + i * 32 * RB);
#pragma unroll
        for (int j = 0; j < SN; ++j)
            fb[j] = *reinterpret_cast<const f16x8*>(base + fb_row + j * 32 * RB);
    };
    auto mfma_all = [&](const f16x8* fa, const f16x8* fb) {
        if constexpr (T::PREC == PREC_BF16) {
            bf16x8 ca[SM];
#pragma unroll
            for (int i = 0; i < SM; ++i) ca[i] = to_bf16x8(fa[i]);
#pragma unroll
            for (int i = 0; i < SM; ++i)
#pragma unroll
                for (int j = 0; j < SN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ca[i], __builtin_bit_cast(bf16x8, fb[j]),
                                                                        acc[i][j], 0, 0, 0);
        } else {
#pragma unroll
            for (int i = 0; i < SM; ++i)
#pragma unroll
                for (int j = 0; j < SN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
    };
    // fp8: one k64 step per ring tile.  A: the lane's 32 k of its row (half
    // tile lane >> 5, all four chunks), converted; B: the matching 32 fp8
    // bytes (chunks 2h, 2h + 1 of the 64-byte row).  Any k order works as
    // long as A and B use the same one.
    const int a8_row = ((lane >> 5) * BM + wm * T::TM + lr) * RB;
    // e4m3 row: the lane's 32 bytes are chunks 2h, 2h + 1 (h = lane >> 5)
    auto read_row8 = [&](const char* rowp) {
        const int c0 = 2 * (lane >> 5);
        const i32x4 lo = *reinterpret_cast<const i32x4*>(rowp + ((c0 ^ swz(lr)) * 16));
        const i32x4 hi = *reinterpret_cast<const i32x4*>(rowp + (((c0 + 1) ^ swz(lr)) * 16));
        return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    };
    auto read_fp8 = [&](int st, i32x8* fa, i32x8* fb) {
        const char* base = smem + st * T::kStageBytes;
#pragma unroll
        for (int i = 0; i < SM; ++i) {
            if constexpr (T::kA8) {
                fa[i] = read_row8(base + fa_row + i * 32 * RB);
            } else {
                f16x8 v[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    v[c] = *reinterpret_cast<const f16x8*>(base + a8_row + i * 32 * RB + ((c ^ swz(lr)) * 16));
                fa[i] = to_fp8x32(v);
            }
        }
#pragma unroll
        for (int j = 0; j < SN; ++j) fb[j] = read_row8(base + fb_row + j * 32 * RB);
    };
    auto mfma_fp8 = [&](const i32x8* fa, const i32x8* fb) {
#pragma unroll
        for (int i = 0; i < SM; ++i)
#pragma unroll
            for (int j = 0; j < SN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa[i], fb[j], acc[i][j], 0, 0, 0, 127,
                                                                            0, 127);
    };

    // K loop.  Tile kt lives in ring slot kt % kStages; at iteration kt the
    // DMAs of tiles kt + 1 .. kt + kAhead are in flight (see Tile).
    constexpr int NDS = SM + SN, NMF = SM * SN;

    // Epilogue operands that do not depend on the product (bias, shortcut
    // tile) are loaded before the first DMA, so their latency hides under
    // the K loop.  Epilogue chunk (pass i, t): lane finishes channels
    // n .. n + 7 of pixel m.
    constexpr int NG = T::kQuad ? SN / 4 : SN;
    constexpr int EC = T::kEChunks / NT;
    auto chunk_mn = [&](int i, int t, int& rl, int& cc, int& m, int& n) {
        const int c = tid + t * NT;
        rl = c / (T::BNO / 8);
        cc = (c - rl * (T::BNO / 8)) * 8;
        m = m0 + (rl >> 5) * T::TM + i * 32 + (rl & 31);
        n = nt * T::BNO + cc;
    };
    float bias_r[SN], wsc_r[T::kFp8 ? SN : 1];
#pragma unroll
    for (int s = 0; s < SN; ++s) {
        int row;
        bool ok;
        if constexpr (T::kQuad) {
            const int n = nt * T::BNO + wn * T::TNO + (s >> 2) * 32 + lr;
            row = 4 * n + (s & 3);
            ok = n < g.N;
        } else {
            row = nt * BN + wn * T::TN + s * 32 + lr;
            ok = row < g.N;
        }
        bias_r[s] = ok ? g.bias[row] : 0.f;
        if constexpr (T::kFp8) wsc_r[s] = ok ? g.wscale[row] : 0.f;
    }
    // shortcut passes held in registers through the K loop: all of them
    // when they fit in 64 VGPRs, else only the first; each epilogue pass then
    // loads the pass kS1Early ahead of it
    constexpr bool kS1 = (EPI == EPI_SHORTCUT || EPI == EPI_SHORTCUT2 || EPI == EPI_SHORTCUT_QUANT);
    constexpr int kS1Early = (T::kFp8 && SM * SN >= 16) ? 0 : (SM * EC * 4 <= 64 && T::kWavesPerSimd == 1) ? SM : 1;
    f16x8 s1r[kS1 ? SM : 1][kS1 ? EC : 1];
    auto load_s1 = [&](int i0, int i1) {
#pragma unroll
        for (int i = i0; i < i1; ++i)
#pragma unroll
            for (int t = 0; t < EC; ++t) {
                int rl, cc, m, n;
                chunk_mn(i, t, rl, cc, m, n);
                const f16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
                s1r[i][t] = (m < g.M && n < g.N) ? *reinterpret_cast<const f16x8*>(g.S1 + (size_t)m * g.lds1 + n)
                                                  : zero8;
            }
    };
    if constexpr (kS1 && kS1Early > 0) load_s1(0, kS1Early);

#pragma unroll
    for (int st = 0; st < T::kAhead; ++st) dma_tile(st, st);
    if constexpr (T::kFp8) {
        i32x8 pa[SM], pb[SN];
        wait_tile_and_barrier<(T::kAhead - 1) * T::kDma, T::kWaitLds>();
        dma_tile(T::kAhead, T::kAhead);
        read_fp8(0, pa, pb);
        for (int kt = 1; kt < KT; ++kt) {
            wait_tile_and_barrier<(T::kAhead - 1) * T::kDma, T::kWaitLds>();
            dma_tile(kt + T::kAhead, (kt + T::kAhead) % T::kStages);
            i32x8 na[SM], nb[SN];
            read_fp8(kt % T::kStages, na, nb);
            mfma_fp8(pa, pb);  // tile kt - 1
#pragma unroll
            for (int i = 0; i < SM; ++i) pa[i] = na[i];
#pragma unroll
            for (int j = 0; j < SN; ++j) pb[j] = nb[j];
        }
        mfma_fp8(pa, pb);
    } else {
        // k16 step s of a tile reads into fragment set s & 1 and runs its
        // MFMAs one step later, so the fragment reads of a step overlap the
        // MFMAs of the one before; the last step of tile kt - 1 runs after
        // the barrier of tile kt.  The tile's DMAs are spread over its steps.
        constexpr int NS = T::kSteps;
        static_assert(NS % 2 == 0, "the last step of a tile lands in set 1");
        wait_tile_and_barrier<(T::kAhead - 1) * T::kDma, T::kWaitLds>();
        dma_tile(T::kAhead, T::kAhead % T::kStages);
        read_frags(0, pos[0], fra[0], frb[0]);
#pragma unroll
        for (int s = 1; s < NS; ++s) {
            read_frags(0, pos[s], fra[s & 1], frb[s & 1]);
            mfma_all(fra[(s - 1) & 1], frb[(s - 1) & 1]);
        }
        for (int kt = 1; kt < KT; ++kt) {
            // this wave's DMAs of tile kt have landed; after the barrier, every wave's
            wait_tile_and_barrier<(T::kAhead - 1) * T::kDma, T::kWaitLds>();
            const int st = kt % T::kStages;
            dma_tile(kt + T::kAhead, (kt + T::kAhead) % T::kStages);
            static_for(std::make_integer_sequence<int, NS>{}, [&](auto sc) {
                constexpr int s = decltype(sc)::value;
                read_frags(st, pos[s], fra[s & 1], frb[s & 1]);
                mfma_all(fra[(s + 1) & 1], frb[(s + 1) & 1]);  // step s - 1 (of tile kt - 1 for s = 0)
                // 32-deep tiles issue all their DMAs in the first step
                interleave<NDS, NMF, NS == 2 ? (s == 0 ? T::kDma : 0)
                                             : (T::kDma * (s + 1)) / NS - (T::kDma * s) / NS>();
            });
        }
        mfma_all(fra[1], frb[1]);
    }
    __syncthreads();  // every wave is done with the ring before the epilogue reuses it

    // ---- epilogue ----
    // bias (and activation) in registers; accumulator lane layout of a 32x32
    // tile: column lane & 31, row (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3).
    float* E = reinterpret_cast<float*>(smem);
    constexpr int EP = T::kEPitch;
    const int hwo = g.outH * g.outW;
    // per-QP row of pixel m: one QP index per image of outH x outW pixels
    auto qrow = [&](int m) -> const f16* {
        const int img = (hwo < g.M) ? m / hwo : 0;
        return g.q + (size_t)(g.qp ? g.qp[img] : 0) * g.qstride;
    };
#pragma unroll
    for (int i = 0; i < SM; ++i) {
        if (i > 0) __syncthreads();  // previous pass fully read back
        if constexpr (kS1 && kS1Early < SM)
            if (i + kS1Early < SM) load_s1(i + kS1Early, i + kS1Early + 1);
        // -- stage rows wm * 32 + r of pass i (tile rows wm * TM + i * 32 + r) --
#pragma unroll
        for (int gI = 0; gI < NG; ++gI) {
            const int col = wn * T::TNO + gI * 32 + lr;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = wm * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                float x;
                if constexpr (T::kQuad) {
                    x = 0.f;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        float a = acc[i][4 * gI + j][v];
                        if constexpr (T::kFp8) a *= wsc_r[4 * gI + j];
                        x += wsilu(a + bias_r[4 * gI + j]);
                    }
                } else {
                    float a = acc[i][gI][v];
                    if constexpr (T::kFp8) a *= wsc_r[gI];
                    x = a + bias_r[gI];
                    if constexpr (EPI == EPI_WSILU) x = wsilu(x);
                }
                E[row * EP + col] = x;
            }
        }
        __syncthreads();
        // -- finish 8 consecutive channels of one pixel per lane --
#pragma unroll
        for (int t = 0; t < EC; ++t) {
            int rl, cc, m, n;
            chunk_mn(i, t, rl, cc, m, n);
            if (m >= g.M || n >= g.N) continue;
            const f32x4 lo = *reinterpret_cast<const f32x4*>(E + rl * EP + cc);
            const f32x4 hi = *reinterpret_cast<const f32x4*>(E + rl * EP + cc + 4);
            float x[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
            if constexpr (kS1) {
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] += (float)s1r[i][t][e];
            }
            if constexpr (EPI == EPI_SHORTCUT2) {
                const f16x8 s = *reinterpret_cast<const f16x8*>(g.S2 + (size_t)m * g.lds2 + n);
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] += (float)s[e];
            }
            if constexpr (EPI == EPI_SHORTCUT_QUANT || EPI == EPI_QUANT) {
                if (g.q_round) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) x[e] = (float)(f16)x[e];
                }
                const f16* q = qrow(m);
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] *= (float)q[n + e];
            }
            f16x8 o;
#pragma unroll
            for (int e = 0; e < 8; ++e) o[e] = (f16)x[e];
            size_t off;
            if constexpr (T::DMODE == D_SUBPEL) {
                // n = sub * Cout + c ; pixel (2h + sub / 2, 2w + sub % 2) of the 2x grid
                const int cout = g.N >> 2;
                const int sub = n / cout;
                const int ch = n - sub * cout;
                const int img = m / hwo;
                const int r = m - img * hwo;
                const int h = r / g.outW;
                const int w = r - h * g.outW;
                off = ((size_t)(img * 2 * g.outH + 2 * h + (sub >> 1)) * (2 * g.outW) + 2 * w + (sub & 1)) * g.ldd + ch;
            } else {
                off = (size_t)m * g.ldd + n;
            }
            if constexpr (T::kFp8) {
                // fp8 products may also (or only) store an e4m3 copy
                if (g.D8) *reinterpret_cast<int2*>(g.D8 + (size_t)m * g.ldd8 + n) = fp8x8_from_f16x8(o);
                if (g.D) *reinterpret_cast<f16x8*>(g.D + off) = o;
            } else {
                *reinterpret_cast<f16x8*>(g.D + off) = o;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// configurations and dispatch
// ---------------------------------------------------------------------------

// Tile configurations (ids for dcvc_gemm_set_tile_policy):
//   1: 256 x 256, wave 128 x 128, 5 stages (1 block / CU)
//   2: 256 x 128, wave 128 x 64, 5 stages (1 block / CU)
//   3: 128 x 128, wave 64 x 64, 5 stages (2 blocks / CU)
//   4: 256 x 128, wave 128 x 64, 3 stages (2 blocks / CU)
//   5: 128 x 256, wave 64 x 128, 3 stages (2 blocks / CU)
//   6: 256 x 256, 8 waves (4 x 2, wave 64 x 128), 4 stages (2 waves / SIMD)
//   7: 256 x 256, 8 waves, 3 stages
//   quad (B rows; 4 B rows per output): 1: 256 x 256 5 st, 2: 128 x 256 5 st,
//   3: 128 x 128 waves 4 x 1, 4: 128 x 256 3 st, 5: 256 x 256 4 st,
//   6 / 7: 256 x 256 8 waves, 4 / 3 st
//   64-deep K tiles (128-byte LDS rows; plain 1x1 products in fp16 / bf16,
//   other launches fall back to the automatic choice):
//   8: 256 x 256, wave 128 x 128, 2 stages (1 block / CU)
//   9: 256 x 256, 8 waves (4 x 2, wave 64 x 128), 2 stages
//   10: 256 x 128 (quad: 128 x 256), 3 stages (1 block / CU)
//   11: 128 x 128 (quad: waves 4 x 1), 2 stages (2 blocks / CU)
//   Tiles for small grids (the 1/16- and 1/64-resolution layers, whose
//   128 x 128 grids leave CUs idle and whose K loops wait on load latency):
//   12: 128 x 128 (quad: 128 x 256), 8 waves (4 x 2, wave 32 x 64 / 32 x 128),
//       5 stages
//   13: 64 x 128 (quad: 64 x 256), waves 2 x 2 (wave 32 x 64 / 32 x 128),
//       5 stages
//   14: as 12 with 64-deep K tiles, 4 stages (quad: 3); plain 1x1 products
//       only, other launches take 12
constexpr int kPolicies = 14;

template <int BM, int BN, int WM, int WN, int ST, int EPI, int AMODE, int DMODE, int PREC = PREC_F16, int KB = 32>
static hipError_t launch_tile(GemmArgs g, hipStream_t stream) {
    using T = Tile<BM, BN, WM, WN, ST, EPI, AMODE, DMODE, PREC, KB>;
    const int tiles_m = ceil_div(g.M, BM);
    const int tiles_n = ceil_div(g.N, T::BNO);
    auto kern = gemm_kernel<T>;
    // once per instantiation; C++ static initialisation is thread-safe, so
    // concurrent first launches from several host threads are fine
    static const hipError_t attr = hipFuncSetAttribute((const void*)kern,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)T::kLdsBytes);
    if (attr != hipSuccess) return attr;
    hipLaunchKernelGGL(kern, dim3(tiles_m * tiles_n, g.batch), dim3(T::kThreads), T::kLdsBytes, stream, g, tiles_n);
    return hipGetLastError();
}

// Automatic choice, from a sweep of the 1080p HT-S layer shapes on MI355X:
// large-M layers run 2 blocks per CU (one block's epilogue overlaps the
// other's K loop), except long-K layers (K >= 1024), which keep the 256 x 256
// tile; smaller grids take 128 x 128 tiles with 8 waves (12), and grids of
// fewer than 256 such tiles (the 1/16- and 1/64-resolution layers) take
// 64 x 128 tiles (13), so that the grid still covers the 256 CUs.  Over the
// autotuner's sweep of every model and resolution, 13 takes 0.89 of the time
// of the 4-wave 128 x 128 tile on those small grids and 12 takes 0.97 on the
// others; the quad GEMM keeps 128 x 128 (4 x 1 waves) there.  The 8-wave
// 256 x 256 tiles (6, 7) are 2-10 % faster on the 1/8-resolution
// 512-channel and long-K shapes in isolation, but inside the codecs, where
// two streams run concurrently, one block per CU cannot share a CU with the
// other stream's blocks and the layers took 5-17 % longer, so they are only
// used when forced.
static int heuristic_tile(const GemmArgs& g, bool quad) {
    const long long m256 = ceil_div(g.M, 256), m128 = ceil_div(g.M, 128);
    if (quad) return (m128 * ceil_div(g.N, 64) >= 480) ? 4 : 3;
    if (g.K >= 1024 && m256 * ceil_div(g.N, 256) >= 240) return 1;
    if (m256 * ceil_div(g.N, 128) >= 480) return 4;
    return (m128 * ceil_div(g.N, 128) < 256) ? 13 : 12;
}

// Forced policy, else the tuning table's entry for this shape, else the
// heuristic.  Every configuration gives the same bits.
static int pick_tile(const GemmArgs& g, int epi, int amode, int dmode) {
    if (g_tile_policy >= 1 && g_tile_policy <= kPolicies) return g_tile_policy;
    uint64_t key;
    if (tune_key(g.prec, epi, amode, dmode, g.M, g.N, g.K, &key)) {
        std::shared_lock<std::shared_mutex> lk(g_tune_mu);
        const auto it = g_tune.find(key);
        if (it != g_tune.end()) return it->second;
    }
    return heuristic_tile(g, epi == EPI_QUAD);
}

template <int EPI, int AMODE, int DMODE, int PREC = PREC_F16>
static hipError_t dispatch_tiles(const GemmArgs& g, hipStream_t stream) {
    // 64-deep K tiles (configurations 8-11) serve plain 1x1 products only
    constexpr bool kPlain = (AMODE == A_PLAIN && DMODE == D_PLAIN);
    const int t = pick_tile(g, EPI, AMODE, DMODE);
    if constexpr (PREC == PREC_FP8 || PREC == PREC_FP8A) {
        // fp8 K tiles are 48 KB (256 x 256) / 24 KB (128 x 128): 3 stages.  The
        // shortcut epilogues (HBM-bound) stay on 128 x 128, whose registers
        // also hold the shortcut tile.
        constexpr bool kShort = (EPI == EPI_SHORTCUT || EPI == EPI_SHORTCUT2 || EPI == EPI_SHORTCUT_QUANT);
        // measured on the 1080p shapes: 256 x 256 pays off only for K >= 1024
        const bool big = !kShort && (g_tile_policy != 0 ? (t == 1 || t == 2) : g.K >= 1024);
        // grids of fewer than 256 128 x 128 tiles take 64 x 128 tiles, as
        // configuration 13 does for fp16
        const bool small = g_tile_policy != 0 ? t == 13 : ceil_div(g.M, 128) * ceil_div(g.N, 128) < 256;
        if constexpr (EPI == EPI_QUAD) {
            if (big) return launch_tile<256, 256, 2, 2, 3, EPI, AMODE, DMODE, PREC>(g, stream);
            return launch_tile<128, 128, 4, 1, 3, EPI, AMODE, DMODE, PREC>(g, stream);
        } else if constexpr (kShort) {
            if (small) return launch_tile<64, 128, 2, 2, 3, EPI, AMODE, DMODE, PREC>(g, stream);
            return launch_tile<128, 128, 2, 2, 3, EPI, AMODE, DMODE, PREC>(g, stream);
        } else {
            if (big) return launch_tile<256, 256, 2, 2, 3, EPI, AMODE, DMODE, PREC>(g, stream);
            if (small) return launch_tile<64, 128, 2, 2, 3, EPI, AMODE, DMODE, PREC>(g, stream);
            return launch_tile<128, 128, 2, 2, 3, EPI, AMODE, DMODE, PREC>(g, stream);
        }
    } else if constexpr (EPI == EPI_QUAD) {
        switch (t) {
        case 1: return launch_tile<256, 256, 2, 2, 5, EPI, AMODE, DMODE, PREC>(g, stream);
        case 2: return launch_tile<128, 256, 2, 2, 5, EPI, AMODE, DMODE, PREC>(g, stream);
        case 4: return launch_tile<128, 256, 2, 2, 3, EPI, AMODE, DMODE, PREC>(g, stream);
        case 5: return launch_tile<256, 256, 2, 2, 4, EPI, AMODE, DMODE, PREC>(g, stream);
        case 6: return launch_tile<256, 256, 4, 2, 4, EPI, AMODE, DMODE, PREC>(g, stream);
        case 7: return launch_tile<256, 256, 4, 2, 3, EPI, AMODE, DMODE, PREC>(g, stream);
        case 14: if constexpr (kPlain) return launch_tile<128, 256, 4, 2, 3, EPI, AMODE, DMODE, PREC, 64>(g, stream);
        [[fallthrough]];
        case 12: return launch_tile<128, 256, 4, 2, 5, EPI, AMODE, DMODE, PREC>(g, stream);
        case 13: return launch_tile<64, 256, 2, 2, 5, EPI, AMODE, DMODE, PREC>(g, stream);
        case 8: if constexpr (kPlain) return launch_tile<256, 256, 2, 2, 2, EPI, AMODE, DMODE, PREC, 64>(g, stream);
        [[fallthrough]];
        case 9: if constexpr (kPlain) return launch_tile<256, 256, 4, 2, 2, EPI, AMODE, DMODE, PREC, 64>(g, stream);
        [[fallthrough]];
        case 10: if constexpr (kPlain) return launch_tile<128, 256, 2, 2, 3, EPI, AMODE, DMODE, PREC, 64>(g, stream);
        [[fallthrough]];
        case 11: if constexpr (kPlain) return launch_tile<128, 128, 4, 1, 2, EPI, AMODE, DMODE, PREC, 64>(g, stream);
        [[fallthrough]];
        default: return launch_tile<128, 128, 4, 1, 5, EPI, AMODE, DMODE, PREC>(g, stream);
        }
    } else {
        switch (t) {
        case 1: return launch_tile<256, 256, 2, 2, 5, EPI, AMODE, DMODE, PREC>(g, stream);
        case 2: return launch_tile<256, 128, 2, 2, 5, EPI, AMODE, DMODE, PREC>(g, stream);
        case 4: return launch_tile<256, 128, 2, 2, 3, EPI, AMODE, DMODE, PREC>(g, stream);
        case 5: return launch_tile<128, 256, 2, 2, 3, EPI, AMODE, DMODE, PREC>(g, stream);
        case 6: return launch_tile<256, 256, 4, 2, 4, EPI, AMODE, DMODE, PREC>(g, stream);
        case 7: return launch_tile<256, 256, 4, 2, 3, EPI, AMODE, DMODE, PREC>(g, stream);
        case 14: if constexpr (kPlain) return launch_tile<128, 128, 4, 2, 4, EPI, AMODE, DMODE, PREC, 64>(g, stream);
        [[fallthrough]];
        case 12: return launch_tile<128, 128, 4, 2, 5, EPI, AMODE, DMODE, PREC>(g, stream);
        case 13: return launch_tile<64, 128, 2, 2, 5, EPI, AMODE, DMODE, PREC>(g, stream);
        case 8: if constexpr (kPlain) return launch_tile<256, 256, 2, 2, 2, EPI, AMODE, DMODE, PREC, 64>(g, stream);
        [[fallthrough]];
        case 9: if constexpr (kPlain) return launch_tile<256, 256, 4, 2, 2, EPI, AMODE, DMODE, PREC, 64>(g, stream);
        [[fallthrough]];
        case 10: if constexpr (kPlain) return launch_tile<256, 128, 2, 2, 3, EPI, AMODE, DMODE, PREC, 64>(g, stream);
        [[fallthrough]];
        case 11: if constexpr (kPlain) return launch_tile<128, 128, 2, 2, 2, EPI, AMODE, DMODE, PREC, 64>(g, stream);
        [[fallthrough]];
        default: return launch_tile<128, 128, 2, 2, 5, EPI, AMODE, DMODE, PREC>(g, stream);
        }
    }
}

template <int PREC>
static hipError_t dispatch_plain(const GemmArgs& g, int epi, hipStream_t stream) {
    switch (epi) {
    case EPI_BIAS: return dispatch_tiles<EPI_BIAS, A_PLAIN, D_PLAIN, PREC>(g, stream);
    case EPI_WSILU: return dispatch_tiles<EPI_WSILU, A_PLAIN, D_PLAIN, PREC>(g, stream);
    case EPI_SHORTCUT: return dispatch_tiles<EPI_SHORTCUT, A_PLAIN, D_PLAIN, PREC>(g, stream);
    case EPI_SHORTCUT2: return dispatch_tiles<EPI_SHORTCUT2, A_PLAIN, D_PLAIN, PREC>(g, stream);
    case EPI_SHORTCUT_QUANT: return dispatch_tiles<EPI_SHORTCUT_QUANT, A_PLAIN, D_PLAIN, PREC>(g, stream);
    case EPI_QUANT: return dispatch_tiles<EPI_QUANT, A_PLAIN, D_PLAIN, PREC>(g, stream);
    case EPI_QUAD: return dispatch_tiles<EPI_QUAD, A_PLAIN, D_PLAIN, PREC>(g, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t gemm_launch(const GemmArgs& g_in, int epi, int amode, int dmode, hipStream_t stream) {
    GemmArgs g = g_in;
    // byte extents of the operands for the buffer range checks
    size_t a_rows;
    int a_last;
    if (amode == A_PLAIN) {
        a_rows = (size_t)g.M;
        a_last = g.K;
    } else {
        a_rows = (size_t)(g.M / (g.outH * g.outW)) * g.inH * g.inW;
        a_last = g.cin;
    }
    const bool w8 = (g.prec == PREC_FP8 || g.prec == PREC_FP8A);
    g.a_bytes = (uint32_t)(((a_rows - 1) * g.lda + a_last) * (g.prec == PREC_FP8A ? 1 : 2));
    g.b_bytes = (uint32_t)((size_t)(epi == EPI_QUAD ? 4 * g.N : g.N) * g.K * (w8 ? 1 : 2));
    // implicit-conv A modes and the sub-pixel store come with the plain bias
    // epilogue only (conv_bias, transposed_conv)
    if (amode != A_PLAIN || dmode != D_PLAIN) {
        if (epi != EPI_BIAS) return hipErrorInvalidValue;
        if (dmode == D_SUBPEL) {
            if (amode == A_PLAIN) return dispatch_tiles<EPI_BIAS, A_PLAIN, D_SUBPEL>(g, stream);
            if (amode == A_S1K3) return dispatch_tiles<EPI_BIAS, A_S1K3, D_SUBPEL>(g, stream);
            return hipErrorInvalidValue;
        }
        switch (amode) {
        case A_S2K2: return dispatch_tiles<EPI_BIAS, A_S2K2, D_PLAIN>(g, stream);
        case A_S2K3: return dispatch_tiles<EPI_BIAS, A_S2K3, D_PLAIN>(g, stream);
        case A_S1K3: return dispatch_tiles<EPI_BIAS, A_S1K3, D_PLAIN>(g, stream);
        default: return hipErrorInvalidValue;
        }
    }
    switch (g.prec) {
    case PREC_F16: return dispatch_plain<PREC_F16>(g, epi, stream);
    case PREC_BF16: return dispatch_plain<PREC_BF16>(g, epi, stream);
    case PREC_FP8: return dispatch_plain<PREC_FP8>(g, epi, stream);
    case PREC_FP8A: return dispatch_plain<PREC_FP8A>(g, epi, stream);
    default: return hipErrorInvalidValue;
    }
}

static bool aligned16(const void* p) { return ((uintptr_t)p) % 16 == 0; }

// Host-side validation shared by every GEMM entry point.
int gemm_validate(const GemmArgs& g, int epi, int amode, int dmode) {
    if (!g.A || !g.B || !g.bias || (!g.D && !g.D8)) return DCVC_ERR_ARG;
    if ((amode != A_PLAIN || dmode != D_PLAIN) && (epi != EPI_BIAS || g.prec != PREC_F16)) return DCVC_ERR_ARG;
    const bool w8 = (g.prec == PREC_FP8 || g.prec == PREC_FP8A);
    if (g.prec < PREC_F16 || g.prec > PREC_FP8A || (w8 && !g.wscale)) return DCVC_ERR_ARG;
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return DCVC_ERR_ARG;
    if (g.K % 64 != 0 || g.N % 8 != 0) return DCVC_ERR_ARG;
    // e4m3 A rows (lda in bytes) keep the 16-byte DMA alignment
    if (g.lda % (g.prec == PREC_FP8A ? 16 : 8) != 0 || g.lda < (amode == A_PLAIN ? g.K : g.cin)) return DCVC_ERR_ARG;
    if (!aligned16(g.A) || !aligned16(g.B)) return DCVC_ERR_ARG;
    if (g.D && (!aligned16(g.D) || g.ldd % 8 != 0)) return DCVC_ERR_ARG;
    if (g.D8 && (!w8 || dmode != D_PLAIN || g.batch != 1 || g.ldd8 < g.N || g.ldd8 % 8 != 0 ||
                 ((uintptr_t)g.D8) % 8 != 0))
        return DCVC_ERR_ARG;
    size_t a_rows = (size_t)g.M;
    if (amode != A_PLAIN) {
        const int taps = (amode == A_S2K2) ? 4 : 9;
        if (g.cin <= 0 || g.cin % 64 != 0 || g.cin * taps != g.K) return DCVC_ERR_ARG;
        if (g.outH <= 0 || g.outW <= 0 || g.inH <= 0 || g.inW <= 0) return DCVC_ERR_ARG;
        if (g.M % (g.outH * g.outW) != 0) return DCVC_ERR_ARG;
        if (amode == A_S2K2 && (g.inH != 2 * g.outH || g.inW != 2 * g.outW)) return DCVC_ERR_ARG;
        if (amode == A_S2K3 && (g.outH != (g.inH + 1) / 2 || g.outW != (g.inW + 1) / 2)) return DCVC_ERR_ARG;
        if (amode == A_S1K3 && (g.outH != g.inH || g.outW != g.inW)) return DCVC_ERR_ARG;
        a_rows = (size_t)(g.M / (g.outH * g.outW)) * g.inH * g.inW;
    }
    // buffer addressing uses 32-bit byte offsets below 2^31
    if (a_rows * (size_t)g.lda * (g.prec == PREC_FP8A ? 1 : 2) >= kOOB) return DCVC_ERR_ARG;
    if ((size_t)(epi == EPI_QUAD ? 4 * g.N : g.N) * g.K * 2 >= kOOB) return DCVC_ERR_ARG;
    if (dmode == D_SUBPEL) {
        if (amode != A_PLAIN && amode != A_S1K3) return DCVC_ERR_ARG;
        if (g.N % 32 != 0 || g.outH <= 0 || g.outW <= 0 || g.M % (g.outH * g.outW) != 0) return DCVC_ERR_ARG;
        if (g.ldd < g.N / 4) return DCVC_ERR_ARG;
    } else if (g.D && g.ldd < g.N) {
        return DCVC_ERR_ARG;
    }
    if (epi == EPI_SHORTCUT || epi == EPI_SHORTCUT2 || epi == EPI_SHORTCUT_QUANT) {
        if (!g.S1 || g.lds1 < g.N || g.lds1 % 8 != 0 || !aligned16(g.S1)) return DCVC_ERR_ARG;
    }
    if (epi == EPI_SHORTCUT2 && (!g.S2 || g.lds2 < g.N || g.lds2 % 8 != 0 || !aligned16(g.S2))) return DCVC_ERR_ARG;
    if ((epi == EPI_SHORTCUT_QUANT || epi == EPI_QUANT) && !g.q) return DCVC_ERR_ARG;
    if (g.batch != 1) {
        // every problem's operands keep the 16-byte alignment of the first
        if (g.batch < 1 || g.batch > 65535 || g.a_div < 1 || amode != A_PLAIN || dmode != D_PLAIN) return DCVC_ERR_ARG;
        if (g.prec == PREC_FP8A || !g.D) return DCVC_ERR_ARG;
        if (g.sA < 0 || g.sB < 0 || g.sBias < 0 || g.sS1 < 0 || g.sS2 < 0 || g.sD < 0) return DCVC_ERR_ARG;
        const long wb = w8 ? 1 : 2;
        if (g.sA % 8 || g.sD % 8 || g.sS1 % 8 || g.sS2 % 8 || (g.sB * wb) % 16) return DCVC_ERR_ARG;
    }
    return DCVC_OK;
}

// ---------------------------------------------------------------------------
// weight conversion
// ---------------------------------------------------------------------------

__global__ void f16_to_bf16_kernel(const f16* __restrict__ w, long n, __bf16* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (__bf16)(float)w[i];
}

// one block per row: max |w|, then q = w / scale as e4m3 (|q| <= 448 by construction)
__global__ __launch_bounds__(256) void rows_to_fp8_kernel(const f16* __restrict__ w, int K, uint8_t* __restrict__ out,
                                                          float* __restrict__ scale) {
    __shared__ float red[256];
    const f16* row = w + (size_t)blockIdx.x * K;
    float m = 0.f;
    for (int k = threadIdx.x; k < K; k += 256) m = fmaxf(m, fabsf((float)row[k]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const float sc = red[0] > 0.f ? red[0] / 448.f : 1.f;
    if (threadIdx.x == 0) scale[blockIdx.x] = sc;
    const float inv = 1.f / sc;
    uint8_t* orow = out + (size_t)blockIdx.x * K;
    for (int k = 2 * threadIdx.x; k < K; k += 512) {
        const float a = fminf(fmaxf((float)row[k] * inv, -448.f), 448.f);
        const float b = fminf(fmaxf((float)row[k + 1] * inv, -448.f), 448.f);
        const int p = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
        orow[k] = (uint8_t)(p & 0xff);
        orow[k + 1] = (uint8_t)((p >> 8) & 0xff);
    }
}

hipError_t weights_to_bf16(const f16* w, long n, void* out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(f16_to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, w, n,
                       static_cast<__bf16*>(out));
    return hipGetLastError();
}

hipError_t weights_to_fp8(const f16* w, int rows, int K, void* out, float* scale, hipStream_t stream) {
    if (rows <= 0) return hipSuccess;
    if (K % 2 != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(rows_to_fp8_kernel, dim3(rows), dim3(256), 0, stream, w, K, static_cast<uint8_t*>(out),
                       scale);
    return hipGetLastError();
}

namespace {
std::atomic<bool> g_fp8_act{true};
}

void set_fp8_activations(bool on) { g_fp8_act.store(on); }
bool fp8_activations() { return g_fp8_act.load(); }

int gemm_set_tile_policy(int policy) {
    if (policy < 0 || policy > kPolicies) return DCVC_ERR_ARG;
    g_tile_policy = policy;
    return DCVC_OK;
}

int gemm_tune_set(const long long* entries, int n) {
    if (n < 0 || (n > 0 && !entries)) return DCVC_ERR_ARG;
    std::unordered_map<uint64_t, int> t;
    for (int i = 0; i < n; ++i) {
        const long long* e = entries + 8LL * i;
        uint64_t key;
        if (!tune_key((int)e[0], (int)e[1], (int)e[2], (int)e[3], e[4], e[5], e[6], &key)) return DCVC_ERR_ARG;
        if (e[7] < 1 || e[7] > kPolicies) return DCVC_ERR_ARG;
        t[key] = (int)e[7];
    }
    std::unique_lock<std::shared_mutex> lk(g_tune_mu);
    g_tune.swap(t);
    return DCVC_OK;
}

int gemm_tile_for(const GemmArgs& g, int epi, int amode, int dmode) { return pick_tile(g, epi, amode, dmode); }

}  // namespace dcvc
