// MFMA GEMM family for the 1x1-conv / implicit-GEMM conv layers of DCVC-UF.
//
//   D[m, n] = epilogue( sum_k A[m, k] * B[n, k] + bias[n] )
//
// A: NHWC fp16 activations (row m = one output pixel, row stride lda), or an
//    implicit im2col view of them for the strided convs.
// B: fp16 weights, row n = output channel, K contiguous (PyTorch conv layout
//    [N, K, 1, 1]; the strided convs use tap-major K = tap * Cin + c).
// Accumulation: fp32 on v_mfma_f32_32x32x16_f16.  Each output element is
// accumulated by one lane in increasing k, 16 at a time, so the result is
// independent of the tile shape, grid, and batch composition.
//
// Reference operators replaced (def_cutlass.h:10-39):
//   conv1x1_bias, conv1x1_bias_wsilu, conv1x1_bias_shortcut(2),
//   conv1x1_bias_shortcut_with_quant, conv1x1_bias_with_quant,
//   conv1x1_bias_wsilu_chunk_add (quad GEMM), conv_bias (k=2/3, s=2),
//   transposed_conv (2x2 s2).  Epilogue op order follows
//   cutlass/cutlass_epilogue.h:79-131,175-224: (acc + b) [-> wsilu] [+ S1]
//   [+ S2 | * q].
#include "common.h"
#include "gemm.h"

namespace dcvc {

template <int WM, int WN, int SM, int SN, int BK, int EPI, int AMODE, int DMODE>
struct GemmCfg {
    static constexpr int kWaves = WM * WN;
    static constexpr int kThreads = kWaves * 64;
    static constexpr int kBM = WM * SM * 32;
    static constexpr int kBNE = WN * SN * 32;                      // B rows per block
    static constexpr bool kQuad = (EPI == EPI_QUAD);
    static constexpr int kBNO = kQuad ? WN * 32 : kBNE;            // output columns per block
    static constexpr int kLdsRow = BK + 8;                         // halves, padded
    static constexpr int kChunksPerRow = BK / 8;                   // 16-byte chunks
    static constexpr int kAChunks = kBM * kChunksPerRow / kThreads;
    static constexpr int kBChunks = kBNE * kChunksPerRow / kThreads;
    static constexpr int kLdsHalves = (kBM + kBNE) * kLdsRow;     // one stage
    static_assert(!kQuad || SN == 4, "quad GEMM needs 4 B slabs per wave");
    static_assert(kBM * kChunksPerRow % kThreads == 0, "A tile split");
    static_assert(kBNE * kChunksPerRow % kThreads == 0, "B tile split");
};

// Global address of the 16-byte A chunk for tile row `row` (output pixel m)
// and K offset k (multiple of 8).  Returns nullptr for a zero (padding) chunk.
template <int AMODE>
DCVC_DEV const f16* a_chunk_ptr(const GemmArgs& g, int m, int k) {
    if (m >= g.M) return nullptr;
    if constexpr (AMODE == A_PLAIN) {
        return g.A + (size_t)m * g.lda + k;
    } else {
        const int hw = g.outH * g.outW;
        const int img = m / hw;
        const int r = m - img * hw;
        const int ho = r / g.outW;
        const int wo = r - ho * g.outW;
        const int cin = g.cin;
        const int tap = k / cin;
        const int c = k - tap * cin;
        int hi, wi;
        if constexpr (AMODE == A_S2K2) {
            hi = 2 * ho + (tap >> 1);
            wi = 2 * wo + (tap & 1);
        } else if constexpr (AMODE == A_S2K3) {
            hi = 2 * ho - 1 + tap / 3;
            wi = 2 * wo - 1 + tap % 3;
        } else {  // A_S1K3
            hi = ho - 1 + tap / 3;
            wi = wo - 1 + tap % 3;
        }
        if (hi < 0 || hi >= g.inH || wi < 0 || wi >= g.inW) return nullptr;
        return g.A + ((size_t)(img * g.inH + hi) * g.inW + wi) * g.lda + c;
    }
}

template <class Cfg, int EPI, int AMODE, int DMODE, int BK, int WM, int WN, int SM, int SN>
__global__ __launch_bounds__(Cfg::kThreads) void gemm_kernel(GemmArgs g) {
    extern __shared__ __attribute__((aligned(16))) f16 lds[];
    constexpr int BM = Cfg::kBM, BNE = Cfg::kBNE, LR = Cfg::kLdsRow, CPR = Cfg::kChunksPerRow;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM;
    const int wn = wave / WM;
    const int m0 = blockIdx.x * BM;
    const int nb = blockIdx.y;  // output-column block

    // ---- per-thread global chunk coordinates (fixed across the K loop) ----
    const f16* a_rowptr[Cfg::kAChunks];
    int a_lds_off[Cfg::kAChunks];
    int a_kk[Cfg::kAChunks];
    int a_m[Cfg::kAChunks];
#pragma unroll
    for (int i = 0; i < Cfg::kAChunks; ++i) {
        const int chunk = tid + i * Cfg::kThreads;
        const int row = chunk / CPR;
        const int kk = (chunk % CPR) * 8;
        a_m[i] = m0 + row;
        a_kk[i] = kk;
        a_lds_off[i] = row * LR + kk;
        a_rowptr[i] = nullptr;
    }
    const f16* b_rowptr[Cfg::kBChunks];
    int b_lds_off[Cfg::kBChunks];
#pragma unroll
    for (int i = 0; i < Cfg::kBChunks; ++i) {
        const int chunk = tid + i * Cfg::kThreads;
        const int row = chunk / CPR;
        const int kk = (chunk % CPR) * 8;
        int grow;
        bool valid;
        if constexpr (Cfg::kQuad) {
            // LDS row -> (wave column, slab j, r): B row 4 * n + j (reference ffn.0 order)
            const int w_n = row / (SN * 32);
            const int j = (row / 32) % SN;
            const int r = row % 32;
            const int n = nb * Cfg::kBNO + w_n * 32 + r;
            valid = n < g.N;
            grow = 4 * n + j;
        } else {
            grow = nb * BNE + row;
            valid = grow < g.N;
        }
        b_rowptr[i] = valid ? g.B + (size_t)grow * g.K + kk : nullptr;
        b_lds_off[i] = (BM + row) * LR + kk;
    }

    f16x8 ra[Cfg::kAChunks], rb[Cfg::kBChunks];
    const f16x8 zero8 = {0, 0, 0, 0, 0, 0, 0, 0};

    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < Cfg::kAChunks; ++i) {
            const f16* p;
            if constexpr (AMODE == A_PLAIN) {
                p = (a_m[i] < g.M) ? g.A + (size_t)a_m[i] * g.lda + k0 + a_kk[i] : nullptr;
            } else {
                p = a_chunk_ptr<AMODE>(g, a_m[i], k0 + a_kk[i]);
            }
            ra[i] = p ? *reinterpret_cast<const f16x8*>(p) : zero8;
        }
#pragma unroll
        for (int i = 0; i < Cfg::kBChunks; ++i)
            rb[i] = b_rowptr[i] ? *reinterpret_cast<const f16x8*>(b_rowptr[i] + k0) : zero8;
    };
    auto lstore = [&](int stage) {
        f16* base = lds + stage * Cfg::kLdsHalves;
#pragma unroll
        for (int i = 0; i < Cfg::kAChunks; ++i) *reinterpret_cast<f16x8*>(base + a_lds_off[i]) = ra[i];
#pragma unroll
        for (int i = 0; i < Cfg::kBChunks; ++i) *reinterpret_cast<f16x8*>(base + b_lds_off[i]) = rb[i];
    };
    (void)a_rowptr;

    f32x16 acc[SM][SN];
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
        for (int j = 0; j < SN; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

    const int KT = g.K / BK;
    gload(0);
    lstore(0);
    __syncthreads();
    const int lr = lane & 31;
    const int lk = (lane >> 5) * 8;
    for (int kt = 0; kt < KT; ++kt) {
        if (kt + 1 < KT) gload((kt + 1) * BK);
        const f16* base = lds + (kt & 1) * Cfg::kLdsHalves;
#pragma unroll
        for (int ks = 0; ks < BK / 16; ++ks) {
            f16x8 fa[SM], fb[SN];
#pragma unroll
            for (int i = 0; i < SM; ++i)
                fa[i] = *reinterpret_cast<const f16x8*>(base + ((wm * SM + i) * 32 + lr) * LR + ks * 16 + lk);
#pragma unroll
            for (int j = 0; j < SN; ++j)
                fb[j] = *reinterpret_cast<const f16x8*>(base + (BM + (wn * SN + j) * 32 + lr) * LR + ks * 16 + lk);
#pragma unroll
            for (int i = 0; i < SM; ++i)
#pragma unroll
                for (int j = 0; j < SN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < KT) lstore((kt + 1) & 1);
        __syncthreads();
    }

    // ---- epilogue ----
    const f16* q = nullptr;
    if constexpr (EPI == EPI_SHORTCUT_QUANT || EPI == EPI_QUANT) {
        const int qp = g.qp ? *g.qp : 0;
        q = g.q + (size_t)qp * g.qstride;
    }
    const int hwo = g.outH * g.outW;
#pragma unroll
    for (int i = 0; i < SM; ++i) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int m = m0 + (wm * SM + i) * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
            if (m >= g.M) continue;
            if constexpr (Cfg::kQuad) {
                const int n = nb * Cfg::kBNO + wn * 32 + lr;
                if (n >= g.N) continue;
                float s = 0.f;
#pragma unroll
                for (int j = 0; j < SN; ++j) s += wsilu(acc[i][j][v] + g.bias[4 * n + j]);
                g.D[(size_t)m * g.ldd + n] = (f16)s;
            } else {
#pragma unroll
                for (int j = 0; j < SN; ++j) {
                    const int n = nb * BNE + (wn * SN + j) * 32 + lr;
                    if (n >= g.N) continue;
                    float x = acc[i][j][v] + g.bias[n];
                    if constexpr (EPI == EPI_WSILU) x = wsilu(x);
                    if constexpr (EPI == EPI_SHORTCUT || EPI == EPI_SHORTCUT2 || EPI == EPI_SHORTCUT_QUANT)
                        x += (float)g.S1[(size_t)m * g.lds1 + n];
                    if constexpr (EPI == EPI_SHORTCUT2) x += (float)g.S2[(size_t)m * g.lds2 + n];
                    if constexpr (EPI == EPI_SHORTCUT_QUANT || EPI == EPI_QUANT) x *= (float)q[n];
                    size_t off;
                    if constexpr (DMODE == D_SUBPEL) {
                        // n = sub * Cout + c ; pixel (2h + dy, 2w + dx) of the 2x grid
                        const int cout = g.N >> 2;
                        const int sub = n / cout;
                        const int c = n - sub * cout;
                        const int img = m / hwo;
                        const int r = m - img * hwo;
                        const int h = r / g.outW;
                        const int w = r - h * g.outW;
                        const int H2 = 2 * g.outH, W2 = 2 * g.outW;
                        off = ((size_t)(img * H2 + 2 * h + (sub >> 1)) * W2 + 2 * w + (sub & 1)) * g.ldd + c;
                    } else {
                        off = (size_t)m * g.ldd + n;
                    }
                    g.D[off] = (f16)x;
                }
            }
        }
    }
}

template <int WM, int WN, int SM, int SN, int BK, int EPI, int AMODE, int DMODE>
static hipError_t launch_cfg(const GemmArgs& g, hipStream_t stream) {
    using Cfg = GemmCfg<WM, WN, SM, SN, BK, EPI, AMODE, DMODE>;
    const int n_out = (EPI == EPI_QUAD) ? g.N : g.N;  // output columns covered by grid.y
    dim3 grid(ceil_div(g.M, Cfg::kBM), ceil_div(n_out, Cfg::kBNO));
    const size_t lds = 2 * Cfg::kLdsHalves * sizeof(f16);
    auto kern = gemm_kernel<Cfg, EPI, AMODE, DMODE, BK, WM, WN, SM, SN>;
    static bool attr_set = false;
    if (!attr_set) {
        hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr_set = true;
    }
    hipLaunchKernelGGL(kern, grid, dim3(Cfg::kThreads), lds, stream, g);
    return hipGetLastError();
}

template <int EPI, int AMODE, int DMODE>
static hipError_t dispatch_tiles(const GemmArgs& g, hipStream_t stream) {
    if constexpr (EPI == EPI_QUAD) {
        // one 32-column output slab per wave column, 4 B slabs per wave
        return launch_cfg<4, 1, 1, 4, 64, EPI, AMODE, DMODE>(g, stream);
    } else {
        if (g.N <= 64 || g.M <= 2048)
            return launch_cfg<4, 1, 1, 2, 64, EPI, AMODE, DMODE>(g, stream);  // 128 x 64
        return launch_cfg<2, 2, 2, 2, 64, EPI, AMODE, DMODE>(g, stream);      // 128 x 128
    }
}

hipError_t gemm_launch(const GemmArgs& g, int epi, int amode, int dmode, hipStream_t stream) {
#define DCVC_EPI_CASE(E)                                                                        \
    case E:                                                                                     \
        if (dmode == D_SUBPEL) return dispatch_tiles<E, A_PLAIN, D_SUBPEL>(g, stream);          \
        switch (amode) {                                                                        \
        case A_PLAIN: return dispatch_tiles<E, A_PLAIN, D_PLAIN>(g, stream);                    \
        case A_S2K2: return dispatch_tiles<E, A_S2K2, D_PLAIN>(g, stream);                      \
        case A_S2K3: return dispatch_tiles<E, A_S2K3, D_PLAIN>(g, stream);                      \
        case A_S1K3: return dispatch_tiles<E, A_S1K3, D_PLAIN>(g, stream);                      \
        default: return hipErrorInvalidValue;                                                   \
        }
    switch (epi) {
        DCVC_EPI_CASE(EPI_BIAS)
        DCVC_EPI_CASE(EPI_WSILU)
        DCVC_EPI_CASE(EPI_SHORTCUT)
        DCVC_EPI_CASE(EPI_SHORTCUT2)
        DCVC_EPI_CASE(EPI_SHORTCUT_QUANT)
        DCVC_EPI_CASE(EPI_QUANT)
    case EPI_QUAD:
        if (amode != A_PLAIN || dmode != D_PLAIN) return hipErrorInvalidValue;
        return dispatch_tiles<EPI_QUAD, A_PLAIN, D_PLAIN>(g, stream);
    default:
        return hipErrorInvalidValue;
    }
#undef DCVC_EPI_CASE
}

// Host-side validation shared by every GEMM entry point.
int gemm_validate(const GemmArgs& g, int epi, int amode, int dmode) {
    if (!g.A || !g.B || !g.D || !g.bias) return DCVC_ERR_ARG;
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return DCVC_ERR_ARG;
    if (g.K % 64 != 0) return DCVC_ERR_ARG;
    if (g.lda % 8 != 0 || g.lda < (amode == A_PLAIN ? g.K : g.cin)) return DCVC_ERR_ARG;
    if (((uintptr_t)g.A) % 16 != 0 || ((uintptr_t)g.B) % 16 != 0) return DCVC_ERR_ARG;
    if (amode != A_PLAIN) {
        const int taps = (amode == A_S2K2) ? 4 : 9;
        if (g.cin <= 0 || g.cin % 64 != 0 || g.cin * taps != g.K) return DCVC_ERR_ARG;
        if (g.outH <= 0 || g.outW <= 0 || g.inH <= 0 || g.inW <= 0) return DCVC_ERR_ARG;
        if (g.M % (g.outH * g.outW) != 0) return DCVC_ERR_ARG;
        if (amode == A_S2K2 && (g.inH != 2 * g.outH || g.inW != 2 * g.outW)) return DCVC_ERR_ARG;
        if (amode == A_S2K3 && (g.outH != (g.inH + 1) / 2 || g.outW != (g.inW + 1) / 2)) return DCVC_ERR_ARG;
        if (amode == A_S1K3 && (g.outH != g.inH || g.outW != g.inW)) return DCVC_ERR_ARG;
    }
    if (dmode == D_SUBPEL) {
        if (g.N % 4 != 0 || g.outH <= 0 || g.outW <= 0 || g.M % (g.outH * g.outW) != 0) return DCVC_ERR_ARG;
        if (g.ldd < g.N / 4) return DCVC_ERR_ARG;
    } else if (g.ldd < g.N) {
        return DCVC_ERR_ARG;
    }
    if ((epi == EPI_SHORTCUT || epi == EPI_SHORTCUT2 || epi == EPI_SHORTCUT_QUANT) && (!g.S1 || g.lds1 < g.N))
        return DCVC_ERR_ARG;
    if (epi == EPI_SHORTCUT2 && (!g.S2 || g.lds2 < g.N)) return DCVC_ERR_ARG;
    if ((epi == EPI_SHORTCUT_QUANT || epi == EPI_QUANT) && !g.q) return DCVC_ERR_ARG;
    return DCVC_OK;
}

}  // namespace dcvc
