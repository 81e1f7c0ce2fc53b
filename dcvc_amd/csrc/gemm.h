// Internal interface of the MFMA GEMM family (see gemm.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "common.h"

namespace dcvc {

enum GemmEpi {
    EPI_BIAS = 0,            // acc + b
    EPI_WSILU = 1,           // wsilu(acc + b)
    EPI_SHORTCUT = 2,        // acc + b + S1
    EPI_SHORTCUT2 = 3,       // acc + b + S1 + S2
    EPI_SHORTCUT_QUANT = 4,  // (acc + b + S1) * q[n]
    EPI_QUANT = 5,           // (acc + b) * q[n]
    EPI_QUAD = 6,            // sum_j wsilu(acc_j + b_j), B rows 4n + j
};

enum GemmAMode {
    A_PLAIN = 0,  // A[m, k] = A + m * lda + k
    A_S2K2 = 1,   // 2x2 stride-2 conv, K = 4 * cin, tap-major
    A_S2K3 = 2,   // 3x3 stride-2 pad-1 conv, K = 9 * cin, tap-major
    A_S1K3 = 3,   // 3x3 stride-1 pad-1 conv, K = 9 * cin, tap-major
};

// Operand precision of the product.  Weights are stored in the precision;
// activations are fp16 in memory and converted as fragments leave LDS, except
// under PREC_FP8A, where a producer has already stored them as e4m3.
enum GemmPrec {
    PREC_F16 = 0,   // fp16 x fp16, v_mfma_f32_32x32x16_f16
    PREC_BF16 = 1,  // bf16 x bf16, v_mfma_f32_32x32x16_bf16
    PREC_FP8 = 2,   // OCP fp8 e4m3 x e4m3, v_mfma_scale_f32_32x32x64_f8f6f4; weights carry
                    // one fp32 scale per row (max |w| / 448), activations are taken at scale 1
    PREC_FP8A = 3,  // PREC_FP8 with A already e4m3 bytes in memory (lda in bytes), as
                    // fp8x8_from_f16x8 makes them: the same bits as PREC_FP8 on the fp16 A
};

enum GemmDMode {
    D_PLAIN = 0,   // D[m, n] = D + m * ldd + n
    D_SUBPEL = 1,  // n = sub * (N/4) + c stored at pixel (2h + sub/2, 2w + sub%2) of the 2x grid
                   // (with A_PLAIN: 1x1 sub-pixel conv; with A_S1K3: 3x3 pad-1 sub-pixel conv)
};

struct GemmArgs {
    const f16* A = nullptr;
    int lda = 0;
    int M = 0, N = 0, K = 0;
    // geometry: GEMM rows are pixels of an (outH x outW) grid, images stacked;
    // implicit-conv modes read an (inH x inW) input grid with cin channels.
    int inH = 0, inW = 0, outH = 1, outW = 1, cin = 0;
    const f16* B = nullptr;
    const float* bias = nullptr;
    const f16* S1 = nullptr;
    int lds1 = 0;
    const f16* S2 = nullptr;
    int lds2 = 0;
    const f16* q = nullptr;     // per-QP table, row stride qstride
    int qstride = 0;
    const int* qp = nullptr;    // device QP index per image of outH x outW pixels (nullptr -> row 0)
    int q_round = 0;            // quant epilogues: round to fp16 before the q multiply,
                                // as a separate fp16 multiply_with_broadcast would
    f16* D = nullptr;                // nullptr: no fp16 output (D8 only)
    int ldd = 0;
    uint8_t* D8 = nullptr;           // optional e4m3 copy of the fp16 output (fp8 products, D_PLAIN, one problem),
    int ldd8 = 0;                    // row stride in bytes: the A operand of a later PREC_FP8A product
    int prec = PREC_F16;             // GemmPrec; B is bf16 / fp8 bytes accordingly
    const float* wscale = nullptr;   // PREC_FP8: per B row dequantisation scale
    // `batch` independent problems of this shape in one launch (A_PLAIN /
    // D_PLAIN only): problem b reads A + (b / a_div) * sA (a_div > 1 shares
    // one input among consecutive problems), B at b * sB weight elements,
    // bias / wscale at b * sBias, S1 / S2 / D at b * their stride (elements)
    int batch = 1, a_div = 1;
    long sA = 0, sB = 0, sBias = 0, sS1 = 0, sS2 = 0, sD = 0;
    // filled by gemm_launch: byte extents of A and B for the buffer range checks
    uint32_t a_bytes = 0, b_bytes = 0;
};

int gemm_validate(const GemmArgs& g, int epi, int amode, int dmode);
// Whether the codec layers in PREC_FP8 store e4m3 copies of their
// activations for the next product (PREC_FP8A) instead of converting fp16
// operands inside every product.  Same bits either way; default on.
void set_fp8_activations(bool on);
bool fp8_activations();
hipError_t gemm_launch(const GemmArgs& g, int epi, int amode, int dmode, hipStream_t stream);
// 0: automatic tile choice; 1-7: force one tile configuration (tests, tuning)
int gemm_set_tile_policy(int policy);
// Replaces the tuning table with n entries of 8 values {prec, epi, amode,
// dmode, M, N, K, configuration}; n = 0 clears it.
int gemm_tune_set(const long long* entries, int n);
// The configuration a launch of this GEMM uses now (policy, table or heuristic).
int gemm_tile_for(const GemmArgs& g, int epi, int amode, int dmode);

// Weight conversion for the low-precision products, on the device: rows x K
// fp16 -> bf16 (round to nearest even), or -> fp8 e4m3 with scale[r] =
// max_k |w[r, k]| / 448 (1 for an all-zero row) and q = w / scale.
hipError_t weights_to_bf16(const f16* w, long n, void* out, hipStream_t stream);
hipError_t weights_to_fp8(const f16* w, int rows, int K, void* out, float* scale, hipStream_t stream);

}  // namespace dcvc
