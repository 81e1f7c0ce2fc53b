// Internal interface of the fused FFN kernel (see ffn.hip): the ffn half of a
// DepthConvBlock, the quad GEMM of ffn.0 and the ffn.2 GEMM with its
// shortcut tail, in one launch with the hidden activation kept in LDS.
#pragma once
#include <hip/hip_runtime.h>
#include "common.h"

namespace dcvc {

struct FfnArgs {
    const f16* A = nullptr;     // block input o [M, C] (row stride lda); also the first shortcut
    int lda = 0;
    int M = 0, C = 0, CM = 0;   // pixels, block channels, hidden channels
    const f16* W0 = nullptr;    // ffn.0 weight [4 CM, C]: rows 4n .. 4n + 3 feed hidden channel n
    const float* b0 = nullptr;  // [4 CM]
    const f16* W2 = nullptr;    // ffn.2 weight [C, CM]
    const float* b2 = nullptr;  // [C]
    const f16* S2 = nullptr;    // EPI_SHORTCUT2: second shortcut
    int lds2 = 0;
    const f16* q = nullptr;     // EPI_SHORTCUT_QUANT: per-QP table, row stride qstride
    int qstride = 0;
    const int* qp = nullptr;
    int q_round = 0;
    f16* D = nullptr;           // output [M, C]; may be A or S2 itself (same row stride)
    int ldd = 0;
};

// Shapes the fused kernel is instantiated for.
bool ffn_supported(int C, int CM);
// epi: EPI_SHORTCUT (D = ffn2 + b + A), EPI_SHORTCUT2 (+ S2) or
// EPI_SHORTCUT_QUANT ((ffn2 + b + A) * q[n]).
int ffn_validate(const FfnArgs& f, int epi);
hipError_t ffn_launch(const FfnArgs& f, int epi, hipStream_t stream);
// Row blocks of one launch (128 pixels each): the kernel runs one block per
// CU, so it pays off only when the grid covers most of the 256 CUs.
inline int ffn_blocks(int M) { return (M + 127) / 128; }
// Use of the fused kernel by DepthConvBlock: 0 (default) never; 1 for fp16
// blocks with C = 256 and a grid of at least 192 blocks; 2 whenever the shape
// is supported and the precision is fp16.  In isolation the fused launch is
// faster at C = 256 (1080p: 42 us against 48 us for the two launches) and
// slower at C = 512 (78 against 74 us): it needs one workgroup per CU, which
// leaves its K-tile loads exposed.  Inside the codecs the decoder runs the
// recon head halves on two streams at once, and two such launches cannot
// share CUs, so the fused blocks took 58 us there against 46 us for the two
// launches; it stays off unless asked for.  Graphs captured before a change
// keep the launches they recorded.
int ffn_set_mode(int mode);
int ffn_mode();
// DepthConvBlock's choice for M pixels under the current mode.
inline bool ffn_use(int M, int C, int CM, int prec) {
    const int mode = ffn_mode();
    if (mode == 0 || prec != 0 || !ffn_supported(C, CM)) return false;
    return mode == 2 || (C == 256 && ffn_blocks(M) >= 192);
}

}  // namespace dcvc
