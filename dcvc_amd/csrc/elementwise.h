// Bandwidth-bound kernels of the DCVC-UF hot path: layout shuffles, per-channel
// scaling, the masked quantisation / scale-index boundary and the stream
// compaction that hands symbols to the CPU entropy coder.
#pragma once
#include <hip/hip_runtime.h>
#include "common.h"

namespace dcvc {

// Geometry of an NHWC activation: n_img images of H x W pixels, C channels,
// consecutive pixels ld halves apart.
struct Act {
    f16* p = nullptr;
    int n_img = 1, H = 0, W = 0, C = 0, ld = 0;
    __host__ __device__ long pixels() const { return (long)n_img * H * W; }
};

// Scale->index threshold table (127 float32 entries + unused slot 0); built
// on the host and kept in device memory by the caller.
void host_scale_index_thresholds(float th[128]);

hipError_t pad_unshuffle(const f16* x, int ldx, int n_img, int H, int W, int Cin, int f, int padB,
                         int padR, f16* y, int ldy, hipStream_t s);
hipError_t pixel_shuffle8_clamp(const f16* x, int ldx, int n_img, int H8, int W8, int C, f16* y,
                                int ldy, hipStream_t s);
// y = x * q[qp[m / hw]][c]: hw pixels per image, one device QP index per
// image (hw <= 0: one image of M pixels)
hipError_t mul_channel(const f16* x, int ldx, long M, int C, const f16* q, int qstride,
                       const int* qp, f16* y, int ldy, hipStream_t s, long hw = 0);
hipError_t round_z(const f16* z, int ldz, long M, int C, f16* zhat, int ldh, int8_t* zi8,
                   hipStream_t s);
hipError_t i8_to_f16(const int8_t* zi8, long M, int C, f16* zhat, int ldh, hipStream_t s);
hipError_t replicate_pad(const f16* x, int ldx, int n_img, int H, int W, int C, int Ho, int Wo,
                         f16* y, int ldy, hipStream_t s);

// HT-S (one entropy round trip, 4-step mean-only spatial prior).
hipError_t hts_enc_step(const Act& y, const f16* scales, int lds, const f16* means, int ldm,
                        const f16* qdec, int ldq, f16* yhat, int ldh, int16_t* sym, uint8_t* cond,
                        const float* th, float thr, int step, hipStream_t s);
hipError_t hts_dec_index(const Act& scales, const float* th, float thr, uint8_t* idx,
                         uint8_t* cond, hipStream_t s);
hipError_t hts_dec_step(const Act& geom, const int8_t* yq, const f16* means, int ldm,
                        const f16* qdec, int ldq, f16* yhat, int ldh, int step, hipStream_t s);

// LD (one entropy round trip, 2-step mean-only spatial prior over channel
// halves, get_mask_2x); indexes from hts_dec_index.
hipError_t ld_enc_step(const Act& y, const f16* scales, int lds, const f16* means, int ldm, const f16* qdec, int ldq,
                       f16* yhat, int ldh, int16_t* sym, uint8_t* cond, const float* th, float thr, int step,
                       hipStream_t s);
hipError_t ld_dec_step(const Act& geom, const int8_t* yq, const f16* means, int ldm, const f16* qdec, int ldq,
                       f16* yhat, int ldh, int step, hipStream_t s);

// DMCI (four entropy round trips over quarter tensors).
hipError_t intra_enc_step(const Act& y, const f16* scales, int lds, const f16* means, int ldm,
                          f16* ysf, int ldsf, const f16* q, int qstride, const int* qp, f16* yout,
                          int ldo, int16_t* sym, uint8_t* cond, const float* th, float thr, int step,
                          hipStream_t s);
hipError_t intra_dec_index(const Act& geom, const f16* scales, int lds, const float* th, float thr,
                           int step, uint8_t* idx, uint8_t* cond, hipStream_t s);
hipError_t intra_dec_step(const Act& geom, const int8_t* yq, const f16* means, int ldm, f16* ysf,
                          int ldsf, const f16* q, int qstride, const int* qp, f16* yout, int ldo,
                          int step, hipStream_t s);

// HT-L (four entropy round trips over quarter tensors, elementwise
// quantisation step max(q_dec, 0.5) from the common params).  The decoder
// side builds its indexes with intra_dec_index.
hipError_t htl_enc_step(const Act& y, const f16* scales, int lds, const f16* means, int ldm, const f16* qdec, int ldq,
                        f16* ysf, int ldsf, f16* yout, int ldo, int16_t* sym, uint8_t* cond, const float* th, float thr,
                        int step, hipStream_t s);
hipError_t htl_dec_step(const Act& geom, const int8_t* yq, const f16* means, int ldm, const f16* qdec, int ldq,
                        f16* ysf, int ldsf, f16* yout, int ldo, int step, hipStream_t s);

// Stable stream compaction by a 0/1 byte condition (element size 1 or 2).
// `block_offsets` needs compact_blocks(N) ints and stays valid for recover.
constexpr int COMPACT_TILE = 4096;
inline int compact_blocks(long N) { return (int)((N + COMPACT_TILE - 1) / COMPACT_TILE); }
hipError_t compact(const void* vals, int esize, const uint8_t* cond, long N, void* out,
                   int* block_offsets, int* count, hipStream_t s);
hipError_t recover(const int8_t* dec, const uint8_t* cond, long N, const int* block_offsets,
                   int8_t* out, hipStream_t s);

}  // namespace dcvc
