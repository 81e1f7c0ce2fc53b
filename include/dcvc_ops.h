/*
 * dcvc_ops.h — kernel-level C ABI of the MI355X DCVC-UF hot path.
 *
 * Each entry point replaces one free function of the reference's
 * inference_extensions_cuda operator layer (src/layers/extensions/inference/
 * def_cutlass.h:10-39 and def_elementwise.h:10-78), taking plain device
 * pointers, element strides and a hipStream_t (as void*) instead of
 * at::Tensor.  Activations are NHWC fp16: a feature map of C channels is an
 * [pixels x C] matrix whose rows are `ld` halves apart (ld >= C), so channel
 * slices of a concatenation buffer are passed as (pointer + offset, ld).
 *
 * All functions are asynchronous on `stream` and return 0 (DCVC_OK) or a
 * nonzero status (1 bad argument, 2 HIP failure).  Nothing here allocates.
 */
#ifndef DCVC_OPS_H
#define DCVC_OPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Library version (major * 10000 + minor * 100 + patch). */
int dcvc_version(void);

/* Fills th[0..127] with the float32 scale->CDF-index thresholds:
 * index(s) = #{k >= 1 : s >= th[k]}, the FP32 value of
 * floor((ln(clamp(s, 0.11, 16)) + 2.2073) * 127 / 4.9799)
 * (def_const.h:6-11, elementwise/stream.cu:77-87).  Host-only. */
int dcvc_scale_index_thresholds(float* th128);

/* GEMM tile configuration: 0 automatic (default), 1-14 force one of the
 * tile configurations for every GEMM launched afterwards (process-wide).
 * No counterpart in the reference: used by tests (all configurations must
 * give bit-identical results) and by the autotuner.  Host-only. */
int dcvc_gemm_set_tile_policy(int policy);

/* Tile-choice database of the autotuner, the counterpart of the reference's
 * per-shape tuned maps HINT_MAP_BIAS (cutlass/conv1x1_bias.cu:8-522) and
 * HINT_MAP_D3x3 (cutlass/d3x3.cu:7-441), which launch_cutlass_helper
 * consults (cutlass/cutlass_kernel.h:233-249).  dcvc_gemm_tune_set replaces
 * the GEMM table with n entries of 8 int64 {prec, epi, amode, dmode, M, N, K,
 * config (1-14)} (prec as dcvc_gemm_lowp, or 3: prec 2 on e4m3 activations,
 * dcvc_gemm_fp8_act): a GEMM of that key launches with that configuration, other
 * shapes keep the built-in heuristic (n = 0 clears).  dcvc_d3x3_tune_set does
 * the same for the depthwise strip length, entries {n_img, H, W, C, rows (4
 * or 8)}.  The *_for functions return the choice a launch makes now (or
 * -1 for bad arguments).  The choice never changes results.  Install before
 * graphs are captured: a captured graph keeps its launches.  Host-only. */
int dcvc_gemm_tune_set(const long long* entries, int n);
int dcvc_gemm_tile_for(int prec, int epi, int amode, int dmode, long M, int N, int K);
int dcvc_d3x3_tune_set(const long long* entries, int n);
int dcvc_d3x3_strip_for(int n_img, int H, int W, int C);

/* Calibration kernel, no counterpart in the reference: measures the rate the
 * matrix cores sustain on given operand values, so a GEMM's MFMA fraction can
 * be read against it as well as against the nominal peak.  `blocks`
 * workgroups of 4 waves each issue 16 * `iters` MFMAs on operands held in
 * registers: prec 0 v_mfma_f32_32x32x16_f16 and 1 v_mfma_f32_32x32x16_bf16
 * (2 * 32 * 32 * 16 FLOP each, operands the first 16 bytes of each 32-byte
 * fragment), prec 2 v_mfma_scale_f32_32x32x64_f8f6f4 on OCP e4m3 (2 * 32 *
 * 32 * 64 FLOP each).  `operands` holds blocks * 2048 fragments of 32 bytes
 * (blocks * 64 KiB, 32-byte aligned); `sink` receives blocks * 256 floats.
 * Asynchronous on `stream`. */
int dcvc_mfma_calibrate(const void* operands, float* sink, int blocks, int iters, int prec, void* stream);

/* ---------------- 1x1-conv GEMMs (MFMA, fp32 accumulate) ----------------
 * A [M x K] fp16 (row stride lda), W [N x K] fp16 (PyTorch [N,K,1,1]),
 * bias [N] fp32, D [M x N] fp16 (row stride ldd).  K % 64 == 0,
 * N % 8 == 0, lda / ldd / shortcut strides % 8 == 0, A / W / D / shortcut
 * pointers 16-byte aligned.
 * D may alias S1 (in-place shortcut), as the reference's out_buf does.
 * q: per-QP table of fp16 rows (row stride qstride); qp: device int* row
 * index, or NULL for row 0. */

/* conv1x1_bias (cutlass/conv1x1_bias.cu:524-540): D = A W^T + b */
int dcvc_conv1x1_bias(const void* A, int lda, int M, int K, const void* W, int N, const float* bias,
                      void* D, int ldd, void* stream);
/* conv1x1_bias_wsilu (conv1x1_bias_wsilu.cu:230-248): D = wsilu(A W^T + b) */
int dcvc_conv1x1_bias_wsilu(const void* A, int lda, int M, int K, const void* W, int N,
                            const float* bias, void* D, int ldd, void* stream);
/* conv1x1_bias_shortcut (conv1x1_bias_shortcut.cu:230-247): D = A W^T + b + S */
int dcvc_conv1x1_bias_shortcut(const void* A, int lda, int M, int K, const void* W, int N,
                               const float* bias, const void* S, int lds, void* D, int ldd,
                               void* stream);
/* conv1x1_bias_shortcut2 (conv1x1_bias_shortcut2.cu:81-100): D = A W^T + b + S1 + S2 */
int dcvc_conv1x1_bias_shortcut2(const void* A, int lda, int M, int K, const void* W, int N,
                                const float* bias, const void* S1, int lds1, const void* S2,
                                int lds2, void* D, int ldd, void* stream);
/* conv1x1_bias_shortcut_with_quant (…_with_quant.cu:50-70): D = (A W^T + b + S) * q[n] */
int dcvc_conv1x1_bias_shortcut_with_quant(const void* A, int lda, int M, int K, const void* W,
                                          int N, const float* bias, const void* S, int lds,
                                          const void* q, int qstride, const int* qp, void* D,
                                          int ldd, void* stream);
/* conv1x1_bias_shortcut followed by multiply_with_broadcast_cuda, the HT-L
 * block tail (dmc_htl_proxy.cpp:39-40, 102-103), in one pass:
 * D = fp16(fp16(A W^T + b + S) * q[n]), both fp16 roundings kept. */
int dcvc_conv1x1_bias_shortcut_then_quant(const void* A, int lda, int M, int K, const void* W,
                                          int N, const float* bias, const void* S, int lds,
                                          const void* q, int qstride, const int* qp, void* D,
                                          int ldd, void* stream);
/* conv1x1_bias_with_quant (conv1x1_bias_with_quant.cu:30-48): D = (A W^T + b) * q[n] */
int dcvc_conv1x1_bias_with_quant(const void* A, int lda, int M, int K, const void* W, int N,
                                 const float* bias, const void* q, int qstride, const int* qp,
                                 void* D, int ldd, void* stream);
/* conv1x1_bias_wsilu_chunk_add, the quad GEMM
 * (conv1x1_bias_wsilu_chunk_add.cu:356-390): W [4N x K] and bias [4N] in the
 * reference ffn.0 order; D[m, n] = sum_{j<4} wsilu(A W[4n+j]^T + b[4n+j]). */
int dcvc_conv1x1_bias_wsilu_chunk_add(const void* A, int lda, int M, int K, const void* W, int N,
                                      const float* bias, void* D, int ldd, void* stream);
/* The ffn half of DepthConvBlock, which the reference's proxy runs as
 * conv1x1_bias_wsilu_chunk_add followed by conv1x1_bias_shortcut /
 * _shortcut2 / _shortcut_with_quant (layers_proxy.cpp:84-97), in one launch:
 *   T = quad(A W0^T + b0)            (W0 [4 CM x C], b0 [4 CM], as above)
 *   D = tail(T W2^T + b2 + A)        (W2 [C x CM], b2 [C])
 * tail by epi: 2 nothing, 3 + S2, 4 * q[n] (after rounding to fp16 first
 * when q_round is nonzero, the HT-L order).  A is [M x C] (row stride lda);
 * D may be A or S2 itself with the same row stride.  Supported: CM = 256,
 * C in {256, 512}.  Bit-identical to the two entry points run in sequence. */
int dcvc_depth_conv_ffn(const void* A, int lda, int M, int C, int CM, const void* W0, const float* b0,
                        const void* W2, const float* b2, int epi, const void* S2, int lds2, const void* q,
                        int qstride, const int* qp, int q_round, void* D, int ldd, void* stream);
/* DepthConvBlock's use of dcvc_depth_conv_ffn inside the codecs (no
 * reference counterpart): 0 never (default: with the decoder's two
 * concurrent streams the one-workgroup-per-CU fused launch is slower than
 * the two launches); 1 for fp16 blocks with C = 256 whose grid covers at
 * least 192 of the 256 CUs (faster when the GPU runs nothing else); 2
 * whenever supported.  Graphs captured before a change keep their
 * launches. */
int dcvc_set_ffn_fusion(int mode);

/* Grouped 1x1-conv product (no reference counterpart: the reference's recon
 * heads launch one conv per frame branch): `batch` problems of one shape in
 * one launch.  Problem b reads A + (b / a_div) * sA (a_div > 1: consecutive
 * problems share one input), W + b * sW, bias + b * sBias, S1 / S2 / D + b *
 * their stride (all strides in elements, multiples of 8 except sBias).  epi:
 * 0 bias, 1 wsilu, 2 shortcut, 3 shortcut2, 6 quad (W rows 4N per problem,
 * sW = 4 N K, sBias = 4 N).  Each problem's result equals the single-problem
 * entry point's bit for bit. */
int dcvc_conv1x1_grouped(const void* A, int lda, long sA, int a_div, int M, int K, const void* W, long sW, int N,
                         const float* bias, long sBias, int epi, const void* S1, int lds1, long sS1, const void* S2,
                         int lds2, long sS2, void* D, int ldd, long sD, int batch, void* stream);

/* Low-precision 1x1-conv product (precision form; no reference counterpart).
 * dcvc_weights_to_lowp converts fp16 weights [rows x K] on the device:
 * prec 1 -> bf16 (Wq: rows*K*2 bytes, scale unused), prec 2 -> fp8 e4m3
 * (Wq: rows*K bytes) with scale[r] = max|w[r,:]| / 448.  dcvc_gemm_lowp runs
 * the product with epilogue `epi` (0 bias, 1 wsilu, 2 shortcut, 3 shortcut2,
 * 4 shortcut_with_quant, 5 with_quant, 6 quad: N outputs from 4N rows) on
 * fp16 activations converted to the operand precision (fp8 at scale 1,
 * saturated at +-448); prec 0 takes fp16 weights.  Other arguments as the
 * conv1x1 entry points; unused shortcut / quant pointers may be NULL. */
int dcvc_weights_to_lowp(const void* W, int rows, int K, int prec, void* Wq, float* scale, void* stream);
int dcvc_gemm_lowp(const void* A, int lda, int M, int K, const void* Wq, int N, const float* scale, int prec,
                   int epi, const float* bias, const void* S1, int lds1, const void* S2, int lds2, const void* q,
                   int qstride, const int* qp, void* D, int ldd, void* stream);

/* fp8 activations in memory (precision form; no reference counterpart).
 * dcvc_gemm_fp8_act is dcvc_gemm_lowp with two additions:
 *   a_fp8 = 1: A holds e4m3 bytes (row stride lda in bytes, lda % 16 == 0)
 *     as an earlier D8 / Y8 output stored them; requires prec 2, and gives
 *     the same bits as prec 2 on the fp16 activations they were made from;
 *   D8 (may be NULL; prec 2 only): an e4m3 copy of the fp16 output [M x N]
 *     (row stride ldd8 bytes, % 8 == 0), saturated at +-448, round to nearest
 *     even at scale 1; D may then be NULL (e4m3 output only).
 * dcvc_d3x3_fp8_out is dcvc_d3x3 with the same optional Y8 copy (Y may be
 * NULL when Y8 is given).  dcvc_set_fp8_activations(0) makes the codec
 * proxies in precision 2 convert fp16 activations inside every product
 * instead of storing e4m3 copies (default 1; the bits are the same either
 * way; takes effect at the next graph capture).  dcvc_get_fp8_activations
 * returns the setting. */
int dcvc_gemm_fp8_act(const void* A, int lda, int a_fp8, int M, int K, const void* Wq, int N, const float* scale,
                      int prec, int epi, const float* bias, const void* S1, int lds1, const void* S2, int lds2,
                      const void* q, int qstride, const int* qp, void* D, int ldd, void* D8, int ldd8, void* stream);
int dcvc_d3x3_fp8_out(const void* X, int ldx, int n_img, int H, int W, int C, const void* Wt, void* Y, int ldy,
                      void* Y8, int ldy8, void* stream);
int dcvc_set_fp8_activations(int on);
int dcvc_get_fp8_activations(void);

/* conv_bias (cutlass/conv_bias.cu:131-151): k x k conv, stride s, pad (k-1)/2,
 * over n_img NHWC images of H x W x Cin.  (k, s) in {(2,2), (3,2), (3,1)}.
 * W is [N][k*k][Cin] (tap-major, i.e. OHWI).  Output [n_img, Ho, Wo, N]. */
int dcvc_conv_bias(const void* X, int ldx, int n_img, int H, int W, int Cin, const void* Wt, int N,
                   const float* bias, int kernel, int stride, void* D, int ldd, void* stream);

/* SubpelConv2x with a 3x3 pad-1 kernel and bias (layers.py:92-103; the
 * proxy's biased branch conv2d + pixel_shuffle_2_cuda, layers_proxy.cpp:
 * 270-276), the shuffle fused into the GEMM store.  W is [4][Cout][9][Cin]
 * (sub-pixel major, taps then input channels), bias [4*Cout] in the same
 * order.  Output [n_img, 2H, 2W, Cout]. */
int dcvc_subpel_conv3x3(const void* X, int ldx, int n_img, int H, int W, int Cin, const void* Wt,
                        int Cout, const float* bias, void* D, int ldd, void* stream);

/* transposed_conv (cutlass/transposed_conv.cu:101-119): SubpelConv2x with a
 * 1x1 kernel, i.e. 1x1 conv to 4*Cout channels then pixel_shuffle(2).
 * W is [4][Cout][Cin] (sub-pixel major: sub = dy*2+dx); bias [4*Cout] in the
 * same order, or NULL for none.  Output [n_img, 2H, 2W, Cout]. */
int dcvc_transposed_conv(const void* X, int ldx, int n_img, int H, int W, int Cin, const void* Wt,
                         int Cout, const float* bias, void* D, int ldd, void* stream);

/* d3x3 (cutlass/d3x3.cu:443-446): depthwise 3x3, pad 1, no bias.
 * Wt is [9][C] (tap ky*3+kx major). */
int dcvc_d3x3(const void* X, int ldx, int n_img, int H, int W, int C, const void* Wt, void* Y,
              int ldy, void* stream);

/* ---------------- layout / scaling ---------------- */
/* pad_and_unshuffle_8_cuda / pixel_unshuffle_cuda (cat_and_pad.cu:7-51,
 * shuffle.cu:134-185): replicate-pad by (padB, padR) then unshuffle by f. */
int dcvc_pad_and_unshuffle(const void* X, int ldx, int n_img, int H, int W, int Cin, int f,
                           int padB, int padR, void* Y, int ldy, void* stream);
/* pixel_shuffle_8_cuda with clamp to [-0.5, 0.5] (shuffle.cu:224-316) */
int dcvc_pixel_shuffle_8(const void* X, int ldx, int n_img, int H8, int W8, int C, void* Y,
                         int ldy, void* stream);
/* multiply_with_broadcast_cuda (stream.cu:484-546): Y = X * q[qp][c] */
int dcvc_multiply_with_broadcast(const void* X, int ldx, long M, int C, const void* q, int qstride,
                                 const int* qp, void* Y, int ldy, void* stream);
/* round_z_cuda (stream.cu:862-894): clamp(round(z), -64, 63) -> fp16 + int8 (NHWC-linear) */
int dcvc_round_z(const void* Z, int ldz, long M, int C, void* Zhat, int ldh, int8_t* Zi8,
                 void* stream);
/* int8_to_dtype_cuda (stream.cu:454-482) */
int dcvc_int8_to_half(const int8_t* Zi8, long M, int C, void* Zhat, int ldh, void* stream);
/* replicate_pad_cuda / slice_cuda (cat_and_pad.cu:53-110): Ho >= H pads, Ho <= H crops */
int dcvc_replicate_pad(const void* X, int ldx, int n_img, int H, int W, int C, int Ho, int Wo,
                       void* Y, int ldy, void* stream);

/* ---------------- quantisation / index boundary ----------------
 * th: device pointer to the 128-float threshold table; thr: skip threshold.
 * sym: int16 (y_q << 8) + index, cond: uint8 (scale > thr), NHWC-linear. */

/* HT-S encoder mask step k in 0..3 (process_with_mask_no_scale*_cuda,
 * stream.cu:548-683, fused with divide_with_clamp_min :422-443 and
 * build_index_enc :130-161).  Y = raw y; S, Mn, Qd = scales, means, q_dec;
 * Yhat is accumulated in place (written at active positions; step 0 zeroes
 * the rest; step 3 scales everything by max(q_dec, 0.5)). */
int dcvc_hts_quant_step(const void* Y, int ldy, int n_img, int H, int W, int C, const void* S,
                        int lds, const void* Mn, int ldm, const void* Qd, int ldq, void* Yhat,
                        int ldh, int16_t* sym, uint8_t* cond, const float* th, float thr, int step,
                        void* stream);
/* build_index_dec_cuda (stream.cu:89-128): dense index + condition */
int dcvc_hts_build_index_dec(const void* S, int lds, int n_img, int H, int W, int C,
                             const float* th, float thr, uint8_t* idx, uint8_t* cond, void* stream);
/* restore_y_cuda / restore_y_and_add(_multiply)_inplace_cuda (stream.cu:685-754) */
int dcvc_hts_restore_step(const int8_t* yq, int n_img, int H, int W, int C, const void* Mn,
                          int ldm, const void* Qd, int ldq, void* Yhat, int ldh, int step,
                          void* stream);
/* DMCI encoder mask step (process_with_mask_cuda + single_part_for_writing_4x
 * + build_index_enc, dmci_proxy.cpp:333-393): quarter outputs sym/cond of
 * n_img*H*W*C/4 elements; Ysf = y_hat_so_far (in place); step 3 writes
 * Yout = (Ysf + y_hat_3) * q[qp][c]. */
int dcvc_intra_quant_step(const void* Y, int ldy, int n_img, int H, int W, int C, const void* S,
                          int lds, const void* Mn, int ldm, void* Ysf, int ldsf, const void* q,
                          int qstride, const int* qp, void* Yout, int ldo, int16_t* sym,
                          uint8_t* cond, const float* th, float thr, int step, void* stream);
/* single_part_for_reading_4x + build_index_dec (dmci_proxy.cpp:667-676) */
int dcvc_intra_build_index_dec(const void* S, int lds, int n_img, int H, int W, int C,
                               const float* th, float thr, int step, uint8_t* idx, uint8_t* cond,
                               void* stream);
/* restore_y_4x* (stream.cu:756-860) */
int dcvc_intra_restore_step(const int8_t* yq, int n_img, int H, int W, int C, const void* Mn,
                            int ldm, void* Ysf, int ldsf, const void* q, int qstride,
                            const int* qp, void* Yout, int ldo, int step, void* stream);

/* LD encoder half `step` of the 2x checkerboard (process_with_mask_no_scale
 * and process_with_mask_no_scale_add_and_multiply_inplace after
 * divide_with_clamp_min, dmc_ld_proxy.cpp:444-452; stream.cu:548-645):
 * as dcvc_hts_quant_step over channel halves, 2 steps, step 1 scales y_hat
 * by max(Qd, .5).  Indexes on the decoder side: dcvc_hts_build_index_dec. */
int dcvc_ld_quant_step(const void* Y, int ldy, int n_img, int H, int W, int C, const void* S,
                       int lds, const void* Mn, int ldm, const void* Qd, int ldq, void* Yhat,
                       int ldh, int16_t* sym, uint8_t* cond, const float* th, float thr, int step,
                       void* stream);
/* restore_y_cuda / restore_y_and_add_multiply_inplace_cuda (dmc_ld_proxy.cpp:578-580) */
int dcvc_ld_restore_step(const int8_t* yq, int n_img, int H, int W, int C, const void* Mn,
                         int ldm, const void* Qd, int ldq, void* Yhat, int ldh, int step,
                         void* stream);

/* HT-L encoder mask step (dmc_htl_proxy.cpp:619-688: divide_with_clamp_min
 * on y, process_with_mask_cuda, single_part_for_writing_4x, build_index_enc;
 * add_and_multiply_with_clamp_min after step 3): quarter outputs as
 * dcvc_intra_quant_step, with the elementwise quantisation step
 * max(Qd, 0.5): y / max(Qd, .5) is quantised and step 3 writes
 * Yout = (Ysf + y_hat_3) * max(Qd, .5).  Indexes on the decoder side come
 * from dcvc_intra_build_index_dec. */
int dcvc_htl_quant_step(const void* Y, int ldy, int n_img, int H, int W, int C, const void* S,
                        int lds, const void* Mn, int ldm, const void* Qd, int ldq, void* Ysf,
                        int ldsf, void* Yout, int ldo, int16_t* sym, uint8_t* cond,
                        const float* th, float thr, int step, void* stream);
/* restore_y_4x(_and_add_inplace) + add_and_multiply_with_clamp_min
 * (dmc_htl_proxy.cpp:803-905) */
int dcvc_htl_restore_step(const int8_t* yq, int n_img, int H, int W, int C, const void* Mn,
                          int ldm, const void* Qd, int ldq, void* Ysf, int ldsf, void* Yout,
                          int ldo, int step, void* stream);

/* conditional_index_part1_cuda (stream.cu:175-340): stable compaction of N
 * elements of esize (1 or 2) bytes where cond != 0; *count (device int)
 * receives the kept count.  block_offsets: dcvc_compact_blocks(N) ints. */
int dcvc_compact_blocks(long N);
int dcvc_conditional_index(const void* vals, int esize, const uint8_t* cond, long N, void* out,
                           int* block_offsets, int* count, void* stream);
/* dcvc_conditional_index without the block_offsets workspace: no reference
 * counterpart.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The block offsets live in a fixed per-GPU device array of 16384 ints, so
 * N is limited to 16384 * 4096 = 67,108,864 elements (larger N returns 1
 * with a message in dcvc_last_error()).  Nothing is allocated.  All calls on
 * one GPU share that array: calls that overlap on one GPU race, and the
 * offsets of one call are not kept for dcvc_conditional_recover. */
int dcvc_conditional_index_noalloc(const void* vals, int esize, const uint8_t* cond, long N, void* out,
                                   int* count, void* stream);
/* conditional_recover_with_type_conversion_cuda (stream.cu:359-420): dense
 * int8 from compacted decoded symbols, 0 where cond == 0; block_offsets as
 * left by dcvc_conditional_index over the same cond. */
int dcvc_conditional_recover(const int8_t* dec, const uint8_t* cond, long N,
                             const int* block_offsets, int8_t* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DCVC_OPS_H */
