/*
 * dcvc_codec.h — codec-level C ABI of the MI355X DCVC-UF hot path.
 *
 * 1. The CPU entropy coder, replacing the reference's MLCodec_extensions_cpp
 *    module (src/cpp/py_rans/bind.cpp:14-40, py_rans.h:15-78) with the same
 *    stream format.
 * 2. The codec proxies, replacing the reference's inference_extensions_cuda
 *    classes DMCIProxy / DMCHTSProxy (src/layers/extensions/inference/
 *    bind.cpp:11-39): set_param / compress / decompress /
 *    add_ref_feature_from_frame over plain device pointers.
 *
 * Return value 0 = success; otherwise a status (1 bad argument, 2 HIP
 * failure, 3 malformed bitstream, 4 coder overflow, 5 call order, 6 budget).
 * dcvc_last_error() gives the message of the calling thread's last failure.
 */
#ifndef DCVC_CODEC_H
#define DCVC_CODEC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* dcvc_last_error(void);

/* ---------------- entropy coder (MLCodec_extensions_cpp) ---------------- */
/* pmf_to_quantized_cdf (py_rans.cpp:35-94): cdf_out receives n + 1 values. */
int dcvc_pmf_to_quantized_cdf(const float* pmf, int n, uint32_t* cdf_out);
/* compute_ec_parallel (dmc_common.cpp:31-35) */
int dcvc_compute_ec_parallel(long symbol_count);

/* RansEncoder (py_rans.h:21-48).  index 0 = z (bit estimator), 1 = y
 * (gaussian).  cdfs is n_cdf rows of row_len int32; sizes = cdf_length. */
void* dcvc_rans_encoder_create(void);
void dcvc_rans_encoder_destroy(void* enc);
int dcvc_rans_encoder_set_cdf(void* enc, const int32_t* cdfs, int n_cdf, int row_len,
                              const int32_t* sizes, int index);
int dcvc_rans_encoder_reset(void* enc);
int dcvc_rans_encoder_set_parallel(void* enc, int n);
/* symbols must stay valid until dcvc_rans_encoder_flush returns */
int dcvc_rans_encoder_encode_y(void* enc, const int16_t* symbols, long n);
int dcvc_rans_encoder_encode_z(void* enc, const int8_t* symbols, long n, int cdf_offset, int ch);
int dcvc_rans_encoder_flush(void* enc);
/* merged stream of the last flush (valid until the next reset / flush) */
int dcvc_rans_encoder_stream(void* enc, const uint8_t** data, long* size);

/* RansDecoder (py_rans.h:50-78) */
void* dcvc_rans_decoder_create(void);
void dcvc_rans_decoder_destroy(void* dec);
int dcvc_rans_decoder_set_cdf(void* dec, const int32_t* cdfs, int n_cdf, int row_len,
                              const int32_t* sizes, int index);
int dcvc_rans_decoder_set_parallel(void* dec, int n);
int dcvc_rans_decoder_set_stream(void* dec, const uint8_t* data, long size);
int dcvc_rans_decoder_decode_z(void* dec, long n, int cdf_offset, int ch, int8_t* out);
int dcvc_rans_decoder_decode_y(void* dec, const uint8_t* indexes, long n, int8_t* out);

/* ---------------- codec proxies (inference_extensions_cuda) ----------------
 * set_param (dmci_proxy.cpp:604-652, dmc_hts_proxy.cpp:712-762): the model's
 * state dict as n named float32 host arrays (shape[i*4 .. i*4+ndim[i]-1]),
 * including gaussian_encoder.{quantized_cdf,cdf_length} and
 * bit_estimator_z.{quantized_cdf,cdf_length}; skip_thres as in the reference.
 * Pictures are NHWC fp16 device buffers (pixel stride ld), values in
 * [-0.5, 0.5].  `stream` is the caller's hipStream_t; the proxy orders its
 * own stream after it and the caller's stream after the proxy's work.
 * Threading: one proxy is not re-entrant; different proxies may run
 * concurrently from different host threads (each captures its graphs on its
 * own stream), except set_param, which uploads on the legacy stream and
 * must not overlap other proxies' calls. */

/* DMCIProxy (dmci_proxy.h:132-254) */
void* dcvc_dmci_create(void);
void dcvc_dmci_destroy(void* h);
int dcvc_dmci_set_param(void* h, int n, const char** names, const float** data, const int64_t* shapes,
                        const int* ndims, float skip_thres);
/* buffers for a Hp x Wp (multiple of 16) picture; x_hat = caller-owned
 * [Hp, Wp, 3] output, the tensor compress/decompress return (a proxy-owned
 * view in the reference, dmci_proxy.cpp:418, :601). */
int dcvc_dmci_prepare(void* h, int Hp, int Wp, void* x_hat, int ld_out);
/* DMCIProxy::compress (dmci_proxy.cpp:296-421): x is [H, W, 3] unpadded. */
int dcvc_dmci_compress(void* h, const void* x, int ldx, int H, int W, int qp, int pad_b, int pad_r,
                       void* stream, const uint8_t** bits, long* nbytes, int* ec_parallel);
/* DMCIProxy::decompress (dmci_proxy.cpp:423-602) */
int dcvc_dmci_decompress(void* h, const uint8_t* bits, long nbytes, int qp, int height, int width,
                         int ec_parallel, void* stream);

/* DMCHTSProxy (dmc_hts_proxy.h:215-355) */
void* dcvc_hts_create(void);
void dcvc_hts_destroy(void* h);
int dcvc_hts_set_param(void* h, int n, const char** names, const float** data, const int64_t* shapes,
                       const int* ndims, float skip_thres);
/* Shares the weights of `src` (set_param done there) with `h`: one weight
 * copy in HBM for many stream contexts.  No reference counterpart (the
 * reference keeps one weight set per proxy). */
int dcvc_hts_share_weights(void* h, void* src);
/* 8 caller-owned [Hp, Wp, 3] reconstruction outputs (dmc_hts_proxy.cpp:347) */
int dcvc_hts_prepare(void* h, int Hp, int Wp, void* const* frames, int ld_out);
/* add_ref_feature_from_frame (dmc_hts_proxy.cpp:492-502): frame [Hp, Wp, 3] */
int dcvc_hts_add_ref_feature_from_frame(void* h, const void* frame, int ld, int Hp, int Wp,
                                        int apply_adaptor, void* stream);
/* compress (dmc_hts_proxy.cpp:504-585): x is [H, W, 24], 8 frames on channels */
int dcvc_hts_compress(void* h, const void* x, int ldx, int H, int W, int qp, int reset_feature_memory,
                      int pad_b, int pad_r, void* stream, const uint8_t** bits, long* nbytes,
                      int* ec_parallel);
/* decompress (dmc_hts_proxy.cpp:587-710): writes the 8 prepared frames */
int dcvc_hts_decompress(void* h, const uint8_t* bits, long nbytes, int qp, int height, int width,
                        int ec_parallel, int reset_feature_memory, void* stream);
/* Switch hipGraph capture on (default) or off (eager launches). */
int dcvc_hts_set_use_graphs(void* h, int on);
int dcvc_dmci_set_use_graphs(void* h, int on);
/* Precision form (no reference counterpart; the reference runs FP16 only):
 * every 1x1-conv product of the proxy in 0 FP16 (default), 1 BF16 or 2 FP8
 * E4M3 with per-output-channel weight scales, FP32 accumulate.  Activations,
 * depthwise, strided / sub-pixel convs and the quantise / index steps stay
 * FP16 / FP32.  Encoder and decoder must use the same precision. */
int dcvc_hts_set_precision(void* h, int prec);
/* Device bytes held by the proxy's activation buffers (its resident stream
 * context; weights excluded).  No reference counterpart: used for admission
 * by HBM footprint. */
long dcvc_hts_device_bytes(void* h);
int dcvc_dmci_set_precision(void* h, int prec);
/* Test hook: copies an internal buffer (0 y_hat|reduced, 1 y, 2 q_dec|scales|
 * means, 3 memory|feature, 4 x|ctx, 5 dense symbols, 6 z) to host memory. */
int dcvc_hts_debug_copy(void* h, int what, void* host, long bytes);
/* DMCI test hook: 0-3 dense symbols of encoder step k, 4-7 their conditions,
 * 8 the final y_hat. */
int dcvc_dmci_debug_copy(void* h, int what, void* host, long bytes);
/* Symbols handed to the entropy coder by the last call (per step for DMCI). */
long dcvc_hts_last_symbol_count(void* h);
/* Batches of streams (batching form; no reference counterpart: the
 * reference holds one proxy per stream, dmc_hts_proxy.cpp:892-1000).  One
 * proxy codes n_img (1-16) streams of one resolution together: every layer
 * is one launch over the n_img pictures, each stream keeps its own QP,
 * temporal context and bitstream, and each stream's bits equal those of a
 * proxy of its own.  dcvc_hts_prepare_batch is dcvc_hts_prepare with the 8
 * outputs of shape [n_img, Hp, Wp, 3]; add_ref_feature_from_frame then takes
 * [n_img, Hp, Wp, 3]; compress_batch takes x [n_img, H, W, 24] and n_img
 * entries in qps (in), bits / nbytes / ec_parallel (out; the bitstreams stay
 * valid until the next compress); decompress_batch takes n_img bitstreams,
 * QPs and ec_parallel values.  The single-stream compress / decompress
 * refuse a proxy prepared for more than one stream.  batch_symbol_count is
 * last_symbol_count of one stream (-1 for a bad index). */
int dcvc_hts_prepare_batch(void* h, int n_img, int Hp, int Wp, void* const* frames, int ld_out);
int dcvc_hts_compress_batch(void* h, const void* x, int ldx, int H, int W, const int* qps, int reset_feature_memory,
                            int pad_b, int pad_r, void* stream, const uint8_t** bits, long* nbytes,
                            int* ec_parallel);
int dcvc_hts_decompress_batch(void* h, const uint8_t* const* bits, const long* nbytes, const int* qps, int height,
                              int width, const int* ec_parallel, int reset_feature_memory, void* stream);
long dcvc_hts_batch_symbol_count(void* h, int img);
/* DMCHTLProxy (dmc_htl_proxy.h; bind.cpp exports it as DMCHTLProxy).  Same
 * call sequence and buffer contracts as the HT-S entry points above; the y
 * latent is coded in four entropy round trips (dmc_htl_proxy.cpp:718-915). */
void* dcvc_htl_create(void);
void dcvc_htl_destroy(void* h);
/* set_param (dmc_htl_proxy.cpp:917-967) */
int dcvc_htl_set_param(void* h, int n, const char** names, const float** data, const int64_t* shapes,
                       const int* ndims, float skip_thres);
/* One weight copy for many stream contexts (no reference counterpart). */
int dcvc_htl_share_weights(void* h, void* src);
/* pre_allocate_tensors (dmc_htl_proxy.cpp:1042-1178) with 8 caller-owned
 * [Hp, Wp, 3] reconstruction outputs */
int dcvc_htl_prepare(void* h, int Hp, int Wp, void* const* frames, int ld_out);
/* add_ref_feature_from_frame (dmc_htl_proxy.cpp:583-593) */
int dcvc_htl_add_ref_feature_from_frame(void* h, const void* frame, int ld, int Hp, int Wp,
                                        int apply_adaptor, void* stream);
/* compress (dmc_htl_proxy.cpp:595-716): x is [H, W, 24], 8 frames on channels */
int dcvc_htl_compress(void* h, const void* x, int ldx, int H, int W, int qp, int reset_feature_memory,
                      int pad_b, int pad_r, void* stream, const uint8_t** bits, long* nbytes,
                      int* ec_parallel);
/* decompress (dmc_htl_proxy.cpp:718-915): writes the 8 prepared frames */
int dcvc_htl_decompress(void* h, const uint8_t* bits, long nbytes, int qp, int height, int width,
                        int ec_parallel, int reset_feature_memory, void* stream);
int dcvc_htl_set_use_graphs(void* h, int on);
/* Precision form, as dcvc_hts_set_precision (no reference counterpart). */
int dcvc_htl_set_precision(void* h, int prec);
long dcvc_htl_device_bytes(void* h);
/* Test hook: 0 y_hat, 1 y, 2 q_dec|scales|means, 3 memory|feature, 4 x|ctx,
 * 5-8 dense symbols of encoder step k, 9-12 their conditions, 13 z. */
int dcvc_htl_debug_copy(void* h, int what, void* host, long bytes);
/* Symbols (encoder) or indexes (decoder) of checkerboard step k in the last call. */
long dcvc_htl_step_symbol_count(void* h, int step);
/* DMCLDProxy (dmc_ld_proxy.h; bind.cpp exports it as DMCLDProxy): the
 * low-delay model, one frame per call, one entropy round trip. */
void* dcvc_ld_create(void);
void dcvc_ld_destroy(void* h);
/* set_param (dmc_ld_proxy.cpp:595-637) */
int dcvc_ld_set_param(void* h, int n, const char** names, const float** data, const int64_t* shapes,
                      const int* ndims, float skip_thres);
/* One weight copy for many stream contexts (no reference counterpart). */
int dcvc_ld_share_weights(void* h, void* src);
/* pre_allocate_tensors (dmc_ld_proxy.cpp:693-794) with a caller-owned
 * [Hp, Wp, 3] reconstruction output */
int dcvc_ld_prepare(void* h, int Hp, int Wp, void* x_hat, int ld_out);
/* add_ref_feature_from_frame (dmc_ld_proxy.cpp:407-418) */
int dcvc_ld_add_ref_feature_from_frame(void* h, const void* frame, int ld, int Hp, int Wp,
                                       int apply_adaptor, void* stream);
/* compress (dmc_ld_proxy.cpp:420-485): x is [H, W, 3], one frame */
int dcvc_ld_compress(void* h, const void* x, int ldx, int H, int W, int qp, int reset_feature_memory,
                     int pad_b, int pad_r, void* stream, const uint8_t** bits, long* nbytes,
                     int* ec_parallel);
/* decompress (dmc_ld_proxy.cpp:487-593): writes the prepared output */
int dcvc_ld_decompress(void* h, const uint8_t* bits, long nbytes, int qp, int height, int width,
                       int ec_parallel, int reset_feature_memory, void* stream);
int dcvc_ld_set_use_graphs(void* h, int on);
/* Precision form, as dcvc_hts_set_precision (no reference counterpart). */
int dcvc_ld_set_precision(void* h, int prec);
long dcvc_ld_device_bytes(void* h);
/* Test hook: 0 y_hat|q_dec|scales|means, 1 y, 3 memory|feature, 4 up|x|ctx,
 * 5 dense symbols, 6 z, 7 hyper|temporal params. */
int dcvc_ld_debug_copy(void* h, int what, void* host, long bytes);
long dcvc_ld_last_symbol_count(void* h);
/* Layer-table trace (no reference counterpart): between begin and end, every
 * GEMM / conv / depthwise launch issued from the host is recorded; end
 * returns the record count (or -status) and fills up to max_records rows of
 * 15 int64: kind (0 1x1, 1 stride-2 conv, 2 sub-pixel conv, 3 depthwise,
 * 4 fused FFN), epilogue, M, N, K, algorithmic FLOPs, algorithmic bytes,
 * launch duration in ns (timed != 0: HIP events recorded around each launch
 * on its own stream; else -1), then the tuning key and choice: A mode,
 * D mode, operand precision, tile configuration (depthwise: strip rows;
 * fused FFN: -1), and the grid the launch works on (images, rows, columns;
 * the output grid for GEMMs, 0 for the fused FFN).  Graph replays launch nothing from the host, so trace with
 * graphs off. */
int dcvc_trace_begin(int timed);
int dcvc_trace_end(long long* records, int max_records);
long dcvc_dmci_step_symbol_count(void* h, int step);

#ifdef __cplusplus
}
#endif

#endif /* DCVC_CODEC_H */
