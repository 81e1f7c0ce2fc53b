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

#ifdef __cplusplus
}
#endif

#endif /* DCVC_CODEC_H */
