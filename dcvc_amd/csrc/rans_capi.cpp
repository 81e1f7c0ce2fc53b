// extern "C" entry points of the CPU entropy coder (include/dcvc_codec.h, part 1).
#include <cstring>
#include <exception>
#include <string>

#include "../../include/dcvc_codec.h"
#include "capi_util.h"
#include "rans.h"

using namespace dcvc;

extern "C" {

int dcvc_pmf_to_quantized_cdf(const float* pmf, int n, uint32_t* cdf_out) {
    return guarded([&] {
        if (!pmf || !cdf_out || n <= 0) throw std::invalid_argument("bad pmf");
        auto cdf = pmf_to_quantized_cdf(std::vector<float>(pmf, pmf + n));
        std::memcpy(cdf_out, cdf.data(), cdf.size() * sizeof(uint32_t));
    });
}

int dcvc_compute_ec_parallel(long symbol_count) { return compute_ec_parallel(symbol_count); }

void* dcvc_rans_encoder_create(void) {
    try {
        return new RansEncoder();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}

void dcvc_rans_encoder_destroy(void* enc) { delete static_cast<RansEncoder*>(enc); }

int dcvc_rans_encoder_set_cdf(void* enc, const int32_t* cdfs, int n_cdf, int row_len, const int32_t* sizes,
                              int index) {
    return guarded([&] { checked<RansEncoder>(enc)->set_cdf(cdfs, n_cdf, row_len, sizes, index); });
}

int dcvc_rans_encoder_reset(void* enc) {
    return guarded([&] { checked<RansEncoder>(enc)->reset(); });
}

int dcvc_rans_encoder_set_parallel(void* enc, int n) {
    return guarded([&] { checked<RansEncoder>(enc)->set_parallel(n); });
}

int dcvc_rans_encoder_encode_y(void* enc, const int16_t* symbols, long n) {
    return guarded([&] {
        if (n < 0 || (n > 0 && !symbols)) throw std::invalid_argument("bad symbols");
        checked<RansEncoder>(enc)->encode_y(symbols, n);
    });
}

int dcvc_rans_encoder_encode_z(void* enc, const int8_t* symbols, long n, int cdf_offset, int ch) {
    return guarded([&] {
        if (n < 0 || (n > 0 && !symbols)) throw std::invalid_argument("bad symbols");
        checked<RansEncoder>(enc)->encode_z(symbols, n, cdf_offset, ch);
    });
}

int dcvc_rans_encoder_flush(void* enc) {
    return guarded([&] { checked<RansEncoder>(enc)->flush(); });
}

int dcvc_rans_encoder_stream(void* enc, const uint8_t** data, long* size) {
    return guarded([&] {
        if (!data || !size) throw std::invalid_argument("null output");
        const auto& s = checked<RansEncoder>(enc)->stream();
        *data = s.data();
        *size = (long)s.size();
    });
}

void* dcvc_rans_decoder_create(void) {
    try {
        return new RansDecoder();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}

void dcvc_rans_decoder_destroy(void* dec) { delete static_cast<RansDecoder*>(dec); }

int dcvc_rans_decoder_set_cdf(void* dec, const int32_t* cdfs, int n_cdf, int row_len, const int32_t* sizes,
                              int index) {
    return guarded([&] { checked<RansDecoder>(dec)->set_cdf(cdfs, n_cdf, row_len, sizes, index); });
}

int dcvc_rans_decoder_set_parallel(void* dec, int n) {
    return guarded([&] { checked<RansDecoder>(dec)->set_parallel(n); });
}

int dcvc_rans_decoder_set_stream(void* dec, const uint8_t* data, long size) {
    return guarded([&] { checked<RansDecoder>(dec)->set_stream(data, size); });
}

int dcvc_rans_decoder_decode_z(void* dec, long n, int cdf_offset, int ch, int8_t* out) {
    return guarded([&] {
        if (n < 0 || (n > 0 && !out)) throw std::invalid_argument("bad output");
        checked<RansDecoder>(dec)->decode_z(n, cdf_offset, ch, out);
    });
}

int dcvc_rans_decoder_decode_y(void* dec, const uint8_t* indexes, long n, int8_t* out) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!out || !indexes))) throw std::invalid_argument("bad arguments");
        checked<RansDecoder>(dec)->decode_y(indexes, n, out);
    });
}

}  // extern "C"
