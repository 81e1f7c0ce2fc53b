// extern "C" entry points of the codec proxies (include/dcvc_codec.h, part 2).
#include "../../include/dcvc_codec.h"
#include "capi_util.h"
#include "proxy_dmci.h"
#include "proxy_hts.h"
#include "proxy_htl.h"
#include "proxy_ld.h"

using namespace dcvc;

static HostDict make_dict(int n, const char** names, const float** data, const int64_t* shapes, const int* ndims) {
    if (n <= 0 || !names || !data || !shapes || !ndims) throw std::invalid_argument("empty state dict");
    HostDict d;
    for (int i = 0; i < n; ++i) {
        if (!names[i] || !data[i] || ndims[i] < 0 || ndims[i] > 4) throw std::invalid_argument("bad state dict entry");
        HostTensor t;
        t.shape.assign(shapes + 4 * i, shapes + 4 * i + ndims[i]);
        const int64_t numel = t.numel();
        if (numel < 0) throw std::invalid_argument("bad shape");
        t.data.assign(data[i], data[i] + numel);
        d.emplace(names[i], std::move(t));
    }
    return d;
}

static hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

template <class T>
static void* create() {
    try {
        return new T();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}

extern "C" {

void* dcvc_dmci_create(void) { return create<DMCIProxy>(); }
void dcvc_dmci_destroy(void* h) { delete static_cast<DMCIProxy*>(h); }

int dcvc_dmci_set_param(void* h, int n, const char** names, const float** data, const int64_t* shapes,
                        const int* ndims, float skip_thres) {
    return guarded([&] { checked<DMCIProxy>(h)->set_param(make_dict(n, names, data, shapes, ndims), skip_thres); });
}

int dcvc_dmci_prepare(void* h, int Hp, int Wp, void* x_hat, int ld_out) {
    return guarded([&] {
        if (!x_hat || ld_out < 3) throw std::invalid_argument("bad output buffer");
        checked<DMCIProxy>(h)->prepare(Hp, Wp, static_cast<f16*>(x_hat), ld_out);
    });
}

int dcvc_dmci_compress(void* h, const void* x, int ldx, int H, int W, int qp, int pad_b, int pad_r, void* stream,
                       const uint8_t** bits, long* nbytes, int* ec_parallel) {
    return guarded([&] {
        if (!x || !bits || !nbytes || !ec_parallel) throw std::invalid_argument("null argument");
        *ec_parallel = checked<DMCIProxy>(h)->compress(static_cast<const f16*>(x), ldx, H, W, qp, pad_b, pad_r,
                                                       S(stream), bits, nbytes);
    });
}

int dcvc_dmci_decompress(void* h, const uint8_t* bits, long nbytes, int qp, int height, int width, int ec_parallel,
                         void* stream) {
    return guarded([&] {
        if (!bits && nbytes > 0) throw std::invalid_argument("null bitstream");
        checked<DMCIProxy>(h)->decompress(bits, nbytes, qp, height, width, ec_parallel, S(stream));
    });
}

int dcvc_dmci_set_use_graphs(void* h, int on) {
    return guarded([&] { checked<DMCIProxy>(h)->set_use_graphs(on != 0); });
}

int dcvc_dmci_set_precision(void* h, int prec) {
    return guarded([&] { checked<DMCIProxy>(h)->set_precision(prec); });
}

int dcvc_dmci_debug_copy(void* h, int what, void* host, long bytes) {
    return guarded([&] {
        if (!host || bytes < 0) throw std::invalid_argument("bad host buffer");
        checked<DMCIProxy>(h)->copy_debug(what, host, bytes);
    });
}

long dcvc_dmci_step_symbol_count(void* h, int step) {
    if (!h || step < 0 || step > 3) return -1;
    return static_cast<DMCIProxy*>(h)->step_count(step);
}

void* dcvc_hts_create(void) { return create<HTSProxy>(); }
void dcvc_hts_destroy(void* h) { delete static_cast<HTSProxy*>(h); }

int dcvc_hts_set_param(void* h, int n, const char** names, const float** data, const int64_t* shapes,
                       const int* ndims, float skip_thres) {
    return guarded([&] { checked<HTSProxy>(h)->set_param(make_dict(n, names, data, shapes, ndims), skip_thres); });
}

int dcvc_hts_share_weights(void* h, void* src) {
    return guarded([&] {
        auto w = checked<HTSProxy>(src)->weights();
        if (!w) throw std::logic_error("source proxy has no parameters");
        checked<HTSProxy>(h)->share_weights(w);
    });
}

int dcvc_hts_prepare(void* h, int Hp, int Wp, void* const* frames, int ld_out) {
    return guarded([&] {
        if (!frames || ld_out < 3) throw std::invalid_argument("bad output buffers");
        f16* f[HTSProxy::kFrames];
        for (int i = 0; i < HTSProxy::kFrames; ++i) {
            if (!frames[i]) throw std::invalid_argument("null output frame");
            f[i] = static_cast<f16*>(frames[i]);
        }
        checked<HTSProxy>(h)->prepare(Hp, Wp, f, ld_out);
    });
}

int dcvc_hts_prepare_batch(void* h, int n_img, int Hp, int Wp, void* const* frames, int ld_out) {
    return guarded([&] {
        if (!frames || ld_out < 3) throw std::invalid_argument("bad output buffers");
        f16* f[HTSProxy::kFrames];
        for (int i = 0; i < HTSProxy::kFrames; ++i) {
            if (!frames[i]) throw std::invalid_argument("null output frame");
            f[i] = static_cast<f16*>(frames[i]);
        }
        checked<HTSProxy>(h)->prepare(Hp, Wp, f, ld_out, n_img);
    });
}

int dcvc_hts_compress_batch(void* h, const void* x, int ldx, int H, int W, const int* qps, int reset_feature_memory,
                            int pad_b, int pad_r, void* stream, const uint8_t** bits, long* nbytes,
                            int* ec_parallel) {
    return guarded([&] {
        if (!x || !qps || !bits || !nbytes || !ec_parallel) throw std::invalid_argument("null argument");
        checked<HTSProxy>(h)->compress_batch(static_cast<const f16*>(x), ldx, H, W, qps, reset_feature_memory != 0,
                                             pad_b, pad_r, S(stream), bits, nbytes, ec_parallel);
    });
}

int dcvc_hts_decompress_batch(void* h, const uint8_t* const* bits, const long* nbytes, const int* qps, int height,
                              int width, const int* ec_parallel, int reset_feature_memory, void* stream) {
    return guarded([&] {
        if (!bits || !nbytes || !qps || !ec_parallel) throw std::invalid_argument("null argument");
        checked<HTSProxy>(h)->decompress_batch(bits, nbytes, qps, height, width, ec_parallel,
                                               reset_feature_memory != 0, S(stream));
    });
}

long dcvc_hts_batch_symbol_count(void* h, int img) {
    if (!h) return -1;
    const HTSProxy* p = static_cast<HTSProxy*>(h);
    return (img >= 0 && img < p->images()) ? p->last_symbol_count(img) : -1;
}

int dcvc_hts_add_ref_feature_from_frame(void* h, const void* frame, int ld, int Hp, int Wp, int apply_adaptor,
                                        void* stream) {
    return guarded([&] {
        if (!frame || ld < 3) throw std::invalid_argument("bad frame");
        checked<HTSProxy>(h)->add_ref_feature_from_frame(static_cast<const f16*>(frame), ld, Hp, Wp, apply_adaptor != 0,
                                                         S(stream));
    });
}

int dcvc_hts_compress(void* h, const void* x, int ldx, int H, int W, int qp, int reset_feature_memory, int pad_b,
                      int pad_r, void* stream, const uint8_t** bits, long* nbytes, int* ec_parallel) {
    return guarded([&] {
        if (!x || !bits || !nbytes || !ec_parallel) throw std::invalid_argument("null argument");
        *ec_parallel = checked<HTSProxy>(h)->compress(static_cast<const f16*>(x), ldx, H, W, qp,
                                                      reset_feature_memory != 0, pad_b, pad_r, S(stream), bits, nbytes);
    });
}

int dcvc_hts_decompress(void* h, const uint8_t* bits, long nbytes, int qp, int height, int width, int ec_parallel,
                        int reset_feature_memory, void* stream) {
    return guarded([&] {
        if (!bits && nbytes > 0) throw std::invalid_argument("null bitstream");
        checked<HTSProxy>(h)->decompress(bits, nbytes, qp, height, width, ec_parallel, reset_feature_memory != 0,
                                         S(stream));
    });
}

int dcvc_hts_set_use_graphs(void* h, int on) {
    return guarded([&] { checked<HTSProxy>(h)->set_use_graphs(on != 0); });
}

long dcvc_hts_device_bytes(void* h) { return h ? (long)static_cast<HTSProxy*>(h)->device_bytes() : -1; }

int dcvc_hts_set_precision(void* h, int prec) {
    return guarded([&] { checked<HTSProxy>(h)->set_precision(prec); });
}

int dcvc_hts_debug_copy(void* h, int what, void* host, long bytes) {
    return guarded([&] {
        if (!host || bytes < 0) throw std::invalid_argument("bad host buffer");
        checked<HTSProxy>(h)->copy_debug(what, host, bytes);
    });
}

long dcvc_hts_last_symbol_count(void* h) { return h ? static_cast<HTSProxy*>(h)->last_symbol_count() : -1; }

void* dcvc_htl_create(void) { return create<HTLProxy>(); }
void dcvc_htl_destroy(void* h) { delete static_cast<HTLProxy*>(h); }

int dcvc_htl_set_param(void* h, int n, const char** names, const float** data, const int64_t* shapes,
                       const int* ndims, float skip_thres) {
    return guarded([&] { checked<HTLProxy>(h)->set_param(make_dict(n, names, data, shapes, ndims), skip_thres); });
}

int dcvc_htl_share_weights(void* h, void* src) {
    return guarded([&] {
        auto w = checked<HTLProxy>(src)->weights();
        if (!w) throw std::logic_error("source proxy has no parameters");
        checked<HTLProxy>(h)->share_weights(w);
    });
}

int dcvc_htl_prepare(void* h, int Hp, int Wp, void* const* frames, int ld_out) {
    return guarded([&] {
        if (!frames || ld_out < 3) throw std::invalid_argument("bad output buffers");
        f16* f[HTLProxy::kFrames];
        for (int i = 0; i < HTLProxy::kFrames; ++i) {
            if (!frames[i]) throw std::invalid_argument("null output frame");
            f[i] = static_cast<f16*>(frames[i]);
        }
        checked<HTLProxy>(h)->prepare(Hp, Wp, f, ld_out);
    });
}

int dcvc_htl_add_ref_feature_from_frame(void* h, const void* frame, int ld, int Hp, int Wp, int apply_adaptor,
                                        void* stream) {
    return guarded([&] {
        if (!frame || ld < 3) throw std::invalid_argument("bad frame");
        checked<HTLProxy>(h)->add_ref_feature_from_frame(static_cast<const f16*>(frame), ld, Hp, Wp, apply_adaptor != 0,
                                                         S(stream));
    });
}

int dcvc_htl_compress(void* h, const void* x, int ldx, int H, int W, int qp, int reset_feature_memory, int pad_b,
                      int pad_r, void* stream, const uint8_t** bits, long* nbytes, int* ec_parallel) {
    return guarded([&] {
        if (!x || !bits || !nbytes || !ec_parallel) throw std::invalid_argument("null argument");
        *ec_parallel = checked<HTLProxy>(h)->compress(static_cast<const f16*>(x), ldx, H, W, qp,
                                                      reset_feature_memory != 0, pad_b, pad_r, S(stream), bits, nbytes);
    });
}

int dcvc_htl_decompress(void* h, const uint8_t* bits, long nbytes, int qp, int height, int width, int ec_parallel,
                        int reset_feature_memory, void* stream) {
    return guarded([&] {
        if (!bits && nbytes > 0) throw std::invalid_argument("null bitstream");
        checked<HTLProxy>(h)->decompress(bits, nbytes, qp, height, width, ec_parallel, reset_feature_memory != 0,
                                         S(stream));
    });
}

int dcvc_htl_set_use_graphs(void* h, int on) {
    return guarded([&] { checked<HTLProxy>(h)->set_use_graphs(on != 0); });
}

int dcvc_htl_set_precision(void* h, int prec) {
    return guarded([&] { checked<HTLProxy>(h)->set_precision(prec); });
}

long dcvc_htl_device_bytes(void* h) { return h ? (long)static_cast<HTLProxy*>(h)->device_bytes() : -1; }

int dcvc_htl_debug_copy(void* h, int what, void* host, long bytes) {
    return guarded([&] {
        if (!host || bytes < 0) throw std::invalid_argument("bad host buffer");
        checked<HTLProxy>(h)->copy_debug(what, host, bytes);
    });
}

long dcvc_htl_step_symbol_count(void* h, int step) {
    if (!h || step < 0 || step > 3) return -1;
    return static_cast<HTLProxy*>(h)->step_count(step);
}

void* dcvc_ld_create(void) { return create<LDProxy>(); }
void dcvc_ld_destroy(void* h) { delete static_cast<LDProxy*>(h); }

int dcvc_ld_set_param(void* h, int n, const char** names, const float** data, const int64_t* shapes,
                      const int* ndims, float skip_thres) {
    return guarded([&] { checked<LDProxy>(h)->set_param(make_dict(n, names, data, shapes, ndims), skip_thres); });
}

int dcvc_ld_share_weights(void* h, void* src) {
    return guarded([&] {
        auto w = checked<LDProxy>(src)->weights();
        if (!w) throw std::logic_error("source proxy has no parameters");
        checked<LDProxy>(h)->share_weights(w);
    });
}

int dcvc_ld_prepare(void* h, int Hp, int Wp, void* x_hat, int ld_out) {
    return guarded([&] {
        if (!x_hat || ld_out < 3) throw std::invalid_argument("bad output buffer");
        checked<LDProxy>(h)->prepare(Hp, Wp, static_cast<f16*>(x_hat), ld_out);
    });
}

int dcvc_ld_add_ref_feature_from_frame(void* h, const void* frame, int ld, int Hp, int Wp, int apply_adaptor,
                                       void* stream) {
    return guarded([&] {
        if (!frame || ld < 3) throw std::invalid_argument("bad frame");
        checked<LDProxy>(h)->add_ref_feature_from_frame(static_cast<const f16*>(frame), ld, Hp, Wp, apply_adaptor != 0,
                                                        S(stream));
    });
}

int dcvc_ld_compress(void* h, const void* x, int ldx, int H, int W, int qp, int reset_feature_memory, int pad_b,
                     int pad_r, void* stream, const uint8_t** bits, long* nbytes, int* ec_parallel) {
    return guarded([&] {
        if (!x || !bits || !nbytes || !ec_parallel) throw std::invalid_argument("null argument");
        *ec_parallel = checked<LDProxy>(h)->compress(static_cast<const f16*>(x), ldx, H, W, qp, reset_feature_memory != 0,
                                                     pad_b, pad_r, S(stream), bits, nbytes);
    });
}

int dcvc_ld_decompress(void* h, const uint8_t* bits, long nbytes, int qp, int height, int width, int ec_parallel,
                       int reset_feature_memory, void* stream) {
    return guarded([&] {
        if (!bits && nbytes > 0) throw std::invalid_argument("null bitstream");
        checked<LDProxy>(h)->decompress(bits, nbytes, qp, height, width, ec_parallel, reset_feature_memory != 0,
                                        S(stream));
    });
}

int dcvc_ld_set_use_graphs(void* h, int on) {
    return guarded([&] { checked<LDProxy>(h)->set_use_graphs(on != 0); });
}

int dcvc_ld_set_precision(void* h, int prec) {
    return guarded([&] { checked<LDProxy>(h)->set_precision(prec); });
}

long dcvc_ld_device_bytes(void* h) { return h ? (long)static_cast<LDProxy*>(h)->device_bytes() : -1; }

int dcvc_ld_debug_copy(void* h, int what, void* host, long bytes) {
    return guarded([&] {
        if (!host || bytes < 0) throw std::invalid_argument("bad host buffer");
        checked<LDProxy>(h)->copy_debug(what, host, bytes);
    });
}

long dcvc_ld_last_symbol_count(void* h) { return h ? static_cast<LDProxy*>(h)->last_symbol_count() : -1; }

int dcvc_trace_begin(int timed) {
    return guarded([&] { trace_set(true, timed != 0); });
}

int dcvc_trace_end(long long* records, int max_records) {
    int n = -1;
    const int st = guarded([&] {
        const std::vector<TraceRecord> t = trace_collect();
        n = (int)t.size();
        if (records)
            for (int i = 0; i < n && i < max_records; ++i) {
                const TraceRecord& r = t[i];
                long long* o = records + 15LL * i;
                o[0] = r.kind;
                o[1] = r.epi;
                o[2] = r.M;
                o[3] = r.N;
                o[4] = r.K;
                o[5] = r.flops;
                o[6] = r.bytes;
                o[7] = r.ns;
                o[8] = r.amode;
                o[9] = r.dmode;
                o[10] = r.prec;
                o[11] = r.cfg;
                o[12] = r.n_img;
                o[13] = r.H;
                o[14] = r.W;
            }
    });
    return st == DCVC_OK ? n : -st;
}

}  // extern "C"
