// DMC low-delay codec proxy: the MI355X counterpart of the reference's
// DMCLDProxy (src/layers/extensions/inference/dmc_ld_proxy.h,
// dmc_ld_proxy.cpp:380-862).  One frame per call.
#pragma once

#include <memory>

#include "rans.h"
#include "runtime.h"

namespace dcvc {

struct LDWeights {
    DevMem mem;
    DCBSeq fa_i, fa_m, fe, enc1, pf, sp, dec1, rh;
    DCB enc2, he0, hd2;
    ConvS2 enc_down;
    RBS2 he1, he2, tpe;
    RBU hd0, hd1;
    Conv1x1 pf_out, sp_out, dec2, rh_head;
    Subpel dec_up;
    f16 *q_encoder = nullptr, *q_decoder = nullptr, *q_feature = nullptr;
    float* thresholds = nullptr;
    float skip_thres = 0.f;
    std::vector<int32_t> cdf[2], cdf_len[2];
    int cdf_row[2] = {0, 0};
    void load(const HostDict& d, float skip);
};

class LDProxy {
public:
    LDProxy();
    ~LDProxy();
    void set_param(const HostDict& d, float skip_thres);
    void share_weights(const std::shared_ptr<LDWeights>& w);
    std::shared_ptr<LDWeights> weights() const { return m_w; }
    // buffers for a padded Hp x Wp picture; x_hat: caller-owned NHWC fp16 [Hp, Wp, 3]
    void prepare(int Hp, int Wp, f16* x_hat, int ld_out);
    void add_ref_feature_from_frame(const f16* frame, int ld, int Hp, int Wp, bool apply_adaptor, hipStream_t s);
    int compress(const f16* x, int ldx, int H, int W, int qp, bool reset, int pad_b, int pad_r, hipStream_t s,
                 const uint8_t** bits, long* nbytes);
    void decompress(const uint8_t* bits, long nbytes, int qp, int height, int width, int ec_parallel, bool reset,
                    hipStream_t s);
    void set_use_graphs(bool on) { m_use_graphs = on; reset_graphs(); }
    void set_precision(int prec) {
        if (prec < PREC_F16 || prec > PREC_FP8) throw std::invalid_argument("precision must be 0, 1 or 2");
        if (prec != PREC_F16 && m_w) m_w->mem.ensure_low_precision();
        m_prec = prec;
        reset_graphs();
    }
    size_t device_bytes() const { return m_buf.bytes(); }
    long last_symbol_count() const { return m_last_count; }
    void copy_debug(int what, void* host, long bytes);

private:
    void reset_graphs();
    void enter(hipStream_t caller);
    void leave(hipStream_t caller);
    void feature_adaptor(bool use_m);
    void context_from_memory();
    void hyper_and_prior();
    void spatial_means();
    void decoder_and_recon(bool recon, bool reset_for_encoder);
    void enc_stage0();
    void enc_stage1(bool reset);
    void dec_stage0(bool memory_has_value);
    void dec_stage1();
    void dec_stage2();
    void dec_stage3();
    Exec exec() { return Exec{m_stream, m_d_qp, &m_scratch, m_prec}; }
    // the decoder's side stream: feature extraction while the host decodes
    // z and the prior runs (independent work, own scratch)
    Exec exec_side() { return Exec{m_side, m_d_qp, &m_scratch2, m_prec}; }

    std::shared_ptr<LDWeights> m_w;
    RansEncoder m_enc;
    RansDecoder m_dec;
    hipStream_t m_stream = nullptr, m_copy = nullptr, m_side = nullptr;
    hipEvent_t m_ev_in = nullptr, m_ev_out = nullptr, m_ev_sym = nullptr, m_ev_fork = nullptr, m_ev_join = nullptr;
    bool m_use_graphs = true;
    int m_prec = PREC_F16;
    Graph m_g_enc0, m_g_enc1[2], m_g_dec0[2], m_g_dec1, m_g_dec2, m_g_dec3;

    int m_Hp = 0, m_Wp = 0, m_H8 = 0, m_W8 = 0, m_H16 = 0, m_W16 = 0, m_H16p = 0, m_W16p = 0;
    int m_H64 = 0, m_W64 = 0;
    bool m_pad4 = false;
    long m_ny = 0, m_nz = 0;

    DevMem m_buf;
    Scratch m_scratch, m_scratch2;
    int* m_d_qp = nullptr;
    int* m_d_count = nullptr;
    int* m_d_offs = nullptr;
    f16 *m_cat_dec = nullptr, *m_cat_fam = nullptr, *m_feature_i = nullptr, *m_enc_out = nullptr;
    f16 *m_y = nullptr, *m_y_pad = nullptr, *m_he_a = nullptr, *m_he_b = nullptr, *m_z = nullptr, *m_z_hat = nullptr;
    f16 *m_hd_a = nullptr, *m_hd_b = nullptr, *m_cat_pf = nullptr, *m_pf_a = nullptr;
    f16 *m_cat_sp = nullptr, *m_sp_a = nullptr, *m_means1 = nullptr, *m_rh_a = nullptr;
    int8_t *m_z_i8 = nullptr, *m_yq = nullptr, *m_dec_c = nullptr;
    int16_t *m_sym = nullptr, *m_symc = nullptr;
    uint8_t *m_cond = nullptr, *m_idx = nullptr, *m_idxc = nullptr;
    f16* m_x_hat = nullptr;
    int m_ld_out = 3;

    int* m_h_count = nullptr;
    int16_t* m_h_sym = nullptr;
    uint8_t* m_h_idx = nullptr;
    int8_t* m_h_dec = nullptr;
    int8_t* m_h_z = nullptr;

    bool m_memory_ready = false;
    bool m_memory_has_value = false;  // decoder side (dmc_ld_proxy.cpp:417, 590)
    bool m_has_ref = false;  // add_ref_feature_from_frame ran since prepare(): m_feature_i holds a frame
    long m_last_count = 0;
    int m_qp = 0;
};

}  // namespace dcvc
