// DMC HT-L codec proxy: the MI355X counterpart of the reference's DMCHTLProxy
// (src/layers/extensions/inference/dmc_htl_proxy.h, dmc_htl_proxy.cpp:476-1178).
#pragma once

#include <array>
#include <memory>

#include "rans.h"
#include "runtime.h"

namespace dcvc {

// Weights of the HT-L model (one copy can serve many stream contexts).
struct HTLWeights {
    DevMem mem;
    DCBSeq fa_i, fa_m, fe, enc, pf, sp, dec;
    ConvS2 enc_down;
    DCB he0, hd2, spa[3];
    RBS2 he1, he2, tpe;
    RBU hd0, hd1;
    Conv1x1 pf_out, reduction, sp_out;
    Subpel dec_up;   // 3x3 pad 1, with bias
    DCBSeq rh[8];    // per frame: 5 blocks 512 -> 256
    Conv1x1 rh_head[8];
    f16 *q_encoder = nullptr, *q_decoder = nullptr, *q_feature = nullptr;
    float* thresholds = nullptr;
    float skip_thres = 0.f;
    // entropy tables: 0 = z (bit estimator), 1 = y (gaussian)
    std::vector<int32_t> cdf[2], cdf_len[2];
    int cdf_row[2] = {0, 0};
    void load(const HostDict& d, float skip);
};

class HTLProxy {
public:
    static constexpr int kFrames = 8;
    HTLProxy();
    ~HTLProxy();
    void set_param(const HostDict& d, float skip_thres);
    void share_weights(const std::shared_ptr<HTLWeights>& w);
    std::shared_ptr<HTLWeights> weights() const { return m_w; }
    // Allocate buffers for a padded Hp x Wp picture; `frames` are 8 caller-owned
    // NHWC fp16 [Hp, Wp, 3] reconstruction outputs (pixel stride ld_out).
    void prepare(int Hp, int Wp, f16* const frames[kFrames], int ld_out);
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
    long step_count(int step) const { return m_counts[step]; }
    void copy_debug(int what, void* host, long bytes);

private:
    void reset_graphs();
    void enter(hipStream_t caller);
    void leave(hipStream_t caller);
    void enc_stage0();      // encoder, hyper encoder and decoder (main stream)
    void enc_temporal();    // temporal prior from memory * q_feature (side stream)
    void enc_stage0b();     // prior fusion, spatial prior steps, compaction
    void enc_stage1(bool reset);
    void dec_stage0(bool memory_has_value);
    void dec_stage1();
    void dec_restore(int step);
    void dec_index(int step);
    void decode_step_on_host(int step);
    void hyper_and_prior();
    void hyper_decode();
    void prior_fusion();
    void spatial_prior(int step);
    void decoder_and_recon(bool recon_all, bool reset_for_encoder);
    // recon head frames [first, last) on `e`'s stream with the given buffers
    void recon_frames(const Exec& e, int first, int last, f16* ra, f16* head);
    void fe_side();
    void fork_side();
    void join_side();
    void feature_adaptor(bool use_m);
    Exec exec() { return Exec{m_stream, m_d_qp, &m_scratch, m_prec}; }
    // the decoder's side stream: feature extraction during the prior and
    // entropy round trips, half of the recon head (own scratch)
    Exec exec_side() { return Exec{m_side, m_d_qp, &m_scratch2, m_prec}; }
    // scales / means of checkerboard step k: the common params for step 0,
    // the spatial prior output after
    const f16* step_scales(int step) const { return step == 0 ? m_common + 256 : m_sp_o; }
    const f16* step_means(int step) const { return step == 0 ? m_common + 512 : m_sp_o + 256; }
    int step_ld(int step) const { return step == 0 ? 768 : 512; }

    std::shared_ptr<HTLWeights> m_w;
    RansEncoder m_enc;
    RansDecoder m_dec;
    hipStream_t m_stream = nullptr, m_copy = nullptr, m_side = nullptr;
    hipEvent_t m_ev_in = nullptr, m_ev_out = nullptr, m_ev_sym = nullptr, m_ev_fork = nullptr, m_ev_join = nullptr;
    bool m_use_graphs = true;
    int m_prec = PREC_F16;
    Graph m_g_enc0, m_g_enc0b, m_g_enc_tp, m_g_enc1[2], m_g_dec0[2], m_g_dec1, m_g_dec[4], m_g_fe, m_g_rh[2];

    // geometry
    int m_Hp = 0, m_Wp = 0, m_H8 = 0, m_W8 = 0, m_H16 = 0, m_W16 = 0, m_H16p = 0, m_W16p = 0;
    int m_H64 = 0, m_W64 = 0;
    bool m_pad4 = false;
    long m_nq = 0, m_nz = 0;   // quarter of y, z elements

    // device buffers
    DevMem m_buf;
    Scratch m_scratch, m_scratch2;
    int* m_d_qp = nullptr;
    int* m_d_counts = nullptr;  // 4 step counts (encoder) / 1 (decoder)
    int* m_d_offs = nullptr;
    f16 *m_cat_enc = nullptr, *m_cat_fam = nullptr, *m_feature_i = nullptr, *m_tin = nullptr, *m_enc_out = nullptr;
    f16 *m_y = nullptr, *m_y_pad = nullptr, *m_he_a = nullptr, *m_he_b = nullptr, *m_z = nullptr, *m_z_hat = nullptr;
    f16 *m_hd_a = nullptr, *m_hd_b = nullptr, *m_cat_pf = nullptr, *m_pf_a = nullptr, *m_common = nullptr;
    f16 *m_cat_sp = nullptr, *m_sp_a = nullptr, *m_sp_o = nullptr, *m_y_hat = nullptr;
    f16 *m_rh_a = nullptr, *m_rh_head = nullptr, *m_rh_a2 = nullptr, *m_rh_head2 = nullptr;
    int8_t *m_z_i8 = nullptr, *m_yq = nullptr, *m_dec_c = nullptr;
    std::array<int16_t*, 4> m_sym{}, m_symc{};
    std::array<uint8_t*, 4> m_condq{};
    uint8_t *m_cond = nullptr, *m_idx = nullptr, *m_idxc = nullptr;
    f16* m_frames[kFrames] = {};
    int m_ld_out = 3;

    // pinned host staging
    int* m_h_counts = nullptr;
    std::array<int16_t*, 4> m_h_sym{};
    uint8_t* m_h_idx = nullptr;
    int8_t* m_h_dec = nullptr;
    int8_t* m_h_z = nullptr;

    bool m_memory_ready = false;     // encoder side: memory/ctx hold the adaptor output
    bool m_memory_has_value = false; // decoder side (dmc_htl_proxy.cpp:591, 912)
    bool m_has_ref = false;  // add_ref_feature_from_frame ran since prepare(): m_feature_i holds a frame
    long m_counts[4] = {0, 0, 0, 0};
    int m_qp = 0;
};

}  // namespace dcvc
