// DMC HT-S codec proxy: the MI355X counterpart of the reference's DMCHTSProxy
// (src/layers/extensions/inference/dmc_hts_proxy.h:215-355,
//  dmc_hts_proxy.cpp:465-1000).
#pragma once

#include <array>
#include <memory>

#include "rans.h"
#include "runtime.h"

namespace dcvc {

// Weights of the HT-S model (one copy can serve many stream contexts).
struct HTSWeights {
    DevMem mem;
    DCBSeq fa_i, fa_m, fe, enc, pf, sp, dec;
    ConvS2 enc_down;
    DCB he0, hd2, spa[3];
    RBS2 he1, he2, tpe;
    RBU hd0, hd1;
    Conv1x1 pf_out, reduction, sp_out;
    Subpel dec_up;
    DCB rh1[4], rh2[8][3];
    Conv1x1 rh_head[8];
    f16 *q_encoder = nullptr, *q_decoder = nullptr, *q_feature = nullptr;
    float* thresholds = nullptr;
    float skip_thres = 0.f;
    // entropy tables: 0 = z (bit estimator), 1 = y (gaussian)
    std::vector<int32_t> cdf[2], cdf_len[2];
    int cdf_row[2] = {0, 0};
    void load(const HostDict& d, float skip);
};

// One proxy codes n_img streams of one resolution as a batch (n_img = 1 is
// the reference's proxy): every layer is one launch over the n_img images,
// each image keeps its own QP, temporal state and bitstream, and its bits
// equal those of a proxy of its own.
class HTSProxy {
public:
    static constexpr int kFrames = 8;
    static constexpr int kMaxImages = DeviceInts::kMax;
    HTSProxy();
    ~HTSProxy();
    void set_param(const HostDict& d, float skip_thres);
    void share_weights(const std::shared_ptr<HTSWeights>& w);
    std::shared_ptr<HTSWeights> weights() const { return m_w; }
    // Allocate buffers for n_img pictures of padded size Hp x Wp; `frames`
    // are 8 caller-owned NHWC fp16 [n_img, Hp, Wp, 3] reconstruction outputs
    // (pixel stride ld_out).
    void prepare(int Hp, int Wp, f16* const frames[kFrames], int ld_out, int n_img = 1);
    int images() const { return m_nimg; }
    // frame: n_img NHWC pictures [n_img, Hp, Wp] (pixel stride ld)
    void add_ref_feature_from_frame(const f16* frame, int ld, int Hp, int Wp, bool apply_adaptor, hipStream_t s);
    // one image (n_img == 1)
    int compress(const f16* x, int ldx, int H, int W, int qp, bool reset, int pad_b, int pad_r, hipStream_t s,
                 const uint8_t** bits, long* nbytes);
    void decompress(const uint8_t* bits, long nbytes, int qp, int height, int width, int ec_parallel, bool reset,
                    hipStream_t s);
    // all n_img images: x is [n_img, H, W] pixels of 24 channels; qps, bits,
    // nbytes, ec_parallel have n_img entries (bits stay valid until the next
    // compress)
    void compress_batch(const f16* x, int ldx, int H, int W, const int* qps, bool reset, int pad_b, int pad_r,
                        hipStream_t s, const uint8_t** bits, long* nbytes, int* ec_parallel);
    void decompress_batch(const uint8_t* const* bits, const long* nbytes, const int* qps, int height, int width,
                          const int* ec_parallel, bool reset, hipStream_t s);
    void set_use_graphs(bool on) { m_use_graphs = on; reset_graphs(); }
    // GemmPrec of every 1x1-conv product (precision form); recaptures the graphs
    void set_precision(int prec) {
        if (prec < PREC_F16 || prec > PREC_FP8) throw std::invalid_argument("precision must be 0, 1 or 2");
        if (prec != PREC_F16 && m_w) m_w->mem.ensure_low_precision();
        m_prec = prec;
        reset_graphs();
    }
    size_t device_bytes() const { return m_buf.bytes(); }
    // diagnostics for tests: last compacted symbol count and y / symbols on host
    long last_symbol_count(int img = 0) const {
        return (img >= 0 && img < (int)m_last_counts.size()) ? m_last_counts[img] : 0;
    }
    void copy_debug(int what, void* host, long bytes);

private:
    void reset_graphs();
    void enter(hipStream_t caller);
    void leave(hipStream_t caller);
    void enc_stage0();      // encoder, hyper encoder and decoder (main stream)
    void enc_temporal();    // temporal prior from memory * q_feature (side stream)
    void enc_stage0b();     // prior fusion, spatial prior steps, compaction
    void enc_stage1(bool reset);
    void dec_stage0(bool memory_has_value);  // feature adaptor -> memory
    void dec_stage0b();                       // temporal prior from the memory
    void dec_stage1();
    void dec_stage2();
    void dec_stage3();
    void hyper_and_prior();
    void hyper_decode();
    void prior_fusion();
    // recon head frames [first, last) on `e`'s stream with the given buffers
    void recon_frames(const Exec& e, int first, int last, f16* common, f16* ra, f16* head);
    void spatial_means(int step);
    void decoder_and_recon(bool recon_all, bool reset_for_encoder);
    Exec exec() { return Exec{m_stream, m_d_qp, &m_scratch, m_prec}; }
    // the decoder's side stream: feature extraction during the prior, and
    // half of the recon head (independent work, own scratch)
    Exec exec_side() { return Exec{m_side, m_d_qp, &m_scratch2, m_prec}; }
    void fork_side();
    void join_side();
    void set_coder_cdfs();
    void set_qps(const int* qps);
    void check_qps(const int* qps) const;

    std::shared_ptr<HTSWeights> m_w;
    int m_nimg = 1;
    RansEncoder m_enc;                                  // images coded one after another
    std::vector<std::vector<uint8_t>> m_bits;           // per image: the last compress's bitstream
    std::vector<std::unique_ptr<RansDecoder>> m_decs;  // per image: z and y decode of one call
    hipStream_t m_stream = nullptr, m_copy = nullptr, m_side = nullptr;
    hipEvent_t m_ev_in = nullptr, m_ev_out = nullptr, m_ev_sym = nullptr, m_ev_fork = nullptr, m_ev_join = nullptr;
    bool m_use_graphs = true;
    int m_prec = PREC_F16;
    Graph m_g_enc0, m_g_enc0b, m_g_enc_tp, m_g_enc1[2], m_g_dec0[2], m_g_dec0b, m_g_dec1, m_g_dec2, m_g_dec3, m_g_rh[2];

    // geometry
    int m_Hp = 0, m_Wp = 0, m_H8 = 0, m_W8 = 0, m_H16 = 0, m_W16 = 0, m_H16p = 0, m_W16p = 0;
    int m_H64 = 0, m_W64 = 0;
    bool m_pad4 = false;
    long m_ny = 0, m_nz = 0;  // per image
    int m_offs_stride = 0;    // compaction block-offset slots per image

    // device buffers
    DevMem m_buf;
    Scratch m_scratch, m_scratch2;
    int* m_d_qp = nullptr;
    int* m_d_count = nullptr;
    int* m_d_offs = nullptr;
    f16 *m_cat_enc = nullptr, *m_cat_fam = nullptr, *m_feature_i = nullptr, *m_tin = nullptr, *m_enc_out = nullptr;
    f16 *m_y = nullptr, *m_y_pad = nullptr, *m_he_a = nullptr, *m_he_b = nullptr, *m_z = nullptr, *m_z_hat = nullptr;
    f16 *m_hd_a = nullptr, *m_hd_b = nullptr, *m_cat_pf = nullptr, *m_pf_a = nullptr, *m_common = nullptr;
    f16 *m_cat_sp = nullptr, *m_sp_a = nullptr, *m_means = nullptr;
    f16 *m_rh_common = nullptr, *m_rh_a = nullptr, *m_rh_head = nullptr;
    f16 *m_rh_common2 = nullptr, *m_rh_a2 = nullptr, *m_rh_head2 = nullptr;
    int8_t *m_z_i8 = nullptr, *m_yq = nullptr, *m_dec_c = nullptr;
    int16_t *m_sym = nullptr, *m_symc = nullptr;
    uint8_t *m_cond = nullptr, *m_idx = nullptr, *m_idxc = nullptr;
    f16* m_frames[kFrames] = {};
    int m_ld_out = 3;

    // pinned host staging
    int* m_h_count = nullptr;
    int16_t* m_h_sym = nullptr;
    uint8_t* m_h_idx = nullptr;
    int8_t* m_h_dec = nullptr;
    int8_t* m_h_z = nullptr;

    bool m_memory_ready = false;     // encoder side: memory/ctx hold the adaptor output
    bool m_memory_has_value = false; // decoder side (dmc_hts_proxy.cpp:617, 707)
    bool m_has_ref = false;  // add_ref_feature_from_frame ran since prepare(): m_feature_i holds a frame
    std::vector<long> m_last_counts;
};

}  // namespace dcvc
