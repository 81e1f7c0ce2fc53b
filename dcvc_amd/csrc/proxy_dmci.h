// DMCI intra codec proxy: the MI355X counterpart of the reference's DMCIProxy
// (src/layers/extensions/inference/dmci_proxy.h:132-254, dmci_proxy.cpp:269-882).
#pragma once

#include <memory>

#include "rans.h"
#include "runtime.h"

namespace dcvc {

struct DMCIWeights {
    DevMem mem;
    DCB enc_1, he0, hd2, spa[3], dec_2;
    DCBSeq enc_2, pf, sp, dec;
    ConvS2 enc_down;
    RBS2 he1, he2;
    RBU hd0, hd1, dec_up;
    Conv1x1 pf_out, reduction, sp_out;
    f16 *q_scale_enc = nullptr, *q_scale_dec = nullptr, *q_scale_y_enc = nullptr, *q_scale_y_dec = nullptr;
    float* thresholds = nullptr;
    float skip_thres = 0.f;
    std::vector<int32_t> cdf[2], cdf_len[2];
    int cdf_row[2] = {0, 0};
    void load(const HostDict& d, float skip);
};

class DMCIProxy {
public:
    DMCIProxy();
    ~DMCIProxy();
    void set_param(const HostDict& d, float skip_thres);
    // buffers for a padded Hp x Wp picture; x_hat: caller-owned NHWC fp16 [Hp, Wp, 3]
    void prepare(int Hp, int Wp, f16* x_hat, int ld_out);
    int compress(const f16* x, int ldx, int H, int W, int qp, int pad_b, int pad_r, hipStream_t s,
                 const uint8_t** bits, long* nbytes);
    void decompress(const uint8_t* bits, long nbytes, int qp, int height, int width, int ec_parallel, hipStream_t s);
    void set_use_graphs(bool on) { m_use_graphs = on; reset_graphs(); }
    // GemmPrec of every 1x1-conv product (precision form); recaptures the graphs
    void set_precision(int prec) {
        if (prec < PREC_F16 || prec > PREC_FP8) throw std::invalid_argument("precision must be 0, 1 or 2");
        if (prec != PREC_F16 && m_w) m_w->mem.ensure_low_precision();
        m_prec = prec;
        reset_graphs();
    }
    long step_count(int step) const { return m_counts[step]; }
    void copy_debug(int what, void* host, long bytes);

private:
    void reset_graphs();
    void enter(hipStream_t caller);
    void leave(hipStream_t caller);
    void enc_stage0();
    void enc_stage1();
    void prior_from_z();
    void spatial_prior(int step);
    void dec_index(int step);
    void dec_restore(int step);
    void decode_step_on_host(int step);
    Exec exec() { return Exec{m_stream, m_d_qp, &m_scratch, m_prec}; }

    std::unique_ptr<DMCIWeights> m_w;
    RansEncoder m_enc;
    RansDecoder m_dec;
    hipStream_t m_stream = nullptr, m_copy = nullptr;
    hipEvent_t m_ev_in = nullptr, m_ev_out = nullptr, m_ev_sym = nullptr;
    bool m_use_graphs = true;
    int m_prec = PREC_F16;
    Graph m_g_enc0, m_g_enc1, m_g_dec[5];

    int m_Hp = 0, m_Wp = 0, m_H8 = 0, m_W8 = 0, m_H16 = 0, m_W16 = 0, m_H16p = 0, m_W16p = 0, m_H64 = 0, m_W64 = 0;
    bool m_pad4 = false;
    long m_nq = 0, m_nz = 0;  // quarter symbols per step, z symbols

    DevMem m_buf;
    Scratch m_scratch;
    int *m_d_qp = nullptr, *m_d_count = nullptr, *m_d_offs = nullptr;
    f16 *m_x_unsh = nullptr, *m_e_a = nullptr, *m_y = nullptr, *m_y_pad = nullptr, *m_he_a = nullptr, *m_he_b = nullptr;
    f16 *m_z = nullptr, *m_z_hat = nullptr, *m_hd_a = nullptr, *m_hd_b = nullptr, *m_hd_c = nullptr, *m_pf_in = nullptr;
    f16 *m_pf_a = nullptr, *m_params = nullptr, *m_cat_sp = nullptr, *m_sp_a = nullptr, *m_sp_o = nullptr;
    f16 *m_y_hat = nullptr, *m_d_a = nullptr, *m_d_b = nullptr;
    int8_t *m_z_i8 = nullptr, *m_yq = nullptr, *m_dec_c = nullptr;
    int16_t *m_symq[4] = {}, *m_symc[4] = {};
    uint8_t* m_condq[4] = {};
    int* m_d_counts = nullptr;
    uint8_t *m_cond = nullptr, *m_idx = nullptr, *m_idxc = nullptr;
    f16* m_x_hat = nullptr;
    int m_ld_out = 3;

    int* m_h_counts = nullptr;
    int16_t* m_h_sym[4] = {};
    uint8_t* m_h_idx = nullptr;
    int8_t* m_h_dec = nullptr;
    int8_t* m_h_z = nullptr;
    long m_counts[4] = {0, 0, 0, 0};
    int m_qp = 0;
};

}  // namespace dcvc
