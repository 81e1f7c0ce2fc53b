// DMC HT-S codec proxy (see proxy_hts.h).
//
// Buffer layout mirrors the reference's concatenation views
// (dmc_hts_proxy.cpp:932-958) so torch.cat never runs:
//   cat_enc [P8, 2048] = x_unshuffled[0:1536) | ctx[1536:2048);
//                        the decoder input is [1024:2048) = up(y_hat) | ctx
//   cat_fam [P8, 1024] = memory[0:512) | feature_p[512:1024)
//   cat_pf  [P16, 768] = hyper params[0:256) | temporal params[256:768)
//   cat_sp  [P16, 512] = y_hat[0:256) | reduced params[256:512)
// Stage graphs follow the reference's lambdas: enc_0 / enc_1[reset] on the
// encoder (dmc_hts_proxy.cpp:516-576) and dec_0[memory] / dec_1 / dec_2 /
// dec_3 on the decoder (:611-704).  The QP is a device scalar read by the
// kernels, so one graph serves all 64 QPs.
// A proxy prepared for n_img pictures holds them one after another in every
// buffer (the rows of picture b follow those of picture b - 1), with one
// device QP per picture; only the compaction, the host coding and the
// symbol buffers are handled picture by picture.
#include "proxy_hts.h"

#include <cstring>

namespace dcvc {

static std::vector<int32_t> to_i32(const HostTensor& t) {
    std::vector<int32_t> v(t.data.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = (int32_t)t.data[i];
    return v;
}

static f16* upload_qtable(DevMem& mem, const HostDict& d, const std::string& key, int C) {
    const HostTensor& t = need(d, key);
    if (t.numel() != 64L * C) throw std::invalid_argument(key + ": expected [64, C]");
    return upload_f16(mem, t.data);
}

void HTSWeights::load(const HostDict& d, float skip) {
    mem.free_all();
    fa_i.load(mem, d, "feature_adaptor_i.conv.", 4);
    fa_m.load(mem, d, "feature_adaptor_m.conv.", 6);
    fe.load(mem, d, "feature_extractor.conv.", 5);
    enc.load(mem, d, "encoder.conv1.", 6);
    enc_down.load(mem, d, "encoder.down.", 3);
    he0.load(mem, d, "hyper_encoder.conv.0.");
    he1.load(mem, d, "hyper_encoder.conv.1.", false);
    he2.load(mem, d, "hyper_encoder.conv.2.", false);
    hd0.load(mem, d, "hyper_decoder.conv.0.", false);
    hd1.load(mem, d, "hyper_decoder.conv.1.", false);
    hd2.load(mem, d, "hyper_decoder.conv.2.");
    tpe.load(mem, d, "temporal_prior_encoder.conv.", false);
    pf.load(mem, d, "y_prior_fusion.conv.", 3);
    pf_out.load(mem, d, "y_prior_fusion.conv.3.");
    reduction.load(mem, d, "y_spatial_prior_reduction.");
    for (int i = 0; i < 3; ++i) spa[i].load(mem, d, "y_spatial_prior_adaptor_" + std::to_string(i + 1) + ".");
    sp.load(mem, d, "y_spatial_prior.conv.", 3);
    sp_out.load(mem, d, "y_spatial_prior.conv.3.");
    dec_up.load(mem, d, "decoder.up.");
    dec.load(mem, d, "decoder.conv1.", 7);
    for (int i = 0; i < 4; ++i) rh1[i].load(mem, d, "recon_head.conv1." + std::to_string(i) + ".0.");
    for (int i = 0; i < 8; ++i) {
        for (int j = 0; j < 3; ++j)
            rh2[i][j].load(mem, d, "recon_head.conv2." + std::to_string(i) + "." + std::to_string(j) + ".");
        rh_head[i].load(mem, d, "recon_head.conv2." + std::to_string(i) + ".3.");
    }
    q_encoder = upload_qtable(mem, d, "q_encoder", 512);
    q_decoder = upload_qtable(mem, d, "q_decoder", 512);
    q_feature = upload_qtable(mem, d, "q_feature", 512);
    float th[128];
    host_scale_index_thresholds(th);
    thresholds = upload_f32(mem, std::vector<float>(th, th + 128));
    skip_thres = skip;
    const char* names[2] = {"bit_estimator_z.", "gaussian_encoder."};
    for (int i = 0; i < 2; ++i) {
        const HostTensor& c = need(d, std::string(names[i]) + "quantized_cdf");
        const HostTensor& l = need(d, std::string(names[i]) + "cdf_length");
        cdf[i] = to_i32(c);
        cdf_len[i] = to_i32(l);
        cdf_row[i] = (int)(c.numel() / l.numel());
    }
    hip_check(hipDeviceSynchronize(), "weight conversion");
}

// ---------------------------------------------------------------------------
HTSProxy::HTSProxy() {
    hip_check(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking), "stream");
    hip_check(hipStreamCreateWithFlags(&m_copy, hipStreamNonBlocking), "stream");
    hip_check(hipStreamCreateWithFlags(&m_side, hipStreamNonBlocking), "stream");
    hip_check(hipEventCreateWithFlags(&m_ev_in, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&m_ev_out, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&m_ev_sym, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&m_ev_fork, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&m_ev_join, hipEventDisableTiming), "event");
}

HTSProxy::~HTSProxy() {
    if (m_stream) hipStreamSynchronize(m_stream);
    if (m_side) hipStreamSynchronize(m_side);
    reset_graphs();
    hipHostFree(m_h_count);
    hipHostFree(m_h_sym);
    hipHostFree(m_h_idx);
    hipHostFree(m_h_dec);
    hipHostFree(m_h_z);
    m_buf.free_all();
    hipEventDestroy(m_ev_in);
    hipEventDestroy(m_ev_out);
    hipEventDestroy(m_ev_sym);
    hipEventDestroy(m_ev_fork);
    hipEventDestroy(m_ev_join);
    hipStreamDestroy(m_side);
    hipStreamDestroy(m_copy);
    hipStreamDestroy(m_stream);
}

void HTSProxy::set_param(const HostDict& d, float skip_thres) {
    auto w = std::make_shared<HTSWeights>();
    w->load(d, skip_thres);
    share_weights(w);
}

void HTSProxy::share_weights(const std::shared_ptr<HTSWeights>& w) {
    m_w = w;
    if (m_prec != PREC_F16) m_w->mem.ensure_low_precision();
    set_coder_cdfs();
    reset_graphs();
}

void HTSProxy::set_coder_cdfs() {
    if (!m_w) return;
    const HTSWeights& w = *m_w;
    for (int i = 0; i < 2; ++i) {
        m_enc.set_cdf(w.cdf[i].data(), (int)w.cdf_len[i].size(), w.cdf_row[i], w.cdf_len[i].data(), i);
        for (auto& d : m_decs)
            d->set_cdf(w.cdf[i].data(), (int)w.cdf_len[i].size(), w.cdf_row[i], w.cdf_len[i].data(), i);
    }
}

void HTSProxy::reset_graphs() {
    // a graph may still be running on the proxy stream (the encoder returns
    // before its context update finishes)
    if (m_stream) hipStreamSynchronize(m_stream);
    if (m_side) hipStreamSynchronize(m_side);
    m_g_enc0.reset();
    m_g_enc0b.reset();
    m_g_enc_tp.reset();
    for (auto& g : m_g_enc1) g.reset();
    for (auto& g : m_g_dec0) g.reset();
    m_g_dec0b.reset();
    m_g_dec1.reset();
    m_g_dec2.reset();
    m_g_dec3.reset();
    for (auto& g : m_g_rh) g.reset();
}

void HTSProxy::prepare(int Hp, int Wp, f16* const frames[kFrames], int ld_out, int n_img) {
    if (Hp <= 0 || Wp <= 0 || Hp % 16 || Wp % 16) throw std::invalid_argument("picture size must be padded to 16");
    if (n_img < 1 || n_img > kMaxImages) throw std::invalid_argument("1 to 16 images per batch");
    bool same = (Hp == m_Hp && Wp == m_Wp && ld_out == m_ld_out && n_img == m_nimg);
    for (int i = 0; i < kFrames && same; ++i) same = (frames[i] == m_frames[i]);
    if (same && m_cat_enc) return;
    hip_check(hipStreamSynchronize(m_stream), "sync");
    reset_graphs();
    m_buf.free_all();
    for (void* p : {(void*)m_h_count, (void*)m_h_sym, (void*)m_h_idx, (void*)m_h_dec, (void*)m_h_z}) hipHostFree(p);
    m_nimg = n_img;
    m_Hp = Hp;
    m_Wp = Wp;
    m_H8 = Hp / 8;
    m_W8 = Wp / 8;
    m_H16 = Hp / 16;
    m_W16 = Wp / 16;
    m_H16p = (m_H16 + 3) / 4 * 4;
    m_W16p = (m_W16 + 3) / 4 * 4;
    m_pad4 = (m_H16p != m_H16 || m_W16p != m_W16);
    m_H64 = m_H16p / 4;
    m_W64 = m_W16p / 4;
    // pixel counts of the whole batch: images are stacked in every buffer
    const long n = n_img;
    const long P8 = n * m_H8 * m_W8, P16 = n * m_H16 * m_W16, P16p = n * m_H16p * m_W16p;
    const long P32p = P16p / 4, P64 = n * m_H64 * m_W64;
    m_ny = (long)m_H16 * m_W16 * 256;
    m_nz = (long)m_H64 * m_W64 * 128;
    m_offs_stride = compact_blocks(m_ny) + 1;
    DevMem& M = m_buf;
    m_scratch.halves = (size_t)P8 * 768;
    m_scratch.t1 = M.alloc_n<f16>(m_scratch.halves);
    m_scratch.t2 = M.alloc_n<f16>(m_scratch.halves);
    m_scratch.so = M.alloc_n<f16>(m_scratch.halves);
    m_scratch.x1 = M.alloc_n<f16>(m_scratch.halves);
    m_scratch.x2 = M.alloc_n<f16>(m_scratch.halves);
    m_scratch2.halves = m_scratch.halves;
    m_scratch2.t1 = M.alloc_n<f16>(m_scratch2.halves);
    m_scratch2.t2 = M.alloc_n<f16>(m_scratch2.halves);
    m_scratch2.so = M.alloc_n<f16>(m_scratch2.halves);
    m_scratch2.x1 = M.alloc_n<f16>(m_scratch2.halves);
    m_scratch2.x2 = M.alloc_n<f16>(m_scratch2.halves);
    m_d_qp = M.alloc_n<int>(n);
    m_d_count = M.alloc_n<int>(n);
    m_d_offs = M.alloc_n<int>(n * m_offs_stride);
    m_cat_enc = M.alloc_n<f16>(P8 * 2048);
    m_cat_fam = M.alloc_n<f16>(P8 * 1024);
    m_feature_i = M.alloc_n<f16>(P8 * 192);
    m_tin = M.alloc_n<f16>(P8 * 512);
    m_enc_out = M.alloc_n<f16>(P8 * 512);
    m_y = M.alloc_n<f16>(P16 * 256);
    m_y_pad = m_pad4 ? M.alloc_n<f16>(P16p * 256) : m_y;
    m_he_a = M.alloc_n<f16>(P16p * 256);
    m_he_b = M.alloc_n<f16>(P32p * 256);
    m_z = M.alloc_n<f16>(P64 * 128);
    m_z_hat = M.alloc_n<f16>(P64 * 128);
    m_hd_a = M.alloc_n<f16>(P32p * 256);
    m_hd_b = M.alloc_n<f16>(P16p * 256);
    m_cat_pf = M.alloc_n<f16>(P16 * 768);
    m_pf_a = M.alloc_n<f16>(P16 * 768);
    m_common = M.alloc_n<f16>(P16 * 768);
    m_cat_sp = M.alloc_n<f16>(P16 * 512);
    m_sp_a = M.alloc_n<f16>(P16 * 512);
    m_means = M.alloc_n<f16>(P16 * 256);
    m_rh_common = M.alloc_n<f16>(P8 * 512);
    m_rh_a = M.alloc_n<f16>(P8 * 256);
    m_rh_head = M.alloc_n<f16>(P8 * 192);
    m_rh_common2 = M.alloc_n<f16>(P8 * 512);
    m_rh_a2 = M.alloc_n<f16>(P8 * 256);
    m_rh_head2 = M.alloc_n<f16>(P8 * 192);
    m_z_i8 = M.alloc_n<int8_t>(n * m_nz);
    m_yq = M.alloc_n<int8_t>(n * m_ny);
    m_dec_c = M.alloc_n<int8_t>(n * m_ny);
    m_sym = M.alloc_n<int16_t>(n * m_ny);
    m_symc = M.alloc_n<int16_t>(n * m_ny);
    m_cond = M.alloc_n<uint8_t>(n * m_ny);
    m_idx = M.alloc_n<uint8_t>(n * m_ny);
    m_idxc = M.alloc_n<uint8_t>(n * m_ny);
    // on the proxy's own stream: a legacy-stream operation here would break a
    // graph capture running on another thread's proxy
    hip_check(hipMemsetAsync(m_cat_enc, 0, P8 * 2048 * sizeof(f16), m_stream), "memset");
    hip_check(hipMemsetAsync(m_cat_fam, 0, P8 * 1024 * sizeof(f16), m_stream), "memset");
    hip_check(hipStreamSynchronize(m_stream), "memset");
    hip_check(hipHostMalloc((void**)&m_h_count, n * sizeof(int)), "pinned");
    hip_check(hipHostMalloc((void**)&m_h_sym, n * m_ny * sizeof(int16_t)), "pinned");
    hip_check(hipHostMalloc((void**)&m_h_idx, n * m_ny), "pinned");
    hip_check(hipHostMalloc((void**)&m_h_dec, n * m_ny), "pinned");
    hip_check(hipHostMalloc((void**)&m_h_z, n * m_nz), "pinned");
    m_bits.assign(n, {});
    m_last_counts.assign(n, 0);
    while ((int)m_decs.size() < n_img) m_decs.push_back(std::make_unique<RansDecoder>());
    m_decs.resize(n_img);
    set_coder_cdfs();
    for (int i = 0; i < kFrames; ++i) m_frames[i] = frames[i];
    m_ld_out = ld_out;
    m_memory_ready = false;
    m_memory_has_value = false;
    m_has_ref = false;
}

void HTSProxy::enter(hipStream_t caller) {
    hip_check(hipEventRecord(m_ev_in, caller), "event record");
    hip_check(hipStreamWaitEvent(m_stream, m_ev_in, 0), "stream wait");
    // side-stream work of an earlier call that threw before joining it may
    // still read the memory: order this call after it
    hip_check(hipStreamWaitEvent(m_stream, m_ev_join, 0), "stream wait");
}

void HTSProxy::leave(hipStream_t caller) {
    hip_check(hipEventRecord(m_ev_out, m_stream), "event record");
    hip_check(hipStreamWaitEvent(caller, m_ev_out, 0), "stream wait");
}

void HTSProxy::add_ref_feature_from_frame(const f16* frame, int ld, int Hp, int Wp, bool apply_adaptor, hipStream_t s) {
    if (!m_w) throw std::logic_error("set_param must be called first");
    if (Hp != m_Hp || Wp != m_Wp) throw std::logic_error("prepare() was not called for this resolution");
    enter(s);
    Exec e = exec();
    // pixel_unshuffle(frame, 8) (dmc_hts_proxy.cpp:496); eager because the frame pointer changes
    hip_check(pad_unshuffle(frame, ld, m_nimg, Hp, Wp, 3, 8, 0, 0, m_feature_i, 192, m_stream), "unshuffle");
    if (apply_adaptor) {
        const Act fi = make_act(m_feature_i, m_nimg, m_H8, m_W8, 192, 192);
        const Act mem = make_act(m_cat_fam, m_nimg, m_H8, m_W8, 512, 1024);
        const Act ctx = make_act(m_cat_enc + 1536, m_nimg, m_H8, m_W8, 512, 2048);
        const Act tmp = make_act(m_scratch.x1, m_nimg, m_H8, m_W8, 512, 512);
        m_w->fa_i.forward(e, fi, mem, tmp);
        m_w->fe.forward(e, mem, ctx, tmp);
    }
    m_memory_ready = apply_adaptor;
    m_memory_has_value = apply_adaptor;
    m_has_ref = true;
    leave(s);
}

// hyper decoder -> crop -> prior fusion -> common params (shared by both sides)
void HTSProxy::hyper_and_prior() {
    hyper_decode();
    prior_fusion();
}

void HTSProxy::hyper_decode() {
    const HTSWeights& w = *m_w;
    Exec e = exec();
    const Act zh = make_act(m_z_hat, m_nimg, m_H64, m_W64, 128, 128);
    const Act hda = make_act(m_hd_a, m_nimg, m_H64 * 2, m_W64 * 2, 256, 256);
    const Act hdb = make_act(m_hd_b, m_nimg, m_H16p, m_W16p, 256, 256);
    const Act hyper = make_act(m_cat_pf, m_nimg, m_H16, m_W16, 256, 768);
    w.hd0.forward(e, zh, hda);
    w.hd1.forward(e, hda, hdb);
    if (m_pad4) {
        w.hd2.forward(e, hdb, hdb);
        hip_check(replicate_pad(m_hd_b, 256, m_nimg, m_H16p, m_W16p, 256, m_H16, m_W16, m_cat_pf, 768, m_stream),
                  "crop");
    } else {
        w.hd2.forward(e, hdb, hyper);
    }
}

void HTSProxy::prior_fusion() {
    const HTSWeights& w = *m_w;
    Exec e = exec();
    const Act catpf = make_act(m_cat_pf, m_nimg, m_H16, m_W16, 768, 768);
    const Act pfa = make_act(m_pf_a, m_nimg, m_H16, m_W16, 768, 768);
    const Act tmp = make_act(m_scratch.x1, m_nimg, m_H16, m_W16, 768, 768);
    w.pf.forward(e, catpf, pfa, tmp);
    w.pf_out.run(e, pfa, make_act(m_common, m_nimg, m_H16, m_W16, 768, 768), EPI_BIAS);
}

void HTSProxy::spatial_means(int step) {
    const HTSWeights& w = *m_w;
    Exec e = exec();
    const Act catsp = make_act(m_cat_sp, m_nimg, m_H16, m_W16, 512, 512);
    const Act spa = make_act(m_sp_a, m_nimg, m_H16, m_W16, 512, 512);
    const Act tmp = make_act(m_scratch.x1, m_nimg, m_H16, m_W16, 512, 512);
    w.spa[step - 1].forward(e, catsp, spa);
    w.sp.forward(e, spa, spa, tmp);
    w.sp_out.run(e, spa, make_act(m_means, m_nimg, m_H16, m_W16, 256, 256), EPI_BIAS);
}

// lambda_enc_0 (dmc_hts_proxy.cpp:516-557)
void HTSProxy::enc_stage0() {
    const HTSWeights& w = *m_w;
    Exec e = exec();
    const Act cat = make_act(m_cat_enc, m_nimg, m_H8, m_W8, 2048, 2048);
    const Act eo = make_act(m_enc_out, m_nimg, m_H8, m_W8, 512, 512);
    const Act tmp8 = make_act(m_scratch.x1, m_nimg, m_H8, m_W8, 512, 512);
    QuantSel qe{w.q_encoder, 512};
    w.enc.forward(e, cat, eo, tmp8, &qe);
    const Act y = make_act(m_y, m_nimg, m_H16, m_W16, 256, 256);
    w.enc_down.run(e, eo, y);
    const Act ypad = make_act(m_y_pad, m_nimg, m_H16p, m_W16p, 256, 256);
    if (m_pad4)
        hip_check(replicate_pad(m_y, 256, m_nimg, m_H16, m_W16, 256, m_H16p, m_W16p, m_y_pad, 256, m_stream), "pad");
    const Act hea = make_act(m_he_a, m_nimg, m_H16p, m_W16p, 256, 256);
    const Act heb = make_act(m_he_b, m_nimg, m_H16p / 2, m_W16p / 2, 256, 256);
    const Act z = make_act(m_z, m_nimg, m_H64, m_W64, 128, 128);
    w.he0.forward(e, ypad, hea);
    w.he1.forward(e, hea, heb);
    w.he2.forward(e, heb, z);
    hip_check(round_z(m_z, 128, (long)m_nimg * m_H64 * m_W64, 128, m_z_hat, 128, m_z_i8, m_stream), "round_z");
    hyper_decode();
}

// temporal prior from memory * q_feature: needs only the memory and qp, so
// it runs on the side stream while the encoder and hyper coder run; it
// writes channels 256-767 of the prior-fusion input, the hyper decoder 0-255
void HTSProxy::enc_temporal() {
    const HTSWeights& w = *m_w;
    Exec e = exec_side();
    hip_check(mul_channel(m_cat_fam, 1024, (long)m_nimg * m_H8 * m_W8, 512, w.q_feature, 512, m_d_qp, m_tin, 512,
                          m_side, (long)m_H8 * m_W8),
              "mul");
    const Act tin = make_act(m_tin, m_nimg, m_H8, m_W8, 512, 512);
    const Act temporal = make_act(m_cat_pf + 256, m_nimg, m_H16, m_W16, 512, 768);
    w.tpe.forward(e, tin, temporal);
}

void HTSProxy::enc_stage0b() {
    const HTSWeights& w = *m_w;
    Exec e = exec();
    prior_fusion();
    const Act y = make_act(m_y, m_nimg, m_H16, m_W16, 256, 256);
    const Act common = make_act(m_common, m_nimg, m_H16, m_W16, 768, 768);
    w.reduction.run(e, common, make_act(m_cat_sp + 256, m_nimg, m_H16, m_W16, 256, 512), EPI_BIAS);
    const f16* qdec = m_common;
    const f16* scales = m_common + 256;
    for (int step = 0; step < 4; ++step) {
        const f16* means = (step == 0) ? m_common + 512 : m_means;
        const int ldm = (step == 0) ? 768 : 256;
        if (step > 0) spatial_means(step);
        hip_check(hts_enc_step(y, scales, 768, means, ldm, qdec, 768, m_cat_sp, 512, m_sym, m_cond, w.thresholds,
                               w.skip_thres, step, m_stream),
                  "quant step");
    }
    // per image: each image's symbols form its own bitstream
    for (int b = 0; b < m_nimg; ++b)
        hip_check(compact(m_sym + b * m_ny, 2, m_cond + b * m_ny, m_ny, m_symc + b * m_ny,
                          m_d_offs + (long)b * m_offs_stride, m_d_count + b, m_stream),
                  "compact");
}

// decoder + recon head (+ reset branch for the encoder's context update)
void HTSProxy::decoder_and_recon(bool recon_all, bool reset_for_encoder) {
    const HTSWeights& w = *m_w;
    Exec e = exec();
    const Act yhat = make_act(m_cat_sp, m_nimg, m_H16, m_W16, 256, 512);
    const Act up = make_act(m_cat_enc + 1024, m_nimg, m_H8, m_W8, 512, 2048);
    w.dec_up.run(e, yhat, up);
    const Act decin = make_act(m_cat_enc + 1024, m_nimg, m_H8, m_W8, 1024, 2048);
    const Act fp = make_act(m_cat_fam + 512, m_nimg, m_H8, m_W8, 512, 1024);
    const Act tmp = make_act(m_scratch.x1, m_nimg, m_H8, m_W8, 512, 512);
    QuantSel qd{w.q_decoder, 512};
    w.dec.forward(e, decin, fp, tmp, &qd);
    if (recon_all) {
        recon_frames(e, 0, kFrames, m_rh_common, m_rh_a, m_rh_head);
    } else if (reset_for_encoder) {
        // ReconHead(for_reset=True): conv1[-1] -> conv2[-1] (video_model_ht.py:254-256)
        const Act common = make_act(m_rh_common, m_nimg, m_H8, m_W8, 512, 512);
        const Act ra = make_act(m_rh_a, m_nimg, m_H8, m_W8, 256, 256);
        w.rh1[3].forward(e, fp, common);
        w.rh2[7][0].forward(e, common, ra, nullptr, false, true);
        w.rh2[7][1].forward(e, ra, ra, nullptr, true, true);
        w.rh2[7][2].forward(e, ra, ra, nullptr, true, false);
        w.rh_head[7].run(e, ra, make_act(m_feature_i, m_nimg, m_H8, m_W8, 192, 192), EPI_BIAS);
    }
}

// ReconHead frames [first, last) (video_model_ht.py:258-265): conv1[i/2] once
// per frame pair, conv2[i], pixel_shuffle(8) + clamp into the frame output;
// frame 7's head output is also the reset feature.
void HTSProxy::recon_frames(const Exec& e, int first, int last, f16* common_buf, f16* ra_buf, f16* head_buf) {
    const HTSWeights& w = *m_w;
    const Act fp = make_act(m_cat_fam + 512, m_nimg, m_H8, m_W8, 512, 1024);
    const Act common = make_act(common_buf, m_nimg, m_H8, m_W8, 512, 512);
    const Act ra = make_act(ra_buf, m_nimg, m_H8, m_W8, 256, 256);
    const Act fi = make_act(m_feature_i, m_nimg, m_H8, m_W8, 192, 192);
    for (int i = first; i < last; ++i) {
        if (i == first || i % 2 == 0) w.rh1[i / 2].forward(e, fp, common);
        w.rh2[i][0].forward(e, common, ra, nullptr, false, true);
        w.rh2[i][1].forward(e, ra, ra, nullptr, true, true);
        w.rh2[i][2].forward(e, ra, ra, nullptr, true, false);
        const Act head = (i == kFrames - 1) ? fi : make_act(head_buf, m_nimg, m_H8, m_W8, 192, 192);
        w.rh_head[i].run(e, ra, head, EPI_BIAS);
        hip_check(pixel_shuffle8_clamp(head.p, 192, m_nimg, m_H8, m_W8, 3, m_frames[i], m_ld_out, e.stream),
                  "shuffle");
    }
}

void HTSProxy::fork_side() {
    hip_check(hipEventRecord(m_ev_fork, m_stream), "event record");
    hip_check(hipStreamWaitEvent(m_side, m_ev_fork, 0), "stream wait");
}

void HTSProxy::join_side() {
    hip_check(hipEventRecord(m_ev_join, m_side), "event record");
    hip_check(hipStreamWaitEvent(m_stream, m_ev_join, 0), "stream wait");
}

// lambda_enc_1 (dmc_hts_proxy.cpp:571-574, 832-842)
void HTSProxy::enc_stage1(bool reset) {
    const HTSWeights& w = *m_w;
    Exec e = exec();
    decoder_and_recon(false, reset);
    const Act mem = make_act(m_cat_fam, m_nimg, m_H8, m_W8, 512, 1024);
    const Act tmp = make_act(m_scratch.x1, m_nimg, m_H8, m_W8, 512, 512);
    if (reset) {
        w.fa_i.forward(e, make_act(m_feature_i, m_nimg, m_H8, m_W8, 192, 192), mem, tmp);
    } else {
        w.fa_m.forward(e, make_act(m_cat_fam, m_nimg, m_H8, m_W8, 1024, 1024), mem, tmp);
    }
    w.fe.forward(e, mem, make_act(m_cat_enc + 1536, m_nimg, m_H8, m_W8, 512, 2048), tmp);
}

void HTSProxy::check_qps(const int* qps) const {
    if (!qps) throw std::invalid_argument("null qp array");
    for (int b = 0; b < m_nimg; ++b)
        if (qps[b] < 0 || qps[b] >= 64) throw std::invalid_argument("qp must be in [0, 64)");
}

void HTSProxy::set_qps(const int* qps) { set_device_ints(m_d_qp, qps, m_nimg, m_stream); }

int HTSProxy::compress(const f16* x, int ldx, int H, int W, int qp, bool reset, int pad_b, int pad_r, hipStream_t s,
                       const uint8_t** bits, long* nbytes) {
    if (m_nimg != 1) throw std::logic_error("this proxy codes a batch: use compress_batch");
    int ec = 0;
    compress_batch(x, ldx, H, W, &qp, reset, pad_b, pad_r, s, bits, nbytes, &ec);
    return ec;
}

void HTSProxy::compress_batch(const f16* x, int ldx, int H, int W, const int* qps, bool reset, int pad_b, int pad_r,
                              hipStream_t s, const uint8_t** bits, long* nbytes, int* ec_parallel) {
    if (!m_w) throw std::logic_error("set_param must be called first");
    check_qps(qps);
    if (H + pad_b != m_Hp || W + pad_r != m_Wp || pad_b < 0 || pad_r < 0 || ldx < 24)
        throw std::invalid_argument("input size does not match the prepared resolution");
    if (!m_memory_ready) throw std::logic_error("add_ref_feature_from_frame(apply_adaptor=True) must precede compress");
    enter(s);
    set_qps(qps);
    // pad_and_unshuffle_8 is eager: x changes every call (dmc_hts_proxy.cpp:512-513)
    hip_check(pad_unshuffle(x, ldx, m_nimg, H, W, 24, 8, pad_b, pad_r, m_cat_enc, 2048, m_stream), "unshuffle");
    fork_side();
    m_g_enc_tp.run(m_side, m_use_graphs, [&] { enc_temporal(); });
    hip_check(hipEventRecord(m_ev_join, m_side), "event record");
    m_g_enc0.run(m_stream, m_use_graphs, [&] { enc_stage0(); });
    hip_check(hipStreamWaitEvent(m_stream, m_ev_join, 0), "stream wait");
    m_g_enc0b.run(m_stream, m_use_graphs, [&] { enc_stage0b(); });
    hip_check(hipEventRecord(m_ev_sym, m_stream), "event");
    // the caller needs nothing from the context update (it only feeds this
    // proxy's next call, ordered on the proxy stream): release the caller's
    // stream now, so a decoder on it overlaps the update
    leave(s);
    m_g_enc1[reset ? 1 : 0].run(m_stream, m_use_graphs, [&] { enc_stage1(reset); });
    // entropy coding on the host while the GPU runs the context update
    hip_check(hipStreamWaitEvent(m_copy, m_ev_sym, 0), "wait");
    hip_check(hipMemcpyAsync(m_h_count, m_d_count, m_nimg * sizeof(int), hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipMemcpyAsync(m_h_z, m_z_i8, m_nimg * m_nz, hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipStreamSynchronize(m_copy), "sync");
    for (int b = 0; b < m_nimg; ++b) {
        const long n = m_h_count[b];
        if (n < 0 || n > m_ny) throw StatusError(DCVC_ERR_HIP, "symbol count out of range");
        m_last_counts[b] = n;
        hip_check(hipMemcpyAsync(m_h_sym + b * m_ny, m_symc + b * m_ny, n * sizeof(int16_t), hipMemcpyDeviceToHost,
                                 m_copy),
                  "d2h");
    }
    hip_check(hipStreamSynchronize(m_copy), "sync");
    for (int b = 0; b < m_nimg; ++b) {
        const long n = m_last_counts[b];
        const int ec = compute_ec_parallel(n);
        m_enc.reset();
        m_enc.set_parallel(ec);
        m_enc.encode_y(m_h_sym + b * m_ny, n);
        m_enc.encode_z(m_h_z + b * m_nz, m_nz, qps[b] * 128, 128);
        m_enc.flush();
        m_bits[b] = m_enc.stream();
        bits[b] = m_bits[b].data();
        nbytes[b] = (long)m_bits[b].size();
        ec_parallel[b] = ec;
    }
    m_memory_ready = true;
}

// lambda_dec_0 (dmc_hts_proxy.cpp:611-615)
void HTSProxy::dec_stage0(bool memory_has_value) {
    const HTSWeights& w = *m_w;
    Exec e = exec();
    const Act mem = make_act(m_cat_fam, m_nimg, m_H8, m_W8, 512, 1024);
    const Act tmp = make_act(m_scratch.x1, m_nimg, m_H8, m_W8, 512, 512);
    if (memory_has_value)
        w.fa_m.forward(e, make_act(m_cat_fam, m_nimg, m_H8, m_W8, 1024, 1024), mem, tmp);
    else
        w.fa_i.forward(e, make_act(m_feature_i, m_nimg, m_H8, m_W8, 192, 192), mem, tmp);
}

// the rest of lambda_dec_0: temporal prior from memory * q_feature, on the
// main stream while the side stream extracts features from the same memory
void HTSProxy::dec_stage0b() {
    const HTSWeights& w = *m_w;
    Exec e = exec();
    hip_check(mul_channel(m_cat_fam, 1024, (long)m_nimg * m_H8 * m_W8, 512, w.q_feature, 512, m_d_qp, m_tin, 512,
                          m_stream, (long)m_H8 * m_W8),
              "mul");
    w.tpe.forward(e, make_act(m_tin, m_nimg, m_H8, m_W8, 512, 512), make_act(m_cat_pf + 256, m_nimg, m_H16, m_W16, 512, 768));
}

// lambda_dec_1 (dmc_hts_proxy.cpp:634-646)
void HTSProxy::dec_stage1() {
    const HTSWeights& w = *m_w;
    hip_check(i8_to_f16(m_z_i8, (long)m_nimg * m_H64 * m_W64, 128, m_z_hat, 128, m_stream), "i8");
    hyper_and_prior();
    const Act scales = make_act(m_common + 256, m_nimg, m_H16, m_W16, 256, 768);
    hip_check(hts_dec_index(scales, w.thresholds, w.skip_thres, m_idx, m_cond, m_stream), "index");
    for (int b = 0; b < m_nimg; ++b)
        hip_check(compact(m_idx + b * m_ny, 1, m_cond + b * m_ny, m_ny, m_idxc + b * m_ny,
                          m_d_offs + (long)b * m_offs_stride, m_d_count + b, m_stream),
                  "compact");
}

// lambda_dec_2 (dmc_hts_proxy.cpp:660): on the side stream, concurrent with
// the prior (it needs only the memory, and only the decoder reads its ctx)
void HTSProxy::dec_stage2() {
    Exec e = exec_side();
    const Act mem = make_act(m_cat_fam, m_nimg, m_H8, m_W8, 512, 1024);
    const Act tmp = make_act(m_scratch2.x1, m_nimg, m_H8, m_W8, 512, 512);
    m_w->fe.forward(e, mem, make_act(m_cat_enc + 1536, m_nimg, m_H8, m_W8, 512, 2048), tmp);
}

// lambda_dec_3 (dmc_hts_proxy.cpp:676-702)
void HTSProxy::dec_stage3() {
    const HTSWeights& w = *m_w;
    Exec e = exec();
    for (int b = 0; b < m_nimg; ++b)
        hip_check(recover(m_dec_c + b * m_ny, m_cond + b * m_ny, m_ny, m_d_offs + (long)b * m_offs_stride,
                          m_yq + b * m_ny, m_stream),
                  "recover");
    const Act g = make_act(nullptr, m_nimg, m_H16, m_W16, 256, 256);
    hip_check(hts_dec_step(g, m_yq, m_common + 512, 768, m_common, 768, m_cat_sp, 512, 0, m_stream), "restore");
    w.reduction.run(e, make_act(m_common, m_nimg, m_H16, m_W16, 768, 768), make_act(m_cat_sp + 256, m_nimg, m_H16, m_W16, 256, 512),
                    EPI_BIAS);
    for (int step = 1; step < 4; ++step) {
        spatial_means(step);
        hip_check(hts_dec_step(g, m_yq, m_means, 256, m_common, 768, m_cat_sp, 512, step, m_stream), "restore");
    }
    decoder_and_recon(false, false);
}

void HTSProxy::decompress(const uint8_t* bits, long nbytes, int qp, int height, int width, int ec_parallel,
                          bool reset, hipStream_t s) {
    if (m_nimg != 1) throw std::logic_error("this proxy codes a batch: use decompress_batch");
    decompress_batch(&bits, &nbytes, &qp, height, width, &ec_parallel, reset, s);
}

void HTSProxy::decompress_batch(const uint8_t* const* bits, const long* nbytes, const int* qps, int height, int width,
                                const int* ec_parallel, bool reset, hipStream_t s) {
    if (!m_w) throw std::logic_error("set_param must be called first");
    check_qps(qps);
    if (!bits || !nbytes || !ec_parallel) throw std::invalid_argument("null argument");
    for (int b = 0; b < m_nimg; ++b)
        if ((!bits[b] && nbytes[b] > 0) || nbytes[b] < 0) throw std::invalid_argument("bad bitstream");
    if ((height + 15) / 16 * 16 != m_Hp || (width + 15) / 16 * 16 != m_Wp)
        throw std::invalid_argument("picture size does not match the prepared resolution");
    if (!m_memory_has_value && !m_has_ref)
        throw std::logic_error("add_ref_feature_from_frame must precede decompress");
    enter(s);
    set_qps(qps);
    m_g_dec0[m_memory_has_value ? 1 : 0].run(m_stream, m_use_graphs, [&] { dec_stage0(m_memory_has_value); });
    // the context from the new memory, on the side stream during the
    // temporal prior and dec_1
    fork_side();
    m_g_dec2.run(m_side, m_use_graphs, [&] { dec_stage2(); });
    hip_check(hipEventRecord(m_ev_join, m_side), "event record");
    m_g_dec0b.run(m_stream, m_use_graphs, [&] { dec_stage0b(); });
    // z decode on the host, concurrent with dec_0 on the GPU
    for (int b = 0; b < m_nimg; ++b) {
        RansDecoder& d = *m_decs[b];
        d.set_parallel(ec_parallel[b]);
        d.set_stream(bits[b], nbytes[b]);
        d.decode_z(m_nz, qps[b] * 128, 128, m_h_z + b * m_nz);
    }
    hip_check(hipMemcpyAsync(m_z_i8, m_h_z, m_nimg * m_nz, hipMemcpyHostToDevice, m_stream), "h2d");
    m_g_dec1.run(m_stream, m_use_graphs, [&] { dec_stage1(); });
    hip_check(hipEventRecord(m_ev_sym, m_stream), "event");
    hip_check(hipStreamWaitEvent(m_copy, m_ev_sym, 0), "wait");
    hip_check(hipMemcpyAsync(m_h_count, m_d_count, m_nimg * sizeof(int), hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipStreamSynchronize(m_copy), "sync");
    for (int b = 0; b < m_nimg; ++b) {
        const long n = m_h_count[b];
        if (n < 0 || n > m_ny) throw StatusError(DCVC_ERR_HIP, "index count out of range");
        m_last_counts[b] = n;
        hip_check(hipMemcpyAsync(m_h_idx + b * m_ny, m_idxc + b * m_ny, n, hipMemcpyDeviceToHost, m_copy), "d2h");
    }
    hip_check(hipStreamSynchronize(m_copy), "sync");
    for (int b = 0; b < m_nimg; ++b) {
        const long n = m_last_counts[b];
        m_decs[b]->decode_y(m_h_idx + b * m_ny, n, m_h_dec + b * m_ny);
        hip_check(hipMemcpyAsync(m_dec_c + b * m_ny, m_h_dec + b * m_ny, n, hipMemcpyHostToDevice, m_stream), "h2d");
    }
    hip_check(hipStreamWaitEvent(m_stream, m_ev_join, 0), "stream wait");  // ctx
    m_g_dec3.run(m_stream, m_use_graphs, [&] { dec_stage3(); });
    // recon head: frames 4-7 on the side stream, 0-3 here
    fork_side();
    m_g_rh[1].run(m_side, m_use_graphs,
                  [&] { recon_frames(exec_side(), kFrames / 2, kFrames, m_rh_common2, m_rh_a2, m_rh_head2); });
    m_g_rh[0].run(m_stream, m_use_graphs, [&] { recon_frames(exec(), 0, kFrames / 2, m_rh_common, m_rh_a, m_rh_head); });
    join_side();
    leave(s);
    // pinned staging buffers are reused by the next call: finish the uploads
    hip_check(hipStreamSynchronize(m_stream), "sync");
    m_memory_has_value = !reset;
    m_memory_ready = false;
}

void HTSProxy::copy_debug(int what, void* host, long bytes) {
    hip_check(hipStreamSynchronize(m_stream), "sync");
    const void* src = nullptr;
    long cap = 0;
    const long n = m_nimg;  // whole-batch buffers, images stacked
    switch (what) {
    case 0: src = m_cat_sp; cap = n * m_H16 * m_W16 * 512 * 2; break;    // y_hat | reduced
    case 1: src = m_y; cap = n * m_H16 * m_W16 * 256 * 2; break;         // y
    case 2: src = m_common; cap = n * m_H16 * m_W16 * 768 * 2; break;    // q_dec | scales | means
    case 3: src = m_cat_fam; cap = n * m_H8 * m_W8 * 1024 * 2; break;    // memory | feature_p
    case 4: src = m_cat_enc; cap = n * m_H8 * m_W8 * 2048 * 2; break;    // x_unshuffled | ctx
    case 5: src = m_sym; cap = n * m_ny * 2; break;                       // dense symbols
    case 6: src = m_z_i8; cap = n * m_nz; break;
    default: throw std::invalid_argument("unknown debug buffer");
    }
    if (bytes > cap) throw std::invalid_argument("debug copy too large");
    hip_check(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, m_stream), "debug copy");
    hip_check(hipStreamSynchronize(m_stream), "debug copy");
}

}  // namespace dcvc
