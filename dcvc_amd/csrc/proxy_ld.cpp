// DMC low-delay codec proxy (see proxy_ld.h).
//
// Buffer layout (NHWC, pixel rows), the reference's concatenation views
// (dmc_ld_proxy.cpp:726-747):
//   cat_dec [P8, 512] = (64 unused) | x_unshuffled[64:256) | ctx[256:512);
//                       the encoder input is [64:512), the decoder input
//                       [0:512) = up(y_hat)[0:256) | ctx
//   cat_fam [P8, 512] = memory[0:256) | feature_p[256:512)
//   cat_pf  [P16, 384] = hyper params[0:128) | temporal params[128:384)
//   cat_sp  [P16, 512] = y_hat[0:128) | q_dec[128:256) | scales[256:384) | means[384:512)
// The temporal params are computed when the memory is (end of every encode,
// dec_0 of every decode) and scaled by q_feature in place in the next stage,
// as the reference does.  One entropy round trip: the two checkerboard
// halves share the common scales, the second half takes its means from the
// spatial prior over the first half's y_hat.
#include "proxy_ld.h"

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

// DMCLDProxy::set_param (dmc_ld_proxy.cpp:595-637): every stride-2 and
// up-sampling block is shortcut-free (:217-222, 244-249, 373-376).
void LDWeights::load(const HostDict& d, float skip) {
    mem.free_all();
    fa_i.load(mem, d, "feature_adaptor_i.conv.", 4);
    fa_m.load(mem, d, "feature_adaptor_m.conv.", 4);
    fe.load(mem, d, "feature_extractor.conv.", 5);
    enc1.load(mem, d, "encoder.conv1.", 2);
    enc2.load(mem, d, "encoder.conv2.");
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
    sp.load(mem, d, "y_spatial_prior.conv.", 2);
    sp_out.load(mem, d, "y_spatial_prior.conv.2.");
    dec_up.load(mem, d, "decoder.up.");
    dec1.load(mem, d, "decoder.conv1.", 3);
    dec2.load(mem, d, "decoder.conv2.");
    rh.load(mem, d, "recon_head.conv.", 3);
    rh_head.load(mem, d, "recon_head.head.");
    q_encoder = upload_qtable(mem, d, "q_encoder", 256);
    q_decoder = upload_qtable(mem, d, "q_decoder", 256);
    q_feature = upload_qtable(mem, d, "q_feature", 256);
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
LDProxy::LDProxy() {
    hip_check(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking), "stream");
    hip_check(hipStreamCreateWithFlags(&m_copy, hipStreamNonBlocking), "stream");
    hip_check(hipStreamCreateWithFlags(&m_side, hipStreamNonBlocking), "stream");
    hip_check(hipEventCreateWithFlags(&m_ev_in, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&m_ev_out, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&m_ev_sym, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&m_ev_fork, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&m_ev_join, hipEventDisableTiming), "event");
}

LDProxy::~LDProxy() {
    if (m_stream) hipStreamSynchronize(m_stream);
    if (m_side) hipStreamSynchronize(m_side);
    reset_graphs();
    for (void* p : {(void*)m_h_count, (void*)m_h_sym, (void*)m_h_idx, (void*)m_h_dec, (void*)m_h_z}) hipHostFree(p);
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

void LDProxy::set_param(const HostDict& d, float skip_thres) {
    auto w = std::make_shared<LDWeights>();
    w->load(d, skip_thres);
    share_weights(w);
}

void LDProxy::share_weights(const std::shared_ptr<LDWeights>& w) {
    m_w = w;
    if (m_prec != PREC_F16) m_w->mem.ensure_low_precision();
    for (int i = 0; i < 2; ++i) {
        m_enc.set_cdf(w->cdf[i].data(), (int)w->cdf_len[i].size(), w->cdf_row[i], w->cdf_len[i].data(), i);
        m_dec.set_cdf(w->cdf[i].data(), (int)w->cdf_len[i].size(), w->cdf_row[i], w->cdf_len[i].data(), i);
    }
    reset_graphs();
}

void LDProxy::reset_graphs() {
    // a graph may still be running on the proxy stream (the encoder returns
    // before its context update finishes)
    if (m_stream) hipStreamSynchronize(m_stream);
    if (m_side) hipStreamSynchronize(m_side);
    m_g_enc0.reset();
    for (auto& g : m_g_enc1) g.reset();
    for (auto& g : m_g_dec0) g.reset();
    m_g_dec1.reset();
    m_g_dec2.reset();
    m_g_dec3.reset();
}

// pre_allocate_tensors (dmc_ld_proxy.cpp:693-794)
void LDProxy::prepare(int Hp, int Wp, f16* x_hat, int ld_out) {
    if (Hp <= 0 || Wp <= 0 || Hp % 16 || Wp % 16) throw std::invalid_argument("picture size must be padded to 16");
    if (Hp == m_Hp && Wp == m_Wp && x_hat == m_x_hat && ld_out == m_ld_out && m_cat_dec) return;
    hip_check(hipStreamSynchronize(m_stream), "sync");
    reset_graphs();
    m_buf.free_all();
    for (void* p : {(void*)m_h_count, (void*)m_h_sym, (void*)m_h_idx, (void*)m_h_dec, (void*)m_h_z}) hipHostFree(p);
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
    const long P8 = (long)m_H8 * m_W8, P16 = (long)m_H16 * m_W16, P16p = (long)m_H16p * m_W16p;
    const long P32p = P16p / 4, P64 = (long)m_H64 * m_W64;
    m_ny = P16 * 128;
    m_nz = P64 * 128;
    DevMem& M = m_buf;
    m_scratch.halves = (size_t)P8 * 512;
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
    m_d_qp = M.alloc_n<int>(1);
    m_d_count = M.alloc_n<int>(1);
    m_d_offs = M.alloc_n<int>(compact_blocks(m_ny) + 1);
    m_cat_dec = M.alloc_n<f16>(P8 * 512);
    m_cat_fam = M.alloc_n<f16>(P8 * 512);
    m_feature_i = M.alloc_n<f16>(P8 * 192);
    m_enc_out = M.alloc_n<f16>(P8 * 256);
    m_y = M.alloc_n<f16>(P16 * 128);
    m_y_pad = m_pad4 ? M.alloc_n<f16>(P16p * 128) : m_y;
    m_he_a = M.alloc_n<f16>(P16p * 128);
    m_he_b = M.alloc_n<f16>(P32p * 128);
    m_z = M.alloc_n<f16>(P64 * 128);
    m_z_hat = M.alloc_n<f16>(P64 * 128);
    m_hd_a = M.alloc_n<f16>(P32p * 128);
    m_hd_b = M.alloc_n<f16>(P16p * 128);
    m_cat_pf = M.alloc_n<f16>(P16 * 384);
    m_pf_a = M.alloc_n<f16>(P16 * 384);
    m_cat_sp = M.alloc_n<f16>(P16 * 512);
    m_sp_a = M.alloc_n<f16>(P16 * 256);
    m_means1 = M.alloc_n<f16>(P16 * 128);
    m_rh_a = M.alloc_n<f16>(P8 * 256);
    m_z_i8 = M.alloc_n<int8_t>(m_nz);
    m_yq = M.alloc_n<int8_t>(m_ny);
    m_dec_c = M.alloc_n<int8_t>(m_ny);
    m_sym = M.alloc_n<int16_t>(m_ny);
    m_symc = M.alloc_n<int16_t>(m_ny);
    m_cond = M.alloc_n<uint8_t>(m_ny);
    m_idx = M.alloc_n<uint8_t>(m_ny);
    m_idxc = M.alloc_n<uint8_t>(m_ny);
    hip_check(hipMemsetAsync(m_cat_dec, 0, P8 * 512 * sizeof(f16), m_stream), "memset");
    hip_check(hipMemsetAsync(m_cat_fam, 0, P8 * 512 * sizeof(f16), m_stream), "memset");
    hip_check(hipStreamSynchronize(m_stream), "memset");
    hip_check(hipHostMalloc((void**)&m_h_count, sizeof(int)), "pinned");
    hip_check(hipHostMalloc((void**)&m_h_sym, m_ny * sizeof(int16_t)), "pinned");
    hip_check(hipHostMalloc((void**)&m_h_idx, m_ny), "pinned");
    hip_check(hipHostMalloc((void**)&m_h_dec, m_ny), "pinned");
    hip_check(hipHostMalloc((void**)&m_h_z, m_nz), "pinned");
    m_x_hat = x_hat;
    m_ld_out = ld_out;
    m_memory_ready = false;
    m_memory_has_value = false;
    m_has_ref = false;
}

void LDProxy::enter(hipStream_t caller) {
    hip_check(hipEventRecord(m_ev_in, caller), "event record");
    hip_check(hipStreamWaitEvent(m_stream, m_ev_in, 0), "stream wait");
    // side-stream work of an earlier call that threw before joining it may
    // still read the memory: order this call after it
    hip_check(hipStreamWaitEvent(m_stream, m_ev_join, 0), "stream wait");
}

void LDProxy::leave(hipStream_t caller) {
    hip_check(hipEventRecord(m_ev_out, m_stream), "event record");
    hip_check(hipStreamWaitEvent(caller, m_ev_out, 0), "stream wait");
}

// memory = FA_M(cat(memory, feature_p)) or FA_I(feature_i) (dmc_ld_proxy.cpp:651-658)
void LDProxy::feature_adaptor(bool use_m) {
    const LDWeights& w = *m_w;
    Exec e = exec();
    const Act mem = make_act(m_cat_fam, 1, m_H8, m_W8, 256, 512);
    const Act tmp = make_act(m_scratch.x1, 1, m_H8, m_W8, 256, 256);
    if (use_m)
        w.fa_m.forward(e, make_act(m_cat_fam, 1, m_H8, m_W8, 512, 512), mem, tmp);
    else
        w.fa_i.forward(e, make_act(m_feature_i, 1, m_H8, m_W8, 192, 192), mem, tmp);
}

// ctx = FE(memory), temporal params = TPE(memory) (dmc_ld_proxy.cpp:414-416, 468-469)
void LDProxy::context_from_memory() {
    const LDWeights& w = *m_w;
    Exec e = exec();
    const Act mem = make_act(m_cat_fam, 1, m_H8, m_W8, 256, 512);
    const Act tmp = make_act(m_scratch.x1, 1, m_H8, m_W8, 256, 256);
    w.fe.forward(e, mem, make_act(m_cat_dec + 256, 1, m_H8, m_W8, 256, 512), tmp);
    w.tpe.forward(e, mem, make_act(m_cat_pf + 128, 1, m_H16, m_W16, 256, 384));
}

void LDProxy::add_ref_feature_from_frame(const f16* frame, int ld, int Hp, int Wp, bool apply_adaptor, hipStream_t s) {
    if (!m_w) throw std::logic_error("set_param must be called first");
    if (Hp != m_Hp || Wp != m_Wp) throw std::logic_error("prepare() was not called for this resolution");
    enter(s);
    hip_check(pad_unshuffle(frame, ld, 1, Hp, Wp, 3, 8, 0, 0, m_feature_i, 192, m_stream), "unshuffle");
    if (apply_adaptor) {
        feature_adaptor(false);
        context_from_memory();
    }
    m_memory_ready = apply_adaptor;
    m_memory_has_value = apply_adaptor;
    m_has_ref = true;
    leave(s);
}

// hyper decoder -> crop; temporal *= q_feature (in place); prior fusion -> common params
void LDProxy::hyper_and_prior() {
    const LDWeights& w = *m_w;
    Exec e = exec();
    const Act zh = make_act(m_z_hat, 1, m_H64, m_W64, 128, 128);
    const Act hda = make_act(m_hd_a, 1, m_H64 * 2, m_W64 * 2, 128, 128);
    const Act hdb = make_act(m_hd_b, 1, m_H16p, m_W16p, 128, 128);
    w.hd0.forward(e, zh, hda);
    w.hd1.forward(e, hda, hdb);
    if (m_pad4) {
        w.hd2.forward(e, hdb, hdb);
        hip_check(replicate_pad(m_hd_b, 128, 1, m_H16p, m_W16p, 128, m_H16, m_W16, m_cat_pf, 384, m_stream), "crop");
    } else {
        w.hd2.forward(e, hdb, make_act(m_cat_pf, 1, m_H16, m_W16, 128, 384));
    }
    hip_check(mul_channel(m_cat_pf + 128, 384, (long)m_H16 * m_W16, 256, w.q_feature, 256, m_d_qp, m_cat_pf + 128, 384,
                          m_stream),
              "mul");
    const Act catpf = make_act(m_cat_pf, 1, m_H16, m_W16, 384, 384);
    const Act pfa = make_act(m_pf_a, 1, m_H16, m_W16, 384, 384);
    const Act tmp = make_act(m_scratch.x1, 1, m_H16, m_W16, 384, 384);
    w.pf.forward(e, catpf, pfa, tmp);
    w.pf_out.run(e, pfa, make_act(m_cat_sp + 128, 1, m_H16, m_W16, 384, 512), EPI_BIAS);
}

// means of the second half: SpatialPrior(cat(y_hat_0, common params))
void LDProxy::spatial_means() {
    const LDWeights& w = *m_w;
    Exec e = exec();
    const Act spa = make_act(m_sp_a, 1, m_H16, m_W16, 256, 256);
    const Act tmp = make_act(m_scratch.x1, 1, m_H16, m_W16, 256, 256);
    w.sp.forward(e, make_act(m_cat_sp, 1, m_H16, m_W16, 512, 512), spa, tmp);
    w.sp_out.run(e, spa, make_act(m_means1, 1, m_H16, m_W16, 128, 128), EPI_BIAS);
}

// lambda_enc_0 (dmc_ld_proxy.cpp:432-455)
void LDProxy::enc_stage0() {
    const LDWeights& w = *m_w;
    Exec e = exec();
    const Act in = make_act(m_cat_dec + 64, 1, m_H8, m_W8, 448, 512);
    const Act eo = make_act(m_enc_out, 1, m_H8, m_W8, 256, 256);
    const Act tmp = make_act(m_scratch.x1, 1, m_H8, m_W8, 256, 256);
    w.enc1.forward(e, in, eo, tmp);
    QuantSel qe{w.q_encoder, 256, false};  // conv1x1_bias_shortcut_with_quant (:67)
    w.enc2.forward(e, eo, eo, &qe);
    const Act y = make_act(m_y, 1, m_H16, m_W16, 128, 128);
    w.enc_down.run(e, eo, y);
    if (m_pad4) hip_check(replicate_pad(m_y, 128, 1, m_H16, m_W16, 128, m_H16p, m_W16p, m_y_pad, 128, m_stream), "pad");
    const Act hea = make_act(m_he_a, 1, m_H16p, m_W16p, 128, 128);
    const Act heb = make_act(m_he_b, 1, m_H16p / 2, m_W16p / 2, 128, 128);
    w.he0.forward(e, make_act(m_y_pad, 1, m_H16p, m_W16p, 128, 128), hea);
    w.he1.forward(e, hea, heb);
    w.he2.forward(e, heb, make_act(m_z, 1, m_H64, m_W64, 128, 128));
    hip_check(round_z(m_z, 128, (long)m_H64 * m_W64, 128, m_z_hat, 128, m_z_i8, m_stream), "round_z");
    hyper_and_prior();
    const f16* qdec = m_cat_sp + 128;
    const f16* scales = m_cat_sp + 256;
    hip_check(ld_enc_step(y, scales, 512, m_cat_sp + 384, 512, qdec, 512, m_cat_sp, 512, m_sym, m_cond, w.thresholds,
                          w.skip_thres, 0, m_stream),
              "quant step");
    spatial_means();
    hip_check(ld_enc_step(y, scales, 512, m_means1, 128, qdec, 512, m_cat_sp, 512, m_sym, m_cond, w.thresholds,
                          w.skip_thres, 1, m_stream),
              "quant step");
    hip_check(compact(m_sym, 2, m_cond, m_ny, m_symc, m_d_offs, m_d_count, m_stream), "compact");
}

// decoder (+ recon head or its reset branch): dmc_ld_proxy.cpp:14-35, 286-305
void LDProxy::decoder_and_recon(bool recon, bool reset_for_encoder) {
    const LDWeights& w = *m_w;
    Exec e = exec();
    w.dec_up.run(e, make_act(m_cat_sp, 1, m_H16, m_W16, 128, 512), make_act(m_cat_dec, 1, m_H8, m_W8, 256, 512));
    const Act d1 = make_act(m_enc_out, 1, m_H8, m_W8, 256, 256);
    const Act tmp = make_act(m_scratch.x1, 1, m_H8, m_W8, 256, 256);
    w.dec1.forward(e, make_act(m_cat_dec, 1, m_H8, m_W8, 512, 512), d1, tmp);
    const Act fp = make_act(m_cat_fam + 256, 1, m_H8, m_W8, 256, 512);
    QuantSel qd{w.q_decoder, 256, false};  // conv1x1_bias_with_quant (:32)
    w.dec2.run(e, d1, fp, EPI_QUANT, nullptr, nullptr, &qd);
    if (recon || reset_for_encoder) {
        const Act ra = make_act(m_rh_a, 1, m_H8, m_W8, 256, 256);
        w.rh.forward(e, fp, ra, tmp);
        w.rh_head.run(e, ra, make_act(m_feature_i, 1, m_H8, m_W8, 192, 192), EPI_BIAS);
        if (recon)
            hip_check(pixel_shuffle8_clamp(m_feature_i, 192, 1, m_H8, m_W8, 3, m_x_hat, m_ld_out, m_stream), "shuffle");
    }
}

// lambda_enc_1 (dmc_ld_proxy.cpp:466-470, 639-649)
void LDProxy::enc_stage1(bool reset) {
    decoder_and_recon(false, reset);
    feature_adaptor(!reset);
    context_from_memory();
}

int LDProxy::compress(const f16* x, int ldx, int H, int W, int qp, bool reset, int pad_b, int pad_r, hipStream_t s,
                      const uint8_t** bits, long* nbytes) {
    if (!m_w) throw std::logic_error("set_param must be called first");
    if (qp < 0 || qp >= 64) throw std::invalid_argument("qp must be in [0, 64)");
    if (H + pad_b != m_Hp || W + pad_r != m_Wp || pad_b < 0 || pad_r < 0 || ldx < 3)
        throw std::invalid_argument("input size does not match the prepared resolution");
    if (!m_memory_ready) throw std::logic_error("add_ref_feature_from_frame(apply_adaptor=True) must precede compress");
    m_qp = qp;
    enter(s);
    set_device_int(m_d_qp, qp, m_stream);
    // pad_and_unshuffle_8 is eager: x changes every call (dmc_ld_proxy.cpp:429)
    hip_check(pad_unshuffle(x, ldx, 1, H, W, 3, 8, pad_b, pad_r, m_cat_dec + 64, 512, m_stream), "unshuffle");
    m_g_enc0.run(m_stream, m_use_graphs, [&] { enc_stage0(); });
    hip_check(hipEventRecord(m_ev_sym, m_stream), "event");
    // the caller needs nothing from the context update (it only feeds this
    // proxy's next call, ordered on the proxy stream): release the caller's
    // stream now, so a decoder on it overlaps the update
    leave(s);
    m_g_enc1[reset ? 1 : 0].run(m_stream, m_use_graphs, [&] { enc_stage1(reset); });
    // entropy coding on the host (worker, :807-820) while the GPU updates the context
    hip_check(hipStreamWaitEvent(m_copy, m_ev_sym, 0), "wait");
    hip_check(hipMemcpyAsync(m_h_count, m_d_count, sizeof(int), hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipMemcpyAsync(m_h_z, m_z_i8, m_nz, hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipStreamSynchronize(m_copy), "sync");
    const long n = *m_h_count;
    if (n < 0 || n > m_ny) throw StatusError(DCVC_ERR_HIP, "symbol count out of range");
    hip_check(hipMemcpyAsync(m_h_sym, m_symc, n * sizeof(int16_t), hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipStreamSynchronize(m_copy), "sync");
    m_last_count = n;
    const int ec = compute_ec_parallel(n);
    m_enc.reset();
    m_enc.set_parallel(ec);
    m_enc.encode_y(m_h_sym, n);
    m_enc.encode_z(m_h_z, m_nz, qp * 128, 128);
    m_enc.flush();
    m_memory_ready = true;
    *bits = m_enc.stream().data();
    *nbytes = (long)m_enc.stream().size();
    return ec;
}

// lambda_dec_0 (dmc_ld_proxy.cpp:509-515)
void LDProxy::dec_stage0(bool memory_has_value) {
    feature_adaptor(memory_has_value);
    m_w->tpe.forward(exec(), make_act(m_cat_fam, 1, m_H8, m_W8, 256, 512),
                     make_act(m_cat_pf + 128, 1, m_H16, m_W16, 256, 384));
}

// lambda_dec_1 (dmc_ld_proxy.cpp:531-547)
void LDProxy::dec_stage1() {
    const LDWeights& w = *m_w;
    hip_check(i8_to_f16(m_z_i8, (long)m_H64 * m_W64, 128, m_z_hat, 128, m_stream), "i8");
    hyper_and_prior();
    const Act scales = make_act(m_cat_sp + 256, 1, m_H16, m_W16, 128, 512);
    hip_check(hts_dec_index(scales, w.thresholds, w.skip_thres, m_idx, m_cond, m_stream), "index");
    hip_check(compact(m_idx, 1, m_cond, m_ny, m_idxc, m_d_offs, m_d_count, m_stream), "compact");
}

// lambda_dec_2 (dmc_ld_proxy.cpp:560): ctx from the new memory.  It needs
// only dec_0's output, so it runs on the side stream from dec_0 on, during
// the host's z decode and dec_1 (the reference starts it after dec_1).
void LDProxy::dec_stage2() {
    Exec e = exec_side();
    const Act mem = make_act(m_cat_fam, 1, m_H8, m_W8, 256, 512);
    const Act tmp = make_act(m_scratch2.x1, 1, m_H8, m_W8, 256, 256);
    m_w->fe.forward(e, mem, make_act(m_cat_dec + 256, 1, m_H8, m_W8, 256, 512), tmp);
}

// lambda_dec_3 (dmc_ld_proxy.cpp:576-585)
void LDProxy::dec_stage3() {
    hip_check(recover(m_dec_c, m_cond, m_ny, m_d_offs, m_yq, m_stream), "recover");
    const Act g = make_act(nullptr, 1, m_H16, m_W16, 128, 128);
    hip_check(ld_dec_step(g, m_yq, m_cat_sp + 384, 512, m_cat_sp + 128, 512, m_cat_sp, 512, 0, m_stream), "restore");
    spatial_means();
    hip_check(ld_dec_step(g, m_yq, m_means1, 128, m_cat_sp + 128, 512, m_cat_sp, 512, 1, m_stream), "restore");
    decoder_and_recon(true, false);
}

void LDProxy::decompress(const uint8_t* bits, long nbytes, int qp, int height, int width, int ec_parallel, bool reset,
                         hipStream_t s) {
    if (!m_w) throw std::logic_error("set_param must be called first");
    if (qp < 0 || qp >= 64) throw std::invalid_argument("qp must be in [0, 64)");
    if ((height + 15) / 16 * 16 != m_Hp || (width + 15) / 16 * 16 != m_Wp)
        throw std::invalid_argument("picture size does not match the prepared resolution");
    if (!m_memory_has_value && !m_has_ref)
        throw std::logic_error("add_ref_feature_from_frame must precede decompress");
    m_qp = qp;
    enter(s);
    set_device_int(m_d_qp, qp, m_stream);
    m_g_dec0[m_memory_has_value ? 1 : 0].run(m_stream, m_use_graphs, [&] { dec_stage0(m_memory_has_value); });
    hip_check(hipEventRecord(m_ev_fork, m_stream), "event record");
    hip_check(hipStreamWaitEvent(m_side, m_ev_fork, 0), "stream wait");
    m_g_dec2.run(m_side, m_use_graphs, [&] { dec_stage2(); });
    hip_check(hipEventRecord(m_ev_join, m_side), "event record");
    m_dec.set_parallel(ec_parallel);
    m_dec.set_stream(bits, nbytes);
    m_dec.decode_z(m_nz, qp * 128, 128, m_h_z);
    hip_check(hipMemcpyAsync(m_z_i8, m_h_z, m_nz, hipMemcpyHostToDevice, m_stream), "h2d");
    m_g_dec1.run(m_stream, m_use_graphs, [&] { dec_stage1(); });
    hip_check(hipEventRecord(m_ev_sym, m_stream), "event");
    hip_check(hipStreamWaitEvent(m_copy, m_ev_sym, 0), "wait");
    hip_check(hipMemcpyAsync(m_h_count, m_d_count, sizeof(int), hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipStreamSynchronize(m_copy), "sync");
    const long n = *m_h_count;
    if (n < 0 || n > m_ny) throw StatusError(DCVC_ERR_HIP, "index count out of range");
    hip_check(hipMemcpyAsync(m_h_idx, m_idxc, n, hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipStreamSynchronize(m_copy), "sync");
    m_last_count = n;
    m_dec.decode_y(m_h_idx, n, m_h_dec);
    hip_check(hipMemcpyAsync(m_dec_c, m_h_dec, n, hipMemcpyHostToDevice, m_stream), "h2d");
    hip_check(hipStreamWaitEvent(m_stream, m_ev_join, 0), "stream wait");  // ctx
    m_g_dec3.run(m_stream, m_use_graphs, [&] { dec_stage3(); });
    leave(s);
    hip_check(hipStreamSynchronize(m_stream), "sync");
    m_memory_has_value = !reset;  // host state, outside the graphs (:590)
    m_memory_ready = false;
}

void LDProxy::copy_debug(int what, void* host, long bytes) {
    hip_check(hipStreamSynchronize(m_stream), "sync");
    const void* src = nullptr;
    long cap = 0;
    switch (what) {
    case 0: src = m_cat_sp; cap = (long)m_H16 * m_W16 * 512 * 2; break;   // y_hat | q_dec | scales | means
    case 1: src = m_y; cap = (long)m_H16 * m_W16 * 128 * 2; break;        // y
    case 3: src = m_cat_fam; cap = (long)m_H8 * m_W8 * 512 * 2; break;    // memory | feature_p
    case 4: src = m_cat_dec; cap = (long)m_H8 * m_W8 * 512 * 2; break;    // up | x_unshuffled | ctx
    case 5: src = m_sym; cap = m_ny * 2; break;                            // dense symbols
    case 6: src = m_z_i8; cap = m_nz; break;
    case 7: src = m_cat_pf; cap = (long)m_H16 * m_W16 * 384 * 2; break;   // hyper | temporal (scaled)
    default: throw std::invalid_argument("unknown debug buffer");
    }
    if (bytes > cap) throw std::invalid_argument("debug copy too large");
    hip_check(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, m_stream), "debug copy");
    hip_check(hipStreamSynchronize(m_stream), "debug copy");
}

}  // namespace dcvc
