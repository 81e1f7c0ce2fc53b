// DMCI intra codec proxy (see proxy_dmci.h).
//
// Encode: enc_0 (encoder, hyper coder, prior, 4 checkerboard steps each
// writing one quarter of the symbols, dmci_proxy.cpp:308-394) then enc_1
// (decoder, :408-410) while the host entropy-codes steps 3..0 and z
// (:818-845).  Decode: z first, then one graph per step alternating with the
// host decoding that step's quarter (:423-602).
#include "proxy_dmci.h"

namespace dcvc {

static std::vector<int32_t> to_i32(const HostTensor& t) {
    std::vector<int32_t> v(t.data.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = (int32_t)t.data[i];
    return v;
}

void DMCIWeights::load(const HostDict& d, float skip) {
    mem.free_all();
    enc_1.load(mem, d, "enc.enc_1.");
    enc_2.load(mem, d, "enc.enc_2.", 6);
    enc_down.load(mem, d, "enc.enc_2.6.", 3);
    he0.load(mem, d, "hyper_enc.conv.0.");
    he1.load(mem, d, "hyper_enc.conv.1.", true);
    he2.load(mem, d, "hyper_enc.conv.2.", true);
    hd0.load(mem, d, "hyper_dec.conv.0.", true);
    hd1.load(mem, d, "hyper_dec.conv.1.", true);
    hd2.load(mem, d, "hyper_dec.conv.2.");
    pf.load(mem, d, "y_prior_fusion.conv.", 3);
    pf_out.load(mem, d, "y_prior_fusion.conv.3.");
    reduction.load(mem, d, "y_spatial_prior_reduction.");
    for (int i = 0; i < 3; ++i) spa[i].load(mem, d, "y_spatial_prior_adaptor_" + std::to_string(i + 1) + ".");
    sp.load(mem, d, "y_spatial_prior.conv.", 3);
    sp_out.load(mem, d, "y_spatial_prior.conv.3.");
    dec_up.load(mem, d, "dec.dec_1.0.", true);
    dec.load(mem, d, "dec.dec_1.", 13, 1);
    dec_2.load(mem, d, "dec.dec_2.");
    auto qt = [&](const char* k, int C) {
        const HostTensor& t = need(d, k);
        if (t.numel() != 64L * C) throw std::invalid_argument(std::string(k) + ": expected [64, C]");
        return upload_f16(mem, t.data);
    };
    q_scale_enc = qt("q_scale_enc", 384);
    q_scale_dec = qt("q_scale_dec", 384);
    q_scale_y_enc = qt("q_scale_y_enc", 256);
    q_scale_y_dec = qt("q_scale_y_dec", 256);
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

DMCIProxy::DMCIProxy() {
    hip_check(hipStreamCreateWithFlags(&m_stream, hipStreamNonBlocking), "stream");
    hip_check(hipStreamCreateWithFlags(&m_copy, hipStreamNonBlocking), "stream");
    hip_check(hipEventCreateWithFlags(&m_ev_in, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&m_ev_out, hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&m_ev_sym, hipEventDisableTiming), "event");
}

DMCIProxy::~DMCIProxy() {
    if (m_stream) hipStreamSynchronize(m_stream);
    reset_graphs();
    hipHostFree(m_h_counts);
    for (auto p : m_h_sym) hipHostFree(p);
    hipHostFree(m_h_idx);
    hipHostFree(m_h_dec);
    hipHostFree(m_h_z);
    m_buf.free_all();
    hipEventDestroy(m_ev_in);
    hipEventDestroy(m_ev_out);
    hipEventDestroy(m_ev_sym);
    hipStreamDestroy(m_copy);
    hipStreamDestroy(m_stream);
}

void DMCIProxy::set_param(const HostDict& d, float skip_thres) {
    auto w = std::make_unique<DMCIWeights>();
    w->load(d, skip_thres);
    m_w = std::move(w);
    if (m_prec != PREC_F16) m_w->mem.ensure_low_precision();
    for (int i = 0; i < 2; ++i) {
        m_enc.set_cdf(m_w->cdf[i].data(), (int)m_w->cdf_len[i].size(), m_w->cdf_row[i], m_w->cdf_len[i].data(), i);
        m_dec.set_cdf(m_w->cdf[i].data(), (int)m_w->cdf_len[i].size(), m_w->cdf_row[i], m_w->cdf_len[i].data(), i);
    }
    reset_graphs();
}

void DMCIProxy::reset_graphs() {
    // a graph may still be running on the proxy stream (the encoder returns
    // before its context update finishes)
    if (m_stream) hipStreamSynchronize(m_stream);
    m_g_enc0.reset();
    m_g_enc1.reset();
    for (auto& g : m_g_dec) g.reset();
}

void DMCIProxy::prepare(int Hp, int Wp, f16* x_hat, int ld_out) {
    if (Hp <= 0 || Wp <= 0 || Hp % 16 || Wp % 16) throw std::invalid_argument("picture size must be padded to 16");
    if (Hp == m_Hp && Wp == m_Wp && x_hat == m_x_hat && ld_out == m_ld_out && m_x_unsh) return;
    hip_check(hipStreamSynchronize(m_stream), "sync");
    reset_graphs();
    m_buf.free_all();
    hipHostFree(m_h_counts);
    for (auto& p : m_h_sym) hipHostFree(p);
    hipHostFree(m_h_idx);
    hipHostFree(m_h_dec);
    hipHostFree(m_h_z);
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
    m_nq = P16 * 64;
    m_nz = P64 * 128;
    DevMem& M = m_buf;
    m_scratch.halves = (size_t)P8 * 384;
    m_scratch.t1 = M.alloc_n<f16>(m_scratch.halves);
    m_scratch.t2 = M.alloc_n<f16>(m_scratch.halves);
    m_scratch.so = M.alloc_n<f16>(m_scratch.halves);
    m_scratch.x1 = M.alloc_n<f16>(m_scratch.halves);
    m_scratch.x2 = M.alloc_n<f16>(m_scratch.halves);
    m_d_qp = M.alloc_n<int>(1);
    m_d_count = M.alloc_n<int>(1);
    m_d_counts = M.alloc_n<int>(4);
    m_d_offs = M.alloc_n<int>(compact_blocks(m_nq) + 1);
    m_x_unsh = M.alloc_n<f16>(P8 * 192);
    m_e_a = M.alloc_n<f16>(P8 * 384);
    m_y = M.alloc_n<f16>(P16 * 256);
    m_y_pad = m_pad4 ? M.alloc_n<f16>(P16p * 256) : m_y;
    m_he_a = M.alloc_n<f16>(P16p * 128);
    m_he_b = M.alloc_n<f16>(P32p * 128);
    m_z = M.alloc_n<f16>(P64 * 128);
    m_z_hat = M.alloc_n<f16>(P64 * 128);
    m_hd_a = M.alloc_n<f16>(P32p * 128);
    m_hd_b = M.alloc_n<f16>(P16p * 128);
    m_hd_c = M.alloc_n<f16>(P16p * 256);
    m_pf_in = m_pad4 ? M.alloc_n<f16>(P16 * 256) : m_hd_c;
    m_pf_a = M.alloc_n<f16>(P16 * 512);
    m_params = M.alloc_n<f16>(P16 * 512);
    m_cat_sp = M.alloc_n<f16>(P16 * 512);
    m_sp_a = M.alloc_n<f16>(P16 * 512);
    m_sp_o = M.alloc_n<f16>(P16 * 512);
    m_y_hat = M.alloc_n<f16>(P16 * 256);
    m_d_a = M.alloc_n<f16>(P8 * 384);
    m_d_b = M.alloc_n<f16>(P8 * 192);
    m_z_i8 = M.alloc_n<int8_t>(m_nz);
    m_yq = M.alloc_n<int8_t>(m_nq);
    m_dec_c = M.alloc_n<int8_t>(m_nq);
    for (auto& p : m_symq) p = M.alloc_n<int16_t>(m_nq);
    for (auto& p : m_condq) p = M.alloc_n<uint8_t>(m_nq);
    for (auto& p : m_symc) p = M.alloc_n<int16_t>(m_nq);
    m_cond = M.alloc_n<uint8_t>(m_nq);
    m_idx = M.alloc_n<uint8_t>(m_nq);
    m_idxc = M.alloc_n<uint8_t>(m_nq);
    hip_check(hipHostMalloc((void**)&m_h_counts, 4 * sizeof(int)), "pinned");
    for (auto& p : m_h_sym) hip_check(hipHostMalloc((void**)&p, m_nq * sizeof(int16_t)), "pinned");
    hip_check(hipHostMalloc((void**)&m_h_idx, m_nq), "pinned");
    hip_check(hipHostMalloc((void**)&m_h_dec, m_nq), "pinned");
    hip_check(hipHostMalloc((void**)&m_h_z, m_nz), "pinned");
    m_x_hat = x_hat;
    m_ld_out = ld_out;
}

void DMCIProxy::enter(hipStream_t caller) {
    hip_check(hipEventRecord(m_ev_in, caller), "event record");
    hip_check(hipStreamWaitEvent(m_stream, m_ev_in, 0), "stream wait");
}

void DMCIProxy::leave(hipStream_t caller) {
    hip_check(hipEventRecord(m_ev_out, m_stream), "event record");
    hip_check(hipStreamWaitEvent(caller, m_ev_out, 0), "stream wait");
}

// hyper decoder -> crop -> prior fusion -> params (scales | means), reduction
void DMCIProxy::prior_from_z() {
    const DMCIWeights& w = *m_w;
    Exec e = exec();
    const Act zh = make_act(m_z_hat, 1, m_H64, m_W64, 128, 128);
    const Act hda = make_act(m_hd_a, 1, m_H64 * 2, m_W64 * 2, 128, 128);
    const Act hdb = make_act(m_hd_b, 1, m_H16p, m_W16p, 128, 128);
    const Act hdc = make_act(m_hd_c, 1, m_H16p, m_W16p, 256, 256);
    w.hd0.forward(e, zh, hda);
    w.hd1.forward(e, hda, hdb);
    w.hd2.forward(e, hdb, hdc);
    if (m_pad4)
        hip_check(replicate_pad(m_hd_c, 256, 1, m_H16p, m_W16p, 256, m_H16, m_W16, m_pf_in, 256, m_stream), "crop");
    const Act pfin = make_act(m_pf_in, 1, m_H16, m_W16, 256, 256);
    const Act pfa = make_act(m_pf_a, 1, m_H16, m_W16, 512, 512);
    const Act tmp = make_act(m_scratch.x1, 1, m_H16, m_W16, 512, 512);
    w.pf.forward(e, pfin, pfa, tmp);
    const Act params = make_act(m_params, 1, m_H16, m_W16, 512, 512);
    w.pf_out.run(e, pfa, params, EPI_BIAS);
    w.reduction.run(e, params, make_act(m_cat_sp + 256, 1, m_H16, m_W16, 256, 512), EPI_BIAS);
}

void DMCIProxy::spatial_prior(int step) {
    const DMCIWeights& w = *m_w;
    Exec e = exec();
    const Act cat = make_act(m_cat_sp, 1, m_H16, m_W16, 512, 512);
    const Act spa = make_act(m_sp_a, 1, m_H16, m_W16, 512, 512);
    const Act tmp = make_act(m_scratch.x1, 1, m_H16, m_W16, 512, 512);
    w.spa[step - 1].forward(e, cat, spa);
    w.sp.forward(e, spa, spa, tmp);
    w.sp_out.run(e, spa, make_act(m_sp_o, 1, m_H16, m_W16, 512, 512), EPI_BIAS);
}

void DMCIProxy::enc_stage0() {
    const DMCIWeights& w = *m_w;
    Exec e = exec();
    const Act xu = make_act(m_x_unsh, 1, m_H8, m_W8, 192, 192);
    const Act ea = make_act(m_e_a, 1, m_H8, m_W8, 384, 384);
    const Act tmp8 = make_act(m_scratch.x1, 1, m_H8, m_W8, 384, 384);
    QuantSel qe{w.q_scale_enc, 384};
    w.enc_1.forward(e, xu, ea, &qe);  // enc_1 then * q_scale_enc (image_model.py:66-68)
    w.enc_2.forward(e, ea, ea, tmp8);
    const Act y = make_act(m_y, 1, m_H16, m_W16, 256, 256);
    w.enc_down.run(e, ea, y);
    if (m_pad4) hip_check(replicate_pad(m_y, 256, 1, m_H16, m_W16, 256, m_H16p, m_W16p, m_y_pad, 256, m_stream), "pad");
    const Act hea = make_act(m_he_a, 1, m_H16p, m_W16p, 128, 128);
    const Act heb = make_act(m_he_b, 1, m_H16p / 2, m_W16p / 2, 128, 128);
    w.he0.forward(e, make_act(m_y_pad, 1, m_H16p, m_W16p, 256, 256), hea);
    w.he1.forward(e, hea, heb);
    w.he2.forward(e, heb, make_act(m_z, 1, m_H64, m_W64, 128, 128));
    hip_check(round_z(m_z, 128, (long)m_H64 * m_W64, 128, m_z_hat, 128, m_z_i8, m_stream), "round_z");
    prior_from_z();
    // y * q_scale_y_enc (dmci_proxy.cpp:325-326), after the hyper encoder consumed y
    hip_check(mul_channel(m_y, 256, (long)m_H16 * m_W16, 256, w.q_scale_y_enc, 256, m_d_qp, m_y, 256, m_stream), "mul");
    for (int step = 0; step < 4; ++step) {
        const f16* sc = (step == 0) ? m_params : m_sp_o;
        if (step > 0) spatial_prior(step);
        hip_check(intra_enc_step(y, sc, 512, sc + 256, 512, m_cat_sp, 512, w.q_scale_y_dec, 256, m_d_qp, m_y_hat, 256,
                                 m_symq[step], m_condq[step], w.thresholds, w.skip_thres, step, m_stream),
                  "quant step");
        hip_check(compact(m_symq[step], 2, m_condq[step], m_nq, m_symc[step], m_d_offs, m_d_counts + step, m_stream),
                  "compact");
    }
}

void DMCIProxy::enc_stage1() {
    const DMCIWeights& w = *m_w;
    Exec e = exec();
    const Act yh = make_act(m_y_hat, 1, m_H16, m_W16, 256, 256);
    const Act da = make_act(m_d_a, 1, m_H8, m_W8, 384, 384);
    const Act tmp = make_act(m_scratch.x1, 1, m_H8, m_W8, 384, 384);
    w.dec_up.forward(e, yh, da);
    QuantSel qd{w.q_scale_dec, 384};
    w.dec.forward(e, da, da, tmp, &qd);  // dec_1 then * q_scale_dec (image_model.py:42-44)
    const Act db = make_act(m_d_b, 1, m_H8, m_W8, 192, 192);
    w.dec_2.forward(e, da, db);
    hip_check(pixel_shuffle8_clamp(m_d_b, 192, 1, m_H8, m_W8, 3, m_x_hat, m_ld_out, m_stream), "shuffle");
}

int DMCIProxy::compress(const f16* x, int ldx, int H, int W, int qp, int pad_b, int pad_r, hipStream_t s,
                        const uint8_t** bits, long* nbytes) {
    if (!m_w) throw std::logic_error("set_param must be called first");
    if (qp < 0 || qp >= 64) throw std::invalid_argument("qp must be in [0, 64)");
    if (H + pad_b != m_Hp || W + pad_r != m_Wp || pad_b < 0 || pad_r < 0 || ldx < 3)
        throw std::invalid_argument("input size does not match the prepared resolution");
    m_qp = qp;
    enter(s);
    set_device_int(m_d_qp, qp, m_stream);
    hip_check(pad_unshuffle(x, ldx, 1, H, W, 3, 8, pad_b, pad_r, m_x_unsh, 192, m_stream), "unshuffle");
    m_g_enc0.run(m_stream, m_use_graphs, [&] { enc_stage0(); });
    hip_check(hipEventRecord(m_ev_sym, m_stream), "event");
    m_g_enc1.run(m_stream, m_use_graphs, [&] { enc_stage1(); });
    leave(s);
    hip_check(hipStreamWaitEvent(m_copy, m_ev_sym, 0), "wait");
    hip_check(hipMemcpyAsync(m_h_counts, m_d_counts, 4 * sizeof(int), hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipMemcpyAsync(m_h_z, m_z_i8, m_nz, hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipStreamSynchronize(m_copy), "sync");
    long total = 0;
    for (int i = 0; i < 4; ++i) {
        m_counts[i] = m_h_counts[i];
        if (m_counts[i] < 0 || m_counts[i] > m_nq) throw StatusError(DCVC_ERR_HIP, "symbol count out of range");
        total += m_counts[i];
        hip_check(hipMemcpyAsync(m_h_sym[i], m_symc[i], m_counts[i] * sizeof(int16_t), hipMemcpyDeviceToHost, m_copy),
                  "d2h");
    }
    hip_check(hipStreamSynchronize(m_copy), "sync");
    const int ec = compute_ec_parallel(total);
    m_enc.reset();
    m_enc.set_parallel(ec);
    for (int i = 3; i >= 0; --i) m_enc.encode_y(m_h_sym[i], m_counts[i]);  // steps 3..0 (dmci_proxy.cpp:838-841)
    m_enc.encode_z(m_h_z, m_nz, qp * 128, 128);
    m_enc.flush();
    *bits = m_enc.stream().data();
    *nbytes = (long)m_enc.stream().size();
    return ec;
}

void DMCIProxy::dec_index(int step) {
    const DMCIWeights& w = *m_w;
    const f16* sc = (step == 0) ? m_params : m_sp_o;
    const Act g = make_act(nullptr, 1, m_H16, m_W16, 256, 256);
    hip_check(intra_dec_index(g, sc, 512, w.thresholds, w.skip_thres, step, m_idx, m_cond, m_stream), "index");
    hip_check(compact(m_idx, 1, m_cond, m_nq, m_idxc, m_d_offs, m_d_count, m_stream), "compact");
}

void DMCIProxy::dec_restore(int step) {
    const DMCIWeights& w = *m_w;
    hip_check(recover(m_dec_c, m_cond, m_nq, m_d_offs, m_yq, m_stream), "recover");
    const f16* mn = (step == 0) ? m_params + 256 : m_sp_o + 256;
    const Act g = make_act(nullptr, 1, m_H16, m_W16, 256, 256);
    hip_check(intra_dec_step(g, m_yq, mn, 512, m_cat_sp, 512, w.q_scale_y_dec, 256, m_d_qp, m_y_hat, 256, step,
                             m_stream),
              "restore");
}

void DMCIProxy::decode_step_on_host(int step) {
    hip_check(hipEventRecord(m_ev_sym, m_stream), "event");
    hip_check(hipStreamWaitEvent(m_copy, m_ev_sym, 0), "wait");
    hip_check(hipMemcpyAsync(m_h_counts, m_d_count, sizeof(int), hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipStreamSynchronize(m_copy), "sync");
    const long n = m_h_counts[0];
    if (n < 0 || n > m_nq) throw StatusError(DCVC_ERR_HIP, "index count out of range");
    m_counts[step] = n;
    hip_check(hipMemcpyAsync(m_h_idx, m_idxc, n, hipMemcpyDeviceToHost, m_copy), "d2h");
    hip_check(hipStreamSynchronize(m_copy), "sync");
    m_dec.decode_y(m_h_idx, n, m_h_dec);
    hip_check(hipMemcpyAsync(m_dec_c, m_h_dec, n, hipMemcpyHostToDevice, m_stream), "h2d");
    // m_h_dec is reused by the next step: make sure the upload has completed
    hip_check(hipStreamSynchronize(m_stream), "sync");
}

void DMCIProxy::decompress(const uint8_t* bits, long nbytes, int qp, int height, int width, int ec_parallel,
                           hipStream_t s) {
    if (!m_w) throw std::logic_error("set_param must be called first");
    if (qp < 0 || qp >= 64) throw std::invalid_argument("qp must be in [0, 64)");
    if ((height + 15) / 16 * 16 != m_Hp || (width + 15) / 16 * 16 != m_Wp)
        throw std::invalid_argument("picture size does not match the prepared resolution");
    m_qp = qp;
    m_dec.set_parallel(ec_parallel);
    m_dec.set_stream(bits, nbytes);
    m_dec.decode_z(m_nz, qp * 128, 128, m_h_z);
    enter(s);
    set_device_int(m_d_qp, qp, m_stream);
    hip_check(hipMemcpyAsync(m_z_i8, m_h_z, m_nz, hipMemcpyHostToDevice, m_stream), "h2d");
    Exec e = exec();
    (void)e;
    m_g_dec[0].run(m_stream, m_use_graphs, [&] {
        hip_check(i8_to_f16(m_z_i8, (long)m_H64 * m_W64, 128, m_z_hat, 128, m_stream), "i8");
        prior_from_z();
        dec_index(0);
    });
    for (int step = 0; step < 4; ++step) {
        decode_step_on_host(step);
        m_g_dec[step + 1].run(m_stream, m_use_graphs, [&] {
            dec_restore(step);
            if (step < 3) {
                spatial_prior(step + 1);
                dec_index(step + 1);
            } else {
                enc_stage1();  // decoder: y_hat -> x_hat
            }
        });
    }
    leave(s);
}

void DMCIProxy::copy_debug(int what, void* host, long bytes) {
    hip_check(hipStreamSynchronize(m_stream), "sync");
    const void* src = nullptr;
    long cap = 0;
    if (what >= 0 && what < 4) {  // dense symbols of encoder step `what`
        src = m_symq[what];
        cap = m_nq * 2;
    } else if (what >= 4 && what < 8) {  // dense conditions of encoder step what-4
        src = m_condq[what - 4];
        cap = m_nq;
    } else if (what == 8) {  // final y_hat
        src = m_y_hat;
        cap = (long)m_H16 * m_W16 * 256 * 2;
    } else {
        throw std::invalid_argument("unknown debug buffer");
    }
    if (bytes > cap) throw std::invalid_argument("debug copy too large");
    hip_check(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, m_stream), "debug copy");
    hip_check(hipStreamSynchronize(m_stream), "debug copy");
}

}  // namespace dcvc
