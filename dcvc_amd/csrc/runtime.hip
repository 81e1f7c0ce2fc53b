// Layer building blocks of the codec proxies (see runtime.h).
#include "runtime.h"

#include <cstring>

namespace dcvc {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw StatusError(DCVC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

const HostTensor& need(const HostDict& d, const std::string& key) {
    auto it = d.find(key);
    if (it == d.end()) throw std::invalid_argument("state dict is missing '" + key + "'");
    return it->second;
}

bool has(const HostDict& d, const std::string& key) { return d.find(key) != d.end(); }

DevMem::~DevMem() { free_all(); }

void* DevMem::alloc(size_t bytes) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes ? bytes : 16), "hipMalloc");
    m_blocks.push_back(p);
    m_bytes += bytes;
    return p;
}

void DevMem::ensure_low_precision() {
    std::lock_guard<std::mutex> lk(m_lowp_mu);
    if (m_lowp) return;
    for (Conv1x1* c : m_convs) c->make_low_precision(*this);
    hip_check(hipStreamSynchronize(nullptr), "weights to low precision");
    m_lowp = true;
}

void DevMem::free_all() {
    for (void* p : m_blocks) hipFree(p);
    m_blocks.clear();
    m_bytes = 0;
}

f16* upload_f16(DevMem& mem, const std::vector<float>& v) {
    std::vector<f16> h(v.size());
    for (size_t i = 0; i < v.size(); ++i) h[i] = (f16)v[i];
    f16* d = mem.alloc_n<f16>(v.size());
    hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(f16), hipMemcpyHostToDevice), "upload");
    return d;
}

float* upload_f32(DevMem& mem, const std::vector<float>& v) {
    float* d = mem.alloc_n<float>(v.size());
    hip_check(hipMemcpy(d, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice), "upload");
    return d;
}

// weights are rounded to fp16 once, as the reference's model.half() does
static inline float to_half_value(float x) { return (float)(f16)x; }

static void trace_gemm(int kind, int epi, long M, int N_out, int K, bool quad, int n_shortcut, hipStream_t s);

// ---------------------------------------------------------------- Conv1x1
void Conv1x1::load_raw(DevMem& mem, const std::vector<float>& wt, const std::vector<float>& bias, int N_, int K_) {
    N = N_;
    K = K_;
    w = upload_f16(mem, wt);
    b = upload_f32(mem, bias);
    mem.register_conv(this);
}

void Conv1x1::make_low_precision(DevMem& mem) {
    const int rows = quad ? 4 * N : N;
    w_bf16 = mem.alloc((size_t)rows * K * 2);
    w_fp8 = mem.alloc((size_t)rows * K);
    w_scale = mem.alloc_n<float>(rows);
    hip_check(weights_to_bf16(w, (long)rows * K, w_bf16, nullptr), "weights to bf16");
    hip_check(weights_to_fp8(w, rows, K, w_fp8, w_scale, nullptr), "weights to fp8");
}

void Conv1x1::load(DevMem& mem, const HostDict& d, const std::string& prefix, bool quad_) {
    const HostTensor& wt = need(d, prefix + "weight");
    if (wt.shape.size() != 4 || wt.shape[2] != 1 || wt.shape[3] != 1) throw std::invalid_argument(prefix + ": not 1x1");
    quad = quad_;
    const int rows = (int)wt.shape[0];
    K = (int)wt.shape[1];
    N = quad ? rows / 4 : rows;
    std::vector<float> bias(rows, 0.f);
    if (has(d, prefix + "bias")) {
        const HostTensor& bt = need(d, prefix + "bias");
        for (int i = 0; i < rows; ++i) bias[i] = to_half_value(bt.data[i]);
    }
    w = upload_f16(mem, wt.data);
    b = upload_f32(mem, bias);
    mem.register_conv(this);
}

void Conv1x1::run(const Exec& e, const Act& a, const Act& dst, int epi, const Act* s1, const Act* s2,
                  const QuantSel* q, const Fp8Io* io) const {
    GemmArgs g;
    g.A = a.p;
    g.lda = a.ld;
    g.M = (int)a.pixels();
    g.K = K;
    g.B = w;
    g.N = N;
    g.bias = b;
    g.D = dst.p;
    g.ldd = dst.ld;
    g.outH = a.H;
    g.outW = a.W;
    g.prec = e.prec;
    if (e.prec == PREC_BF16) {
        g.B = static_cast<const f16*>(w_bf16);
    } else if (e.prec == PREC_FP8) {
        g.B = static_cast<const f16*>(w_fp8);
        g.wscale = w_scale;
        if (io && io->a8) {
            g.A = reinterpret_cast<const f16*>(io->a8);
            g.lda = io->lda8;
            g.prec = PREC_FP8A;
        }
        if (io && io->d8) {
            g.D8 = io->d8;
            g.ldd8 = io->ldd8;
            if (!io->d16) g.D = nullptr;
        }
    }
    if (s1) {
        g.S1 = s1->p;
        g.lds1 = s1->ld;
    }
    if (s2) {
        g.S2 = s2->p;
        g.lds2 = s2->ld;
    }
    if (q) {
        g.q = q->table;
        g.qstride = q->stride;
        g.qp = e.d_qp;
        g.q_round = q->round_first ? 1 : 0;
    }
    if (a.C != K) throw std::invalid_argument("conv1x1: input channels do not match the weight");
    const int v = gemm_validate(g, epi, A_PLAIN, D_PLAIN);
    if (v != DCVC_OK) throw StatusError(v, "conv1x1: invalid GEMM arguments");
    trace_gemm(0, epi, g.M, N, K, epi == EPI_QUAD,
               (epi == EPI_SHORTCUT || epi == EPI_SHORTCUT_QUANT) ? 1 : (epi == EPI_SHORTCUT2 ? 2 : 0), e.stream);
    if (trace_enabled())
        trace_key(A_PLAIN, D_PLAIN, g.prec, gemm_tile_for(g, epi, A_PLAIN, D_PLAIN), a.n_img, a.H, a.W);
    hip_check(gemm_launch(g, epi, A_PLAIN, D_PLAIN, e.stream), "gemm launch");
    trace_launched(e.stream);
}

// ---------------------------------------------------------------- DW3x3
void DW3x3::load(DevMem& mem, const HostDict& d, const std::string& key) {
    const HostTensor& wt = need(d, key);  // [C, 1, 3, 3]
    C = (int)wt.shape[0];
    std::vector<float> t(9 * (size_t)C);
    for (int c = 0; c < C; ++c)
        for (int k = 0; k < 9; ++k) t[(size_t)k * C + c] = wt.data[(size_t)c * 9 + k];
    w = upload_f16(mem, t);
}

void DW3x3::run(const Exec& e, const Act& in, const Act& out, uint8_t* y8, int ldy8, bool y16) const {
    DwArgs a;
    a.x = in.p;
    a.ldx = in.ld;
    a.wt = w;
    a.y = (y16 || !y8) ? out.p : nullptr;
    a.ldy = out.ld;
    a.y8 = y8;
    a.ldy8 = ldy8;
    a.n_img = in.n_img;
    a.H = in.H;
    a.W = in.W;
    a.C = C;
    const int v = dw3x3_validate(a);
    if (v != DCVC_OK) throw StatusError(v, "d3x3: invalid arguments");
    if (trace_enabled()) {
        const long long M = in.pixels();
        trace_record(TraceRecord{3, 0, (int)M, C, 9, 18LL * M * C, 2LL * (2 * M * C + 9LL * C), -1}, e.stream);
        trace_key(0, 0, 0, dw3x3_strip_for(a), a.n_img, a.H, a.W);
    }
    hip_check(dw3x3_launch(a, e.stream), "d3x3 launch");
    trace_launched(e.stream);
}

// ---------------------------------------------------------------- DCB
void DCB::load(DevMem& mem, const HostDict& d, const std::string& p, bool shortcut_) {
    shortcut = shortcut_;
    has_adaptor = has(d, p + "adaptor.weight");
    if (has_adaptor) ad.load(mem, d, p + "adaptor.");
    dc0.load(mem, d, p + "dc.0.");
    dw.load(mem, d, p + "dc.2.weight");
    // dc.3 bias' = dc.3.bias + W3 . dc.2.bias (layers_proxy.cpp:175-178), in double
    {
        const HostTensor& w3 = need(d, p + "dc.3.weight");
        const int n = (int)w3.shape[0], k = (int)w3.shape[1];
        std::vector<float> b(n, 0.f);
        const HostTensor* b3 = has(d, p + "dc.3.bias") ? &need(d, p + "dc.3.bias") : nullptr;
        const HostTensor* bd = has(d, p + "dc.2.bias") ? &need(d, p + "dc.2.bias") : nullptr;
        for (int o = 0; o < n; ++o) {
            double s = b3 ? (double)to_half_value(b3->data[o]) : 0.0;
            if (bd)
                for (int c = 0; c < k; ++c)
                    s += (double)to_half_value(w3.data[(size_t)o * k + c]) * (double)to_half_value(bd->data[c]);
            b[o] = (float)s;
        }
        dc3.load_raw(mem, w3.data, b, n, k);
    }
    ffn0.load(mem, d, p + "ffn.0.", true);
    ffn2.load(mem, d, p + "ffn.2.");
    cout = dc0.K;
    cmid = dc0.N;
    cin = has_adaptor ? ad.K : cout;
    if (dw.C != cmid || dc3.K != cmid || dc3.N != cout || ffn0.K != cout || ffn0.N != cmid || ffn2.K != cmid ||
        ffn2.N != cout)
        throw std::invalid_argument(p + ": DepthConvBlock layer shapes do not chain");
}

void DCB::forward(const Exec& e, const Act& in, const Act& out, const QuantSel* q, bool chain_in,
                  bool chain_out) const {
    if (in.C != cin || out.C != cout) throw std::invalid_argument("DCB: channel mismatch");
    Scratch& sc = *e.sc;
    const size_t need_mid = (size_t)in.pixels() * cmid;
    if (need_mid > sc.halves || (size_t)in.pixels() * cout > sc.halves) throw std::logic_error("DCB: scratch too small");
    Act x = in;
    if (has_adaptor) {
        ad.run(e, in, out, EPI_BIAS);
        x = out;
    }
    const Act t1 = make_act(sc.t1, in.n_img, in.H, in.W, cmid, cmid);
    const Act t2 = make_act(sc.t2, in.n_img, in.H, in.W, cmid, cmid);
    const Act o = shortcut ? make_act(sc.so, in.n_img, in.H, in.W, cout, cout) : out;
    if (shortcut && q) throw std::logic_error("DCB: quant tail with block shortcut is not used by the models");
    if (e.prec == PREC_FP8 && fp8_activations()) {
        // e4m3 activations between the products: the depthwise output and
        // the quad output exist only as e4m3 (each feeds one product), dc.3's
        // output also as an e4m3 copy (the quad's A; ffn.2 adds the fp16 one).
        // t2's buffer holds the depthwise bytes, then dc.3's copy behind them
        // (M * (cmid + cout) bytes fit in its M * max(cmid, cout) halves).
        // A chained block input / output copy sits at the start of t2's
        // buffer: read by dc.0 before the depthwise conv overwrites it, and
        // written by ffn.2 after dc.3 and the quad have read their operands.
        uint8_t* t2_8 = reinterpret_cast<uint8_t*>(sc.t2);
        uint8_t* o8 = t2_8 + (size_t)in.pixels() * cmid;
        uint8_t* t1_8 = reinterpret_cast<uint8_t*>(sc.t1);
        if (chain_in && !has_adaptor) {
            Fp8Io io;
            io.a8 = t2_8;
            io.lda8 = cin;
            dc0.run(e, x, t1, EPI_WSILU, nullptr, nullptr, nullptr, &io);
        } else {
            dc0.run(e, x, t1, EPI_WSILU);
        }
        dw.run(e, t1, t2, t2_8, cmid, false);
        Fp8Io io3;
        io3.a8 = t2_8;
        io3.lda8 = cmid;
        io3.d8 = o8;
        io3.ldd8 = cout;
        dc3.run(e, t2, o, EPI_SHORTCUT, &x, nullptr, nullptr, &io3);
        Fp8Io io0;
        io0.a8 = o8;
        io0.lda8 = cout;
        io0.d8 = t1_8;
        io0.ldd8 = cmid;
        io0.d16 = false;
        ffn0.run(e, o, t1, EPI_QUAD, nullptr, nullptr, nullptr, &io0);
        Fp8Io io2;
        io2.a8 = t1_8;
        io2.lda8 = cmid;
        if (chain_out) {
            io2.d8 = t2_8;
            io2.ldd8 = cout;
        }
        if (shortcut) {
            ffn2.run(e, t1, out, EPI_SHORTCUT2, &o, &x, nullptr, &io2);
        } else if (q) {
            ffn2.run(e, t1, out, EPI_SHORTCUT_QUANT, &o, nullptr, q, &io2);
        } else {
            ffn2.run(e, t1, out, EPI_SHORTCUT, &o, nullptr, nullptr, &io2);
        }
        return;
    }
    dc0.run(e, x, t1, EPI_WSILU);
    dw.run(e, t1, t2);
    dc3.run(e, t2, o, EPI_SHORTCUT, &x);
    // (the fused FFN reads one QP per launch: batches of images take the two GEMMs)
    if (ffn_use((int)in.pixels(), cout, cmid, e.prec) && (!q || in.n_img == 1)) {
        // ffn.0 and ffn.2 in one launch, the hidden activation kept on chip
        FfnArgs f;
        f.A = o.p;
        f.lda = o.ld;
        f.M = (int)in.pixels();
        f.C = cout;
        f.CM = cmid;
        f.W0 = ffn0.w;
        f.b0 = ffn0.b;
        f.W2 = ffn2.w;
        f.b2 = ffn2.b;
        f.D = out.p;
        f.ldd = out.ld;
        int epi = EPI_SHORTCUT;
        if (shortcut) {
            epi = EPI_SHORTCUT2;
            f.S2 = x.p;
            f.lds2 = x.ld;
        } else if (q) {
            epi = EPI_SHORTCUT_QUANT;
            f.q = q->table;
            f.qstride = q->stride;
            f.qp = e.d_qp;
            f.q_round = q->round_first ? 1 : 0;
        }
        const int v = ffn_validate(f, epi);
        if (v != DCVC_OK) throw StatusError(v, "DCB: invalid fused FFN arguments");
        if (trace_enabled()) {
            const long long M = f.M;
            const long long wb = 5LL * cmid * cout;  // ffn.0 (4 CM x C) + ffn.2 (C x CM)
            trace_record(TraceRecord{4, epi, f.M, cout, cmid, 2LL * M * wb,
                                     2LL * (M * cout * (shortcut ? 3 : 2) + wb), -1},
                         e.stream);
        }
        hip_check(ffn_launch(f, epi, e.stream), "fused ffn launch");
        trace_launched(e.stream);
        return;
    }
    ffn0.run(e, o, t1, EPI_QUAD);
    if (shortcut) {
        ffn2.run(e, t1, out, EPI_SHORTCUT2, &o, &x);
    } else if (q) {
        ffn2.run(e, t1, out, EPI_SHORTCUT_QUANT, &o, nullptr, q);
    } else {
        ffn2.run(e, t1, out, EPI_SHORTCUT, &o);
    }
}

void DCBSeq::load(DevMem& mem, const HostDict& d, const std::string& prefix, int n, int first) {
    blocks.resize(n - first);
    for (int i = first; i < n; ++i) blocks[i - first].load(mem, d, prefix + std::to_string(i) + ".");
}

void DCBSeq::forward(const Exec& e, const Act& in, const Act& out, const Act& tmp, const QuantSel* q) const {
    const int n = (int)blocks.size();
    if (n == 1) {
        blocks[0].forward(e, in, out, q);
        return;
    }
    Act t = channels(tmp, 0, blocks[0].cout);
    blocks[0].forward(e, in, t, nullptr, false, true);
    for (int i = 1; i + 1 < n; ++i) blocks[i].forward(e, t, t, nullptr, true, true);
    blocks[n - 1].forward(e, t, out, q, true, false);
}

// ---------------------------------------------------------------- ConvS2
void ConvS2::load(DevMem& mem, const HostDict& d, const std::string& prefix, int k_) {
    k = k_;
    const HostTensor& wt = need(d, prefix + "weight");
    N = (int)wt.shape[0];
    std::vector<float> t;
    if (k == 2) {
        // down.weight [N, 4*Cin, 1, 1], channel c*4 + dy*2 + dx of pixel_unshuffle(x, 2)
        Cin = (int)wt.shape[1] / 4;
        t.resize((size_t)N * 4 * Cin);
        for (int o = 0; o < N; ++o)
            for (int c = 0; c < Cin; ++c)
                for (int tap = 0; tap < 4; ++tap)
                    t[((size_t)o * 4 + tap) * Cin + c] = wt.data[((size_t)o * Cin + c) * 4 + tap];
    } else {
        // Conv2d weight [N, Cin, 3, 3] -> OHWI
        Cin = (int)wt.shape[1];
        t.resize((size_t)N * 9 * Cin);
        for (int o = 0; o < N; ++o)
            for (int c = 0; c < Cin; ++c)
                for (int tap = 0; tap < 9; ++tap)
                    t[((size_t)o * 9 + tap) * Cin + c] = wt.data[((size_t)o * Cin + c) * 9 + tap];
    }
    w = upload_f16(mem, t);
    std::vector<float> bias(N, 0.f);
    if (has(d, prefix + "bias"))
        for (int i = 0; i < N; ++i) bias[i] = to_half_value(need(d, prefix + "bias").data[i]);
    b = upload_f32(mem, bias);
}

void ConvS2::run(const Exec& e, const Act& in, const Act& out) const {
    GemmArgs g;
    g.A = in.p;
    g.lda = in.ld;
    g.inH = in.H;
    g.inW = in.W;
    g.outH = out.H;
    g.outW = out.W;
    g.cin = Cin;
    g.M = (int)out.pixels();
    g.K = k * k * Cin;
    g.B = w;
    g.N = N;
    g.bias = b;
    g.D = out.p;
    g.ldd = out.ld;
    const int amode = (k == 2) ? A_S2K2 : A_S2K3;
    const int v = gemm_validate(g, EPI_BIAS, amode, D_PLAIN);
    if (v != DCVC_OK || in.C != Cin || out.C != N) throw StatusError(DCVC_ERR_ARG, "strided conv: invalid arguments");
    if (trace_enabled()) {
        // input read once + weights + output
        const long long M = out.pixels();
        trace_record(TraceRecord{1, EPI_BIAS, (int)M, N, g.K, 2LL * M * N * g.K,
                                 2LL * (in.pixels() * (long long)Cin + (long long)N * g.K + M * N), -1},
                     e.stream);
    }
    if (trace_enabled())
        trace_key(amode, D_PLAIN, g.prec, gemm_tile_for(g, EPI_BIAS, amode, D_PLAIN), out.n_img, out.H, out.W);
    hip_check(gemm_launch(g, EPI_BIAS, amode, D_PLAIN, e.stream), "strided conv launch");
    trace_launched(e.stream);
}

// ---------------------------------------------------------------- Subpel
void Subpel::load(DevMem& mem, const HostDict& d, const std::string& prefix) {
    const HostTensor& wt = need(d, prefix + "conv.0.weight");  // [4*Cout, Cin, k, k]
    Cout = (int)wt.shape[0] / 4;
    Cin = (int)wt.shape[1];
    k = wt.shape.size() == 4 ? (int)wt.shape[2] : 1;
    if ((k != 1 && k != 3) || (wt.shape.size() == 4 && wt.shape[3] != k))
        throw std::invalid_argument(prefix + ": sub-pixel conv kernel must be 1x1 or 3x3");
    const int taps = k * k;
    // conv channel c*4 + sub lands at sub-pixel sub of channel c (pixel_shuffle(2));
    // GEMM row sub*Cout + c, K index tap*Cin + ci
    std::vector<float> t((size_t)4 * Cout * taps * Cin);
    std::vector<float> bias((size_t)4 * Cout, 0.f);
    const HostTensor* bt = has(d, prefix + "conv.0.bias") ? &need(d, prefix + "conv.0.bias") : nullptr;
    for (int c = 0; c < Cout; ++c)
        for (int sub = 0; sub < 4; ++sub) {
            const size_t row = (size_t)sub * Cout + c;
            const size_t src = (size_t)c * 4 + sub;
            for (int ci = 0; ci < Cin; ++ci)
                for (int tap = 0; tap < taps; ++tap)
                    t[(row * taps + tap) * Cin + ci] = wt.data[(src * Cin + ci) * taps + tap];
            if (bt) bias[row] = to_half_value(bt->data[src]);
        }
    w = upload_f16(mem, t);
    b = upload_f32(mem, bias);
}

void Subpel::run(const Exec& e, const Act& in, const Act& out) const {
    GemmArgs g;
    g.A = in.p;
    g.lda = in.ld;
    g.outH = in.H;
    g.outW = in.W;
    g.M = (int)in.pixels();
    g.K = k * k * Cin;
    g.B = w;
    g.N = 4 * Cout;
    g.bias = b;
    g.D = out.p;
    g.ldd = out.ld;
    const int amode = (k == 3) ? A_S1K3 : A_PLAIN;
    if (k == 3) {
        g.inH = in.H;
        g.inW = in.W;
        g.cin = Cin;
    }
    const int v = gemm_validate(g, EPI_BIAS, amode, D_SUBPEL);
    if (v != DCVC_OK || in.C != Cin || out.C != Cout || out.H != 2 * in.H || out.W != 2 * in.W)
        throw StatusError(DCVC_ERR_ARG, "subpel conv: invalid arguments");
    if (k == 1) {
        trace_gemm(2, EPI_BIAS, g.M, 4 * Cout, Cin, false, 0, e.stream);
    } else if (trace_enabled()) {
        // input read once + weights + output
        const long long M = in.pixels();
        trace_record(TraceRecord{2, EPI_BIAS, (int)M, 4 * Cout, g.K, 2LL * M * 4 * Cout * g.K,
                                 2LL * (M * Cin + 4LL * Cout * g.K + M * 4LL * Cout), -1},
                     e.stream);
    }
    if (trace_enabled())
        trace_key(amode, D_SUBPEL, g.prec, gemm_tile_for(g, EPI_BIAS, amode, D_SUBPEL), out.n_img, out.H, out.W);
    hip_check(gemm_launch(g, EPI_BIAS, amode, D_SUBPEL, e.stream), "subpel launch");
    trace_launched(e.stream);
}

void RBS2::load(DevMem& mem, const HostDict& d, const std::string& p, bool shortcut) {
    down.load(mem, d, p + "down.", 2);
    conv.load(mem, d, p + "conv.", shortcut);
}

void RBS2::forward(const Exec& e, const Act& in, const Act& out) const {
    down.run(e, in, out);
    conv.forward(e, out, out);
}

void RBU::load(DevMem& mem, const HostDict& d, const std::string& p, bool shortcut) {
    up.load(mem, d, p + "up.");
    conv.load(mem, d, p + "conv.", shortcut);
}

void RBU::forward(const Exec& e, const Act& in, const Act& out) const {
    up.run(e, in, out);
    conv.forward(e, out, out);
}

// ---------------------------------------------------------------- trace
namespace {
struct TraceEntry {
    TraceRecord r;
    hipEvent_t e0 = nullptr, e1 = nullptr;  // timed traces: around the launch, on its stream
};
std::vector<TraceEntry> g_trace;
bool g_trace_on = false;
bool g_trace_timed = false;
}  // namespace

bool trace_enabled() { return g_trace_on; }

void trace_record(const TraceRecord& r, hipStream_t s) {
    if (!g_trace_on) return;
    TraceEntry t;
    t.r = r;
    if (g_trace_timed) {
        hip_check(hipEventCreate(&t.e0), "trace event");
        hip_check(hipEventCreate(&t.e1), "trace event");
        hip_check(hipEventRecord(t.e0, s), "trace event");
    }
    g_trace.push_back(t);
}

void trace_key(int amode, int dmode, int prec, int cfg, int n_img, int H, int W) {
    if (!g_trace_on || g_trace.empty()) return;
    TraceRecord& r = g_trace.back().r;
    r.amode = amode;
    r.dmode = dmode;
    r.prec = prec;
    r.cfg = cfg;
    r.n_img = n_img;
    r.H = H;
    r.W = W;
}

void trace_launched(hipStream_t s) {
    if (!g_trace_on || !g_trace_timed || g_trace.empty()) return;
    hip_check(hipEventRecord(g_trace.back().e1, s), "trace event");
}

static void trace_gemm(int kind, int epi, long M, int N_out, int K, bool quad, int n_shortcut, hipStream_t s) {
    if (!g_trace_on) return;
    const long n_eff = quad ? 4L * N_out : N_out;
    TraceRecord r{kind, epi, (int)M, N_out, K, 2LL * M * n_eff * K,
                  2LL * (M * (long long)K + n_eff * (long long)K + M * (long long)N_out * (1 + n_shortcut)), -1};
    trace_record(r, s);
}

void trace_set(bool on, bool timed) {
    for (auto& t : g_trace) {
        if (t.e0) hipEventDestroy(t.e0);
        if (t.e1) hipEventDestroy(t.e1);
    }
    g_trace.clear();
    g_trace_on = on;
    g_trace_timed = timed;
}

std::vector<TraceRecord> trace_collect() {
    std::vector<TraceRecord> out;
    out.reserve(g_trace.size());
    for (auto& t : g_trace) {
        TraceRecord r = t.r;
        if (t.e0 && t.e1) {
            float ms = 0.f;
            hip_check(hipEventSynchronize(t.e1), "trace event");
            hip_check(hipEventElapsedTime(&ms, t.e0, t.e1), "trace event");
            r.ns = (long long)(ms * 1e6);
        }
        out.push_back(r);
    }
    trace_set(false, false);
    return out;
}

// ---------------------------------------------------------------- graphs
void Graph::reset() {
    if (m_exec) {
        hipGraphExecDestroy(m_exec);
        m_exec = nullptr;
    }
}

__global__ void set_int_kernel(int* dst, int v) { *dst = v; }

void set_device_int(int* dst, int v, hipStream_t s) {
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, dst, v);
    hip_check(hipGetLastError(), "set_device_int");
}

// the values travel as a kernel argument, so no host buffer must outlive the call
__global__ void set_ints_kernel(int* dst, DeviceInts v) {
    if ((int)threadIdx.x < v.n) dst[threadIdx.x] = v.v[threadIdx.x];
}

void set_device_ints(int* dst, const int* v, int n, hipStream_t s) {
    if (n < 1 || n > DeviceInts::kMax) throw std::invalid_argument("set_device_ints: 1..16 values");
    DeviceInts a;
    a.n = n;
    for (int i = 0; i < n; ++i) a.v[i] = v[i];
    hipLaunchKernelGGL(set_ints_kernel, dim3(1), dim3(64), 0, s, dst, a);
    hip_check(hipGetLastError(), "set_device_ints");
}


}  // namespace dcvc
