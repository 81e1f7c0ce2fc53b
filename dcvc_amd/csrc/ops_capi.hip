// extern "C" kernel-level entry points declared in include/dcvc_ops.h.
#include "../../include/dcvc_ops.h"

#include <mutex>
#include <string>

#include "capi_util.h"
#include "common.h"
#include "dwconv.h"
#include "elementwise.h"
#include "ffn.h"
#include "gemm.h"

using namespace dcvc;

static inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
static inline int st(hipError_t e) { return e == hipSuccess ? DCVC_OK : DCVC_ERR_HIP; }
static inline const f16* H(const void* p) { return reinterpret_cast<const f16*>(p); }
static inline f16* Hm(void* p) { return reinterpret_cast<f16*>(p); }

static int run_gemm(GemmArgs& g, int epi, int amode, int dmode, void* stream) {
    const int v = gemm_validate(g, epi, amode, dmode);
    if (v != DCVC_OK) return v;
    return st(gemm_launch(g, epi, amode, dmode, S(stream)));
}

static GemmArgs plain_args(const void* A, int lda, int M, int K, const void* W, int N, const float* bias, void* D,
                           int ldd) {
    GemmArgs g;
    g.A = H(A);
    g.lda = lda;
    g.M = M;
    g.K = K;
    g.B = H(W);
    g.N = N;
    g.bias = bias;
    g.D = Hm(D);
    g.ldd = ldd;
    g.outH = 1;
    g.outW = M;
    return g;
}

// Block offsets of dcvc_conditional_index_noalloc: one int per COMPACT_TILE
// elements, in a file-scope device array so that the entry point allocates
// nothing.  Each GPU holds its own copy for the life of the process; every
// call on that GPU shares it.
constexpr int kNoallocMaxBlocks = 16384;
constexpr long kNoallocMaxElements = (long)kNoallocMaxBlocks * COMPACT_TILE;  // 67,108,864
static __device__ __align__(256) int scratch[kNoallocMaxBlocks];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static int* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current.  The host-side name `scratch` is not a
// device address and is never passed to a kernel.
static int noalloc_scratch(int** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_last_error("dcvc_conditional_index_noalloc: no current HIP device");
        return DCVC_ERR_HIP;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return DCVC_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<int*>(p);
    }
    *out = g_scratch_ptr[dev];
    return DCVC_OK;
}

extern "C" {

int dcvc_version(void) { return 100; }

int dcvc_scale_index_thresholds(float* th128) {
    if (!th128) return DCVC_ERR_ARG;
    host_scale_index_thresholds(th128);
    return DCVC_OK;
}

int dcvc_gemm_set_tile_policy(int policy) { return gemm_set_tile_policy(policy); }

int dcvc_gemm_tune_set(const long long* entries, int n) { return gemm_tune_set(entries, n); }

int dcvc_gemm_tile_for(int prec, int epi, int amode, int dmode, long M, int N, int K) {
    if (M <= 0 || N <= 0 || K <= 0 || epi < EPI_BIAS || epi > EPI_QUAD || amode < A_PLAIN || amode > A_S1K3 ||
        dmode < D_PLAIN || dmode > D_SUBPEL || prec < PREC_F16 || prec > PREC_FP8A || M > 0x7fffffffL)
        return -DCVC_ERR_ARG;
    GemmArgs g;
    g.M = (int)M;
    g.N = N;
    g.K = K;
    g.prec = prec;
    return gemm_tile_for(g, epi, amode, dmode);
}

int dcvc_d3x3_tune_set(const long long* entries, int n) { return dw3x3_tune_set(entries, n); }

int dcvc_d3x3_strip_for(int n_img, int H_, int W_, int C) {
    if (n_img <= 0 || H_ <= 0 || W_ <= 0 || C <= 0) return -DCVC_ERR_ARG;
    DwArgs a;
    a.n_img = n_img;
    a.H = H_;
    a.W = W_;
    a.C = C;
    return dw3x3_strip_for(a);
}

int dcvc_conv1x1_bias(const void* A, int lda, int M, int K, const void* W, int N, const float* bias, void* D,
                      int ldd, void* stream) {
    GemmArgs g = plain_args(A, lda, M, K, W, N, bias, D, ldd);
    return run_gemm(g, EPI_BIAS, A_PLAIN, D_PLAIN, stream);
}

int dcvc_conv1x1_bias_wsilu(const void* A, int lda, int M, int K, const void* W, int N, const float* bias, void* D,
                            int ldd, void* stream) {
    GemmArgs g = plain_args(A, lda, M, K, W, N, bias, D, ldd);
    return run_gemm(g, EPI_WSILU, A_PLAIN, D_PLAIN, stream);
}

int dcvc_conv1x1_bias_shortcut(const void* A, int lda, int M, int K, const void* W, int N, const float* bias,
                               const void* Sc, int lds, void* D, int ldd, void* stream) {
    GemmArgs g = plain_args(A, lda, M, K, W, N, bias, D, ldd);
    g.S1 = H(Sc);
    g.lds1 = lds;
    return run_gemm(g, EPI_SHORTCUT, A_PLAIN, D_PLAIN, stream);
}

int dcvc_conv1x1_bias_shortcut2(const void* A, int lda, int M, int K, const void* W, int N, const float* bias,
                                const void* S1, int lds1, const void* S2, int lds2, void* D, int ldd,
                                void* stream) {
    GemmArgs g = plain_args(A, lda, M, K, W, N, bias, D, ldd);
    g.S1 = H(S1);
    g.lds1 = lds1;
    g.S2 = H(S2);
    g.lds2 = lds2;
    return run_gemm(g, EPI_SHORTCUT2, A_PLAIN, D_PLAIN, stream);
}

int dcvc_conv1x1_bias_shortcut_with_quant(const void* A, int lda, int M, int K, const void* W, int N,
                                          const float* bias, const void* Sc, int lds, const void* q, int qstride,
                                          const int* qp, void* D, int ldd, void* stream) {
    GemmArgs g = plain_args(A, lda, M, K, W, N, bias, D, ldd);
    g.S1 = H(Sc);
    g.lds1 = lds;
    g.q = H(q);
    g.qstride = qstride;
    g.qp = qp;
    return run_gemm(g, EPI_SHORTCUT_QUANT, A_PLAIN, D_PLAIN, stream);
}

int dcvc_conv1x1_bias_shortcut_then_quant(const void* A, int lda, int M, int K, const void* W, int N,
                                          const float* bias, const void* Sc, int lds, const void* q, int qstride,
                                          const int* qp, void* D, int ldd, void* stream) {
    GemmArgs g = plain_args(A, lda, M, K, W, N, bias, D, ldd);
    g.S1 = H(Sc);
    g.lds1 = lds;
    g.q = H(q);
    g.qstride = qstride;
    g.qp = qp;
    g.q_round = 1;
    return run_gemm(g, EPI_SHORTCUT_QUANT, A_PLAIN, D_PLAIN, stream);
}

int dcvc_conv1x1_bias_with_quant(const void* A, int lda, int M, int K, const void* W, int N, const float* bias,
                                 const void* q, int qstride, const int* qp, void* D, int ldd, void* stream) {
    GemmArgs g = plain_args(A, lda, M, K, W, N, bias, D, ldd);
    g.q = H(q);
    g.qstride = qstride;
    g.qp = qp;
    return run_gemm(g, EPI_QUANT, A_PLAIN, D_PLAIN, stream);
}

int dcvc_conv1x1_bias_wsilu_chunk_add(const void* A, int lda, int M, int K, const void* W, int N,
                                      const float* bias, void* D, int ldd, void* stream) {
    GemmArgs g = plain_args(A, lda, M, K, W, N, bias, D, ldd);
    return run_gemm(g, EPI_QUAD, A_PLAIN, D_PLAIN, stream);
}

int dcvc_depth_conv_ffn(const void* A, int lda, int M, int C, int CM, const void* W0, const float* b0,
                        const void* W2, const float* b2, int epi, const void* S2, int lds2, const void* q,
                        int qstride, const int* qp, int q_round, void* D, int ldd, void* stream) {
    FfnArgs f;
    f.A = H(A);
    f.lda = lda;
    f.M = M;
    f.C = C;
    f.CM = CM;
    f.W0 = H(W0);
    f.b0 = b0;
    f.W2 = H(W2);
    f.b2 = b2;
    f.S2 = H(S2);
    f.lds2 = lds2;
    f.q = H(q);
    f.qstride = qstride;
    f.qp = qp;
    f.q_round = q_round;
    f.D = Hm(D);
    f.ldd = ldd;
    const int v = ffn_validate(f, epi);
    if (v != DCVC_OK) return v;
    return st(ffn_launch(f, epi, S(stream)));
}

int dcvc_set_ffn_fusion(int mode) { return ffn_set_mode(mode); }

int dcvc_conv1x1_grouped(const void* A, int lda, long sA, int a_div, int M, int K, const void* W, long sW, int N,
                         const float* bias, long sBias, int epi, const void* S1, int lds1, long sS1, const void* S2,
                         int lds2, long sS2, void* D, int ldd, long sD, int batch, void* stream) {
    if (epi != EPI_BIAS && epi != EPI_WSILU && epi != EPI_SHORTCUT && epi != EPI_SHORTCUT2 && epi != EPI_QUAD)
        return DCVC_ERR_ARG;
    GemmArgs g = plain_args(A, lda, M, K, W, N, bias, D, ldd);
    g.S1 = H(S1);
    g.lds1 = lds1;
    g.S2 = H(S2);
    g.lds2 = lds2;
    g.batch = batch;
    g.a_div = a_div;
    g.sA = sA;
    g.sB = sW;
    g.sBias = sBias;
    g.sS1 = sS1;
    g.sS2 = sS2;
    g.sD = sD;
    return run_gemm(g, epi, A_PLAIN, D_PLAIN, stream);
}

int dcvc_weights_to_lowp(const void* W, int rows, int K, int prec, void* Wq, float* scale, void* stream) {
    if (!W || !Wq || rows <= 0 || K <= 0 || K % 2 != 0) return DCVC_ERR_ARG;
    if (prec == PREC_BF16) return st(weights_to_bf16(H(W), (long)rows * K, Wq, S(stream)));
    if (prec == PREC_FP8) {
        if (!scale) return DCVC_ERR_ARG;
        return st(weights_to_fp8(H(W), rows, K, Wq, scale, S(stream)));
    }
    return DCVC_ERR_ARG;
}

int dcvc_gemm_lowp(const void* A, int lda, int M, int K, const void* Wq, int N, const float* scale, int prec,
                   int epi, const float* bias, const void* S1, int lds1, const void* S2, int lds2, const void* q,
                   int qstride, const int* qp, void* D, int ldd, void* stream) {
    if (epi < EPI_BIAS || epi > EPI_QUAD) return DCVC_ERR_ARG;
    GemmArgs g = plain_args(A, lda, M, K, Wq, N, bias, D, ldd);
    g.prec = prec;
    g.wscale = scale;
    g.S1 = H(S1);
    g.lds1 = lds1;
    g.S2 = H(S2);
    g.lds2 = lds2;
    g.q = H(q);
    g.qstride = qstride;
    g.qp = qp;
    return run_gemm(g, epi, A_PLAIN, D_PLAIN, stream);
}

int dcvc_gemm_fp8_act(const void* A, int lda, int a_fp8, int M, int K, const void* Wq, int N, const float* scale,
                      int prec, int epi, const float* bias, const void* S1, int lds1, const void* S2, int lds2,
                      const void* q, int qstride, const int* qp, void* D, int ldd, void* D8, int ldd8, void* stream) {
    if (epi < EPI_BIAS || epi > EPI_QUAD || prec < PREC_F16 || prec > PREC_FP8 || (a_fp8 != 0 && a_fp8 != 1))
        return DCVC_ERR_ARG;
    if (a_fp8 && prec != PREC_FP8) return DCVC_ERR_ARG;
    GemmArgs g = plain_args(A, lda, M, K, Wq, N, bias, D, ldd);
    g.prec = a_fp8 ? PREC_FP8A : prec;
    g.wscale = scale;
    g.S1 = H(S1);
    g.lds1 = lds1;
    g.S2 = H(S2);
    g.lds2 = lds2;
    g.q = H(q);
    g.qstride = qstride;
    g.qp = qp;
    g.D8 = static_cast<uint8_t*>(D8);
    g.ldd8 = ldd8;
    return run_gemm(g, epi, A_PLAIN, D_PLAIN, stream);
}

int dcvc_set_fp8_activations(int on) {
    if (on != 0 && on != 1) return DCVC_ERR_ARG;
    set_fp8_activations(on != 0);
    return DCVC_OK;
}

int dcvc_get_fp8_activations(void) { return fp8_activations() ? 1 : 0; }

int dcvc_conv_bias(const void* X, int ldx, int n_img, int H_, int W_, int Cin, const void* Wt, int N,
                   const float* bias, int kernel, int stride, void* D, int ldd, void* stream) {
    int amode;
    int Ho, Wo;
    if (kernel == 2 && stride == 2) {
        amode = A_S2K2;
        if (H_ % 2 || W_ % 2) return DCVC_ERR_ARG;
        Ho = H_ / 2;
        Wo = W_ / 2;
    } else if (kernel == 3 && stride == 2) {
        amode = A_S2K3;
        Ho = (H_ + 1) / 2;
        Wo = (W_ + 1) / 2;
    } else if (kernel == 3 && stride == 1) {
        amode = A_S1K3;
        Ho = H_;
        Wo = W_;
    } else {
        return DCVC_ERR_ARG;
    }
    GemmArgs g;
    g.A = H(X);
    g.lda = ldx;
    g.inH = H_;
    g.inW = W_;
    g.outH = Ho;
    g.outW = Wo;
    g.cin = Cin;
    g.M = n_img * Ho * Wo;
    g.K = kernel * kernel * Cin;
    g.B = H(Wt);
    g.N = N;
    g.bias = bias;
    g.D = Hm(D);
    g.ldd = ldd;
    return run_gemm(g, EPI_BIAS, amode, D_PLAIN, stream);
}

int dcvc_transposed_conv(const void* X, int ldx, int n_img, int H_, int W_, int Cin, const void* Wt, int Cout,
                         const float* bias, void* D, int ldd, void* stream) {
    GemmArgs g;
    g.A = H(X);
    g.lda = ldx;
    g.outH = H_;
    g.outW = W_;
    g.M = n_img * H_ * W_;
    g.K = Cin;
    g.B = H(Wt);
    g.N = 4 * Cout;
    g.bias = bias;
    g.D = Hm(D);
    g.ldd = ldd;
    if (!bias) return DCVC_ERR_ARG;  // callers pass a zero vector for bias-free layers
    return run_gemm(g, EPI_BIAS, A_PLAIN, D_SUBPEL, stream);
}

int dcvc_subpel_conv3x3(const void* X, int ldx, int n_img, int H_, int W_, int Cin, const void* Wt, int Cout,
                        const float* bias, void* D, int ldd, void* stream) {
    GemmArgs g;
    g.A = H(X);
    g.lda = ldx;
    g.inH = H_;
    g.inW = W_;
    g.outH = H_;
    g.outW = W_;
    g.cin = Cin;
    g.M = n_img * H_ * W_;
    g.K = 9 * Cin;
    g.B = H(Wt);
    g.N = 4 * Cout;
    g.bias = bias;
    g.D = Hm(D);
    g.ldd = ldd;
    if (!bias) return DCVC_ERR_ARG;
    return run_gemm(g, EPI_BIAS, A_S1K3, D_SUBPEL, stream);
}

int dcvc_d3x3(const void* X, int ldx, int n_img, int H_, int W_, int C, const void* Wt, void* Y, int ldy,
              void* stream) {
    if (!Y) return DCVC_ERR_ARG;
    return dcvc_d3x3_fp8_out(X, ldx, n_img, H_, W_, C, Wt, Y, ldy, nullptr, 0, stream);
}

int dcvc_d3x3_fp8_out(const void* X, int ldx, int n_img, int H_, int W_, int C, const void* Wt, void* Y, int ldy,
                      void* Y8, int ldy8, void* stream) {
    DwArgs a;
    a.x = H(X);
    a.ldx = ldx;
    a.wt = H(Wt);
    a.y = Hm(Y);
    a.ldy = ldy;
    a.y8 = static_cast<uint8_t*>(Y8);
    a.ldy8 = ldy8;
    a.n_img = n_img;
    a.H = H_;
    a.W = W_;
    a.C = C;
    const int v = dw3x3_validate(a);
    if (v != DCVC_OK) return v;
    return st(dw3x3_launch(a, S(stream)));
}

int dcvc_pad_and_unshuffle(const void* X, int ldx, int n_img, int H_, int W_, int Cin, int f, int padB, int padR,
                           void* Y, int ldy, void* stream) {
    if (!X || !Y || f <= 0 || padB < 0 || padR < 0 || (H_ + padB) % f || (W_ + padR) % f) return DCVC_ERR_ARG;
    if ((Cin * f * f) % 8 || ldy % 8 || ldy < Cin * f * f || ldx < Cin) return DCVC_ERR_ARG;
    return st(pad_unshuffle(H(X), ldx, n_img, H_, W_, Cin, f, padB, padR, Hm(Y), ldy, S(stream)));
}

int dcvc_pixel_shuffle_8(const void* X, int ldx, int n_img, int H8, int W8, int C, void* Y, int ldy, void* stream) {
    if (!X || !Y || ldx < C * 64 || ldy < C) return DCVC_ERR_ARG;
    return st(pixel_shuffle8_clamp(H(X), ldx, n_img, H8, W8, C, Hm(Y), ldy, S(stream)));
}

int dcvc_multiply_with_broadcast(const void* X, int ldx, long M, int C, const void* q, int qstride, const int* qp,
                                 void* Y, int ldy, void* stream) {
    if (!X || !Y || !q || C % 8 || ldx % 8 || ldy % 8 || ldx < C || ldy < C) return DCVC_ERR_ARG;
    return st(mul_channel(H(X), ldx, M, C, H(q), qstride, qp, Hm(Y), ldy, S(stream)));
}

int dcvc_round_z(const void* Z, int ldz, long M, int C, void* Zhat, int ldh, int8_t* Zi8, void* stream) {
    if (!Z || !Zhat || !Zi8) return DCVC_ERR_ARG;
    return st(round_z(H(Z), ldz, M, C, Hm(Zhat), ldh, Zi8, S(stream)));
}

int dcvc_int8_to_half(const int8_t* Zi8, long M, int C, void* Zhat, int ldh, void* stream) {
    if (!Zi8 || !Zhat) return DCVC_ERR_ARG;
    return st(i8_to_f16(Zi8, M, C, Hm(Zhat), ldh, S(stream)));
}

int dcvc_replicate_pad(const void* X, int ldx, int n_img, int H_, int W_, int C, int Ho, int Wo, void* Y, int ldy,
                       void* stream) {
    if (!X || !Y || C % 8 || ldx % 8 || ldy % 8) return DCVC_ERR_ARG;
    return st(replicate_pad(H(X), ldx, n_img, H_, W_, C, Ho, Wo, Hm(Y), ldy, S(stream)));
}

static Act act(const void* p, int ld, int n_img, int H_, int W_, int C) {
    Act a;
    a.p = const_cast<f16*>(H(p));
    a.ld = ld;
    a.n_img = n_img;
    a.H = H_;
    a.W = W_;
    a.C = C;
    return a;
}

int dcvc_hts_quant_step(const void* Y, int ldy, int n_img, int H_, int W_, int C, const void* Sc, int lds,
                        const void* Mn, int ldm, const void* Qd, int ldq, void* Yhat, int ldh, int16_t* sym,
                        uint8_t* cond, const float* th, float thr, int step, void* stream) {
    if (!Y || !Sc || !Mn || !Qd || !Yhat || !sym || !cond || !th || step < 0 || step > 3 || C % 4) return DCVC_ERR_ARG;
    return st(hts_enc_step(act(Y, ldy, n_img, H_, W_, C), H(Sc), lds, H(Mn), ldm, H(Qd), ldq, Hm(Yhat), ldh, sym,
                           cond, th, thr, step, S(stream)));
}

int dcvc_hts_build_index_dec(const void* Sc, int lds, int n_img, int H_, int W_, int C, const float* th, float thr,
                             uint8_t* idx, uint8_t* cond, void* stream) {
    if (!Sc || !idx || !cond || !th) return DCVC_ERR_ARG;
    return st(hts_dec_index(act(Sc, lds, n_img, H_, W_, C), th, thr, idx, cond, S(stream)));
}

int dcvc_hts_restore_step(const int8_t* yq, int n_img, int H_, int W_, int C, const void* Mn, int ldm,
                          const void* Qd, int ldq, void* Yhat, int ldh, int step, void* stream) {
    if (!yq || !Mn || !Qd || !Yhat || step < 0 || step > 3 || C % 4) return DCVC_ERR_ARG;
    return st(hts_dec_step(act(nullptr, 0, n_img, H_, W_, C), yq, H(Mn), ldm, H(Qd), ldq, Hm(Yhat), ldh, step,
                           S(stream)));
}

int dcvc_intra_quant_step(const void* Y, int ldy, int n_img, int H_, int W_, int C, const void* Sc, int lds,
                          const void* Mn, int ldm, void* Ysf, int ldsf, const void* q, int qstride, const int* qp,
                          void* Yout, int ldo, int16_t* sym, uint8_t* cond, const float* th, float thr, int step,
                          void* stream) {
    if (!Y || !Sc || !Mn || !Ysf || !sym || !cond || !th || step < 0 || step > 3 || C % 4) return DCVC_ERR_ARG;
    if (step == 3 && (!q || !Yout)) return DCVC_ERR_ARG;
    return st(intra_enc_step(act(Y, ldy, n_img, H_, W_, C), H(Sc), lds, H(Mn), ldm, Hm(Ysf), ldsf, H(q), qstride,
                             qp, Hm(Yout), ldo, sym, cond, th, thr, step, S(stream)));
}

int dcvc_intra_build_index_dec(const void* Sc, int lds, int n_img, int H_, int W_, int C, const float* th,
                               float thr, int step, uint8_t* idx, uint8_t* cond, void* stream) {
    if (!Sc || !idx || !cond || !th || step < 0 || step > 3 || C % 4) return DCVC_ERR_ARG;
    return st(intra_dec_index(act(nullptr, 0, n_img, H_, W_, C), H(Sc), lds, th, thr, step, idx, cond, S(stream)));
}

int dcvc_intra_restore_step(const int8_t* yq, int n_img, int H_, int W_, int C, const void* Mn, int ldm, void* Ysf,
                            int ldsf, const void* q, int qstride, const int* qp, void* Yout, int ldo, int step,
                            void* stream) {
    if (!yq || !Mn || !Ysf || step < 0 || step > 3 || C % 4) return DCVC_ERR_ARG;
    if (step == 3 && (!q || !Yout)) return DCVC_ERR_ARG;
    return st(intra_dec_step(act(nullptr, 0, n_img, H_, W_, C), yq, H(Mn), ldm, Hm(Ysf), ldsf, H(q), qstride, qp,
                             Hm(Yout), ldo, step, S(stream)));
}

int dcvc_ld_quant_step(const void* Y, int ldy, int n_img, int H_, int W_, int C, const void* Sc, int lds,
                       const void* Mn, int ldm, const void* Qd, int ldq, void* Yhat, int ldh, int16_t* sym,
                       uint8_t* cond, const float* th, float thr, int step, void* stream) {
    if (!Y || !Sc || !Mn || !Qd || !Yhat || !sym || !cond || !th || step < 0 || step > 1 || C % 2) return DCVC_ERR_ARG;
    return st(ld_enc_step(act(Y, ldy, n_img, H_, W_, C), H(Sc), lds, H(Mn), ldm, H(Qd), ldq, Hm(Yhat), ldh, sym, cond,
                          th, thr, step, S(stream)));
}

int dcvc_ld_restore_step(const int8_t* yq, int n_img, int H_, int W_, int C, const void* Mn, int ldm, const void* Qd,
                         int ldq, void* Yhat, int ldh, int step, void* stream) {
    if (!yq || !Mn || !Qd || !Yhat || step < 0 || step > 1 || C % 2) return DCVC_ERR_ARG;
    return st(ld_dec_step(act(nullptr, 0, n_img, H_, W_, C), yq, H(Mn), ldm, H(Qd), ldq, Hm(Yhat), ldh, step,
                          S(stream)));
}

int dcvc_htl_quant_step(const void* Y, int ldy, int n_img, int H_, int W_, int C, const void* Sc, int lds,
                        const void* Mn, int ldm, const void* Qd, int ldq, void* Ysf, int ldsf, void* Yout, int ldo,
                        int16_t* sym, uint8_t* cond, const float* th, float thr, int step, void* stream) {
    if (!Y || !Sc || !Mn || !Qd || !Ysf || !sym || !cond || !th || step < 0 || step > 3 || C % 4) return DCVC_ERR_ARG;
    if (step == 3 && !Yout) return DCVC_ERR_ARG;
    return st(htl_enc_step(act(Y, ldy, n_img, H_, W_, C), H(Sc), lds, H(Mn), ldm, H(Qd), ldq, Hm(Ysf), ldsf, Hm(Yout),
                           ldo, sym, cond, th, thr, step, S(stream)));
}

int dcvc_htl_restore_step(const int8_t* yq, int n_img, int H_, int W_, int C, const void* Mn, int ldm, const void* Qd,
                          int ldq, void* Ysf, int ldsf, void* Yout, int ldo, int step, void* stream) {
    if (!yq || !Mn || !Qd || !Ysf || step < 0 || step > 3 || C % 4) return DCVC_ERR_ARG;
    if (step == 3 && !Yout) return DCVC_ERR_ARG;
    return st(htl_dec_step(act(nullptr, 0, n_img, H_, W_, C), yq, H(Mn), ldm, H(Qd), ldq, Hm(Ysf), ldsf, Hm(Yout), ldo,
                           step, S(stream)));
}

int dcvc_compact_blocks(long N) { return compact_blocks(N); }

int dcvc_conditional_index(const void* vals, int esize, const uint8_t* cond, long N, void* out, int* block_offsets,
                           int* count, void* stream) {
    if ((esize != 1 && esize != 2) || N < 0 || !count || (N > 0 && (!vals || !cond || !out || !block_offsets)))
        return DCVC_ERR_ARG;
    return st(compact(vals, esize, cond, N, out, block_offsets, count, S(stream)));
}

int dcvc_conditional_index_noalloc(const void* vals, int esize, const uint8_t* cond, long N, void* out, int* count,
                                   void* stream) {
    if ((esize != 1 && esize != 2) || N < 0 || !count || (N > 0 && (!vals || !cond || !out))) {
        set_last_error("dcvc_conditional_index_noalloc: bad argument");
        return DCVC_ERR_ARG;
    }
    if (N > kNoallocMaxElements) {
        set_last_error("dcvc_conditional_index_noalloc: N = " + std::to_string(N) + " exceeds the maximum of " +
                       std::to_string(kNoallocMaxElements) + " elements");
        return DCVC_ERR_ARG;
    }
    int* offs = nullptr;
    const int r = noalloc_scratch(&offs);
    if (r != DCVC_OK) return r;
    // compact() writes every block offset it later reads (count, then scan
    // in place), so the scratch left by an earlier call needs no clearing.
    return st(compact(vals, esize, cond, N, out, offs, count, S(stream)));
}

int dcvc_conditional_recover(const int8_t* dec, const uint8_t* cond, long N, const int* block_offsets, int8_t* out,
                             void* stream) {
    if (N < 0 || (N > 0 && (!dec || !cond || !block_offsets || !out))) return DCVC_ERR_ARG;
    return st(recover(dec, cond, N, block_offsets, out, S(stream)));
}

}  // extern "C"
