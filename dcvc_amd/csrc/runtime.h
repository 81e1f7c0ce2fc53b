// Host runtime of the codec proxies: device memory, weight loading with the
// reference's load-time transforms, and the layer building blocks
// (DepthConvBlock, stride-2 convs, sub-pixel up-sampling) over NHWC buffers.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <mutex>
#include <vector>

#include "capi_util.h"
#include "common.h"
#include "dwconv.h"
#include "elementwise.h"
#include "ffn.h"
#include "gemm.h"

namespace dcvc {

void hip_check(hipError_t e, const char* what);

// Named float32 host tensors handed over by set_param (the model state dict).
struct HostTensor {
    std::vector<int64_t> shape;
    std::vector<float> data;
    int64_t numel() const {
        int64_t n = 1;
        for (auto d : shape) n *= d;
        return n;
    }
};
using HostDict = std::map<std::string, HostTensor>;
const HostTensor& need(const HostDict& d, const std::string& key);
bool has(const HostDict& d, const std::string& key);

// Owner of device allocations (freed together).
struct Conv1x1;

class DevMem {
public:
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem();
    void* alloc(size_t bytes);
    template <class T>
    T* alloc_n(size_t n) {
        return static_cast<T*>(alloc(n * sizeof(T)));
    }
    void free_all();
    size_t bytes() const { return m_bytes; }
    // 1x1 convolutions whose weights live here.  Their bf16 / fp8 copies
    // (the precision form) are made on the first request, on the null
    // stream, so proxies that only run fp16 hold no extra copies.
    void register_conv(Conv1x1* c) { m_convs.push_back(c); }
    void ensure_low_precision();

private:
    std::vector<void*> m_blocks;
    size_t m_bytes = 0;
    std::vector<Conv1x1*> m_convs;
    bool m_lowp = false;
    std::mutex m_lowp_mu;
};

f16* upload_f16(DevMem& mem, const std::vector<float>& v);
float* upload_f32(DevMem& mem, const std::vector<float>& v);

// A view of an NHWC activation buffer (see Act in elementwise.h).
inline Act make_act(f16* p, int n_img, int H, int W, int C, int ld) {
    Act a;
    a.p = p;
    a.n_img = n_img;
    a.H = H;
    a.W = W;
    a.C = C;
    a.ld = ld;
    return a;
}
inline Act channels(const Act& a, int c0, int c) {
    Act v = a;
    v.p = a.p + c0;
    v.C = c;
    return v;
}

// Per-QP channel scale: row *qp of a [64, C] fp16 table.  round_first rounds
// the block output to fp16 before the multiply, which reproduces a separate
// fp16 multiply_with_broadcast after the block (the HT-L proxy's order) inside
// the fused epilogue.
struct QuantSel {
    const f16* table = nullptr;
    int stride = 0;
    bool round_first = false;
};

// Scratch activations shared by all layers of one proxy (one stream, so the
// layers never use them concurrently).
struct Scratch {
    f16* t1 = nullptr;   // DCB: dc.0 output / ffn.0 (quad) output
    f16* t2 = nullptr;   // DCB: depthwise output
    f16* so = nullptr;   // DCB: dc branch output when the block shortcut keeps x
    f16* x1 = nullptr;   // sequence temporaries
    f16* x2 = nullptr;
    size_t halves = 0;   // capacity of each buffer
};

struct Exec {
    hipStream_t stream = nullptr;
    const int* d_qp = nullptr;
    Scratch* sc = nullptr;
    int prec = PREC_F16;  // GemmPrec of the 1x1-conv products
};

// e4m3 operands of a product in the precision form (Exec::prec == PREC_FP8;
// ignored otherwise): A read from an e4m3 copy of the activation, and an
// e4m3 copy of the output stored beside, or instead of, the fp16 output.
struct Fp8Io {
    const uint8_t* a8 = nullptr;
    int lda8 = 0;
    uint8_t* d8 = nullptr;
    int ldd8 = 0;
    bool d16 = true;  // false: the e4m3 copy is the only output
};

struct Conv1x1 {
    f16* w = nullptr;
    float* b = nullptr;
    void* w_bf16 = nullptr;   // low-precision copies (PREC_BF16 / PREC_FP8), DevMem::ensure_low_precision
    void* w_fp8 = nullptr;
    float* w_scale = nullptr;
    int N = 0, K = 0;
    bool quad = false;
    // weight [N,K,1,1] (quad: [4N,K,1,1] in the reference ffn.0 order)
    void load(DevMem& mem, const HostDict& d, const std::string& prefix, bool quad_ = false);
    void load_raw(DevMem& mem, const std::vector<float>& wt, const std::vector<float>& bias, int N_, int K_);
    // bf16 and fp8 (per-row scale) copies of the fp16 weight, on the null stream
    void make_low_precision(DevMem& mem);
    void run(const Exec& e, const Act& a, const Act& dst, int epi, const Act* s1 = nullptr,
             const Act* s2 = nullptr, const QuantSel* q = nullptr, const Fp8Io* io = nullptr) const;
};

struct DW3x3 {
    f16* w = nullptr;  // [9][C]
    int C = 0;
    void load(DevMem& mem, const HostDict& d, const std::string& key);
    // y8: optional e4m3 copy of the output (ldy8 bytes per pixel); y16 false
    // stores only that copy
    void run(const Exec& e, const Act& in, const Act& out, uint8_t* y8 = nullptr, int ldy8 = 0,
             bool y16 = true) const;
};

// DepthConvBlock (layers.py:128-159) as the proxy runs it
// (layers_proxy.cpp:71-100, 160-206): the depthwise bias is folded into dc.3,
// ffn.0 runs as the quad GEMM, and the tail is shortcut / shortcut2 /
// shortcut-with-quant.
struct DCB {
    bool has_adaptor = false, shortcut = false;
    int cin = 0, cout = 0, cmid = 0;
    Conv1x1 ad, dc0, dc3, ffn0, ffn2;
    DW3x3 dw;
    void load(DevMem& mem, const HostDict& d, const std::string& p, bool shortcut_ = false);
    // out may equal in unless the block has an adaptor; q multiplies the
    // block output per channel (conv1x1_bias_shortcut_with_quant).
    // In precision fp8 with e4m3 activations, chain_out also stores an e4m3
    // copy of the output in the Exec's scratch, and chain_in makes dc.0 read
    // that copy: set it only when `in` is the output of the previous call on
    // the same Exec, made with chain_out.  Both flags are ignored otherwise.
    void forward(const Exec& e, const Act& in, const Act& out, const QuantSel* q = nullptr, bool chain_in = false,
                 bool chain_out = false) const;
};

// A chain of DepthConvBlocks: block 0 reads `in`, intermediate results live in
// `tmp`, the last block writes `out` (in and out may be the same view only if
// block 0 has no adaptor).
struct DCBSeq {
    std::vector<DCB> blocks;
    void load(DevMem& mem, const HostDict& d, const std::string& prefix, int n, int first = 0);
    void forward(const Exec& e, const Act& in, const Act& out, const Act& tmp, const QuantSel* q = nullptr) const;
};

// k x k stride-2 conv (k = 2: 1x1 over pixel_unshuffle(x, 2), the
// ResidualBlockWithStride2 down path; k = 3: Conv2d(3, stride 2, pad 1)).
struct ConvS2 {
    f16* w = nullptr;  // [N][k*k][Cin]
    float* b = nullptr;
    int N = 0, Cin = 0, k = 0;
    void load(DevMem& mem, const HostDict& d, const std::string& prefix, int k_);
    void run(const Exec& e, const Act& in, const Act& out) const;
};

// SubpelConv2x: conv (1x1, or 3x3 pad 1) to 4*Cout then pixel_shuffle(2)
// (layers.py:92-103; the proxy's transposed_conv / conv2d + pixel_shuffle_2,
// layers_proxy.cpp:269-276), with the shuffle fused into the GEMM store.
struct Subpel {
    f16* w = nullptr;  // [4][Cout][k*k][Cin] sub-pixel major
    float* b = nullptr;
    int Cout = 0, Cin = 0, k = 1;
    void load(DevMem& mem, const HostDict& d, const std::string& prefix);
    void run(const Exec& e, const Act& in, const Act& out) const;
};

struct RBS2 {  // ResidualBlockWithStride2 (layers.py:176-188)
    ConvS2 down;
    DCB conv;
    void load(DevMem& mem, const HostDict& d, const std::string& p, bool shortcut);
    void forward(const Exec& e, const Act& in, const Act& out) const;
};

struct RBU {  // ResidualBlockUpsample (layers.py:162-173)
    Subpel up;
    DCB conv;
    void load(DevMem& mem, const HostDict& d, const std::string& p, bool shortcut);
    void forward(const Exec& e, const Act& in, const Act& out) const;
};

// hipGraph wrapper: capture a lambda once, replay afterwards.
class Graph {
public:
    ~Graph() { reset(); }
    void reset();
    bool ready() const { return m_exec != nullptr; }
    template <class Fn>
    void run(hipStream_t s, bool use_graph, Fn&& fn) {
        if (!use_graph) {
            fn();
            return;
        }
        if (!m_exec) capture(s, fn);
        hip_check(hipGraphLaunch(m_exec, s), "hipGraphLaunch");
    }

private:
    template <class Fn>
    void capture(hipStream_t s, Fn&& fn) {
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            fn();
        } catch (...) {
            hipGraph_t g = nullptr;
            hipStreamEndCapture(s, &g);
            if (g) hipGraphDestroy(g);
            throw;
        }
        hipGraph_t g = nullptr;
        hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
        hipError_t err = hipGraphInstantiate(&m_exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        hip_check(err, "hipGraphInstantiate");
    }
    hipGraphExec_t m_exec = nullptr;
};

void set_device_int(int* dst, int v, hipStream_t s);
// dst[0..n) = v[0..n), n <= 16, ordered on stream s (a kernel, so graph-safe)
struct DeviceInts {
    static constexpr int kMax = 16;
    int n = 0;
    int v[kMax] = {};
};
void set_device_ints(int* dst, const int* v, int n, hipStream_t s);

// Launch trace: one record per GEMM / conv / depthwise launch issued by the
// layer classes while tracing is on (used to build the per-chunk layer table).
// A timed trace brackets each launch with HIP events on its stream.
struct TraceRecord {
    int kind;   // 0 1x1 GEMM, 1 stride-2 conv, 2 sub-pixel conv, 3 depthwise 3x3,
                // 4 fused FFN (N = block channels, K = hidden channels)
    int epi;    // GemmEpi for kinds 0-2 and 4 (the tail)
    int M, N, K;            // N = output channels (quad: 4N weight rows)
    long long flops;        // algorithmic multiply-adds x 2
    long long bytes;        // algorithmic HBM bytes: inputs + weights + outputs once
    long long ns;           // launch duration (timed traces), else -1
    // tuning key and choice: A / D modes, operand precision, tile
    // configuration (depthwise: strip rows; fused FFN: -1), and the output
    // grid (images, rows, columns)
    int amode = 0, dmode = 0, prec = 0, cfg = -1;
    int n_img = 0, H = 0, W = 0;
};
void trace_record(const TraceRecord& r, hipStream_t s);
// Sets the tuning key of the last record.
void trace_key(int amode, int dmode, int prec, int cfg, int n_img, int H, int W);
void trace_launched(hipStream_t s);
bool trace_enabled();
void trace_set(bool on, bool timed);
// Waits for the timed events, returns the records and ends the trace.
std::vector<TraceRecord> trace_collect();

}  // namespace dcvc
