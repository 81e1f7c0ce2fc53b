// CPU rANS entropy coder with the stream format of the reference's
// MLCodec_extensions_cpp (src/cpp/py_rans/rans.cpp:13-181, py_rans.cpp:35-249):
//   * 32-bit state, 16-bit frequency precision, byte-wise renormalisation
//     with lower bound 2^23, state flushed as 4 little-endian bytes;
//   * signed symbols mapped to 0, 1, -1, 2, -2, ... (value = 2|s| - [s > 0]);
//     values >= max_value escape through 2-bit bypass bins (count in unary
//     groups of 3, then the raw value, least significant group first);
//   * up to 8 independent sub-streams, paired (2k, 2k+1) into one buffer with
//     the second stream reversed and overlapping on shared trailing bytes,
//     preceded by int32 cumulative group offsets.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <vector>

namespace dcvc {

constexpr int kMaxEcParallel = 8;           // py_rans.h:15
constexpr int kMinSymbolsPerStream = 32768;  // def_const.h:18

int compute_ec_parallel(long symbol_count);  // dmc_common.cpp:31-35

// pmf -> 16-bit quantised CDF of pmf.size()+1 entries (py_rans.cpp:35-94)
std::vector<uint32_t> pmf_to_quantized_cdf(const std::vector<float>& pmf);

// Fixed worker pool: run(n, fn) executes fn(0..n-1) on n threads and waits.
class WorkerPool {
public:
    explicit WorkerPool(int n);
    ~WorkerPool();
    void run(int n, const std::function<void(int)>& fn);

private:
    void loop(int id);
    std::vector<std::thread> m_threads;
    std::mutex m_mu;
    std::condition_variable m_cv_start, m_cv_done;
    const std::function<void(int)>* m_fn = nullptr;
    int m_active = 0;
    int m_pending = 0;
    uint64_t m_gen = 0;
    bool m_stop = false;
};

struct CdfTable {
    // one row per distribution: cdf values and the escape symbol max_value
    std::vector<int32_t> cdf;      // n_cdf x row_len
    std::vector<int32_t> max_value;
    // encoder entry per (row, symbol), n_cdf x row_len: the division by freq
    // as a multiply (Alverson's integer reciprocal), x / freq ==
    // (x * rcp >> 32) >> shift for x < 2^31, and the coding step
    // x' = (x / f) * 2^16 + x % f + start == x + bias + (x / f) * cmpl
    struct EncEntry {
        uint32_t rcp, bias, freq;
        uint16_t cmpl;  // 2^16 - freq (freq >= 1 wherever it is used)
        uint8_t shift, pad;
    };
    std::vector<EncEntry> enc;
    // decoder search start: lut[row * 256 + b] is the symbol whose interval
    // holds cum = 256 b, so the search for cum walks forward from
    // lut[cum >> 8] (usually zero steps) instead of from symbol 0
    std::vector<uint8_t> lut;
    int row_len = 0;
    int n_cdf = 0;
    bool valid() const { return n_cdf > 0; }
};

// Encoder: queue y / z symbol batches, then flush() codes every sub-stream
// in parallel and merges them.  Symbol arrays must stay alive until flush().
class RansEncoder {
public:
    RansEncoder();
    void set_cdf(const int32_t* cdfs, int n_cdf, int row_len, const int32_t* sizes, int index);
    void reset();
    void set_parallel(int n);
    int parallel() const { return m_parallel; }
    void encode_y(const int16_t* symbols, long n);           // (sym << 8) | cdf index
    void encode_z(const int8_t* symbols, long n, int cdf_offset, int ch);
    void flush();                                            // codes and merges
    const std::vector<uint8_t>& stream() const { return m_out; }

private:
    struct Task {
        const int16_t* y = nullptr;
        const int8_t* z = nullptr;
        long begin = 0, count = 0;
        int cdf_offset = 0, ch = 1;
    };
    CdfTable m_tab[2];  // 0: z (bit estimator), 1: y (gaussian)
    int m_parallel = 1;
    std::vector<Task> m_tasks[kMaxEcParallel];
    std::vector<uint8_t> m_sub[kMaxEcParallel];
    std::vector<uint8_t> m_out;
    WorkerPool m_pool;
};

// Decoder: set_stream() splits the merged buffer into sub-streams; each
// decode call runs the sub-decoders in parallel and returns when done.
// A truncated or corrupted stream raises std::runtime_error instead of
// reading out of bounds.
class RansDecoder {
public:
    RansDecoder();
    void set_cdf(const int32_t* cdfs, int n_cdf, int row_len, const int32_t* sizes, int index);
    void set_parallel(int n);
    void set_stream(const uint8_t* data, long size);
    void decode_z(long n, int cdf_offset, int ch, int8_t* out);
    void decode_y(const uint8_t* indexes, long n, int8_t* out);

    // one sub-stream being decoded
    struct Sub {
        std::vector<uint8_t> bytes;  // the sub-stream, then two zero bytes
        size_t size = 0;             // sub-stream bytes
        size_t pos = 0;
        uint32_t state = 0;
        bool overrun = false;
    };

private:
    CdfTable m_tab[2];
    int m_parallel = 1;
    Sub m_sub[kMaxEcParallel];
    WorkerPool m_pool;
};

}  // namespace dcvc
