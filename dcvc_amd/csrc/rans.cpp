// CPU rANS coder (see rans.h for the stream format it shares with the
// reference's MLCodec_extensions_cpp).
#include "rans.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>

namespace dcvc {

namespace {
constexpr int kProbBits = 16;
constexpr uint32_t kLow = 1u << 23;                 // state lower bound
constexpr int kRenormShift = 23 - kProbBits + 8;    // 15
constexpr uint32_t kProbMask = (1u << kProbBits) - 1;
constexpr int kBypassBits = 2;
constexpr uint32_t kBypassMax = (1u << kBypassBits) - 1;
constexpr int kMaxBypassGroups = 16;

// Output of one sub-stream's encoder: bytes in emission order (the encoder
// produces them last to first, so the stream is their reverse).  `p` and `end` live in locals of the coding loop, so the
// byte stores cannot force reloads of the table pointers.
struct Sink {
    uint8_t* p;
    uint8_t* end;
    inline void push(uint8_t b) { *p++ = b; }
};

// most bytes one symbol can emit: two from the coding step, one per bypass
// bin, and room for the coding step's unconditional store
constexpr int kMaxSymbolBytes = 2 + 2 * kMaxBypassGroups + 4 + 1;

// One coding step, r' = (r / freq) * 2^16 + r % freq + start, with the
// division done by the entry's precomputed reciprocal.  The renormalisation
// emits at most two bytes (r < 2^31, freq << 15 >= 2^15), without branches:
// each byte is stored, and the sink advances only when it is emitted (the
// sink always has room for a symbol's bytes plus one).
inline void enc_put(uint32_t& r, Sink& out, const CdfTable::EncEntry& e) {
    const uint32_t r_max = e.freq << kRenormShift;
    uint32_t b = r >= r_max;
    out.p[0] = static_cast<uint8_t>(r);
    out.p += b;
    r = b ? r >> 8 : r;
    b = r >= r_max;
    out.p[0] = static_cast<uint8_t>(r);
    out.p += b;
    r = b ? r >> 8 : r;
    const uint32_t q = (uint32_t)(((uint64_t)r * e.rcp) >> 32) >> e.shift;
    r = r + e.bias + q * (uint32_t)e.cmpl;
}

// Push one bypass bin: emit a byte first if r >= 2^29 (at most one: r < 2^31).
constexpr uint32_t kBinLimit = 1u << (kProbBits - kBypassBits + kRenormShift);  // 2^29
inline void enc_put_bin(uint32_t& r, Sink& out, uint32_t bin) {
    if (r >= kBinLimit) {
        out.push(static_cast<uint8_t>(r));
        r >>= 8;
    }
    r = (r << kBypassBits) | bin;
}

// Encode one symbol (py_rans rans.cpp:144-180 semantics): values at or
// above max_value code max_value, then the excess in bypass bins.  The
// decoder reads the bin count (kBypassMax bins, then the remainder) and the
// excess low group first, so they are pushed in the reverse order.
inline void enc_symbol(uint32_t& r, Sink& out, int32_t s, int32_t max_value, const CdfTable::EncEntry* row) {
    int32_t value = (s < 0 ? -s : s) * 2 - (s > 0 ? 1 : 0);
    if (__builtin_expect(value >= max_value, 0)) {
        const uint32_t raw = static_cast<uint32_t>(value - max_value);
        value = max_value;
        const int groups = raw ? (32 - __builtin_clz(raw) + kBypassBits - 1) / kBypassBits : 0;
        for (int j = groups - 1; j >= 0; --j) enc_put_bin(r, out, (raw >> (kBypassBits * j)) & kBypassMax);
        const int q = groups / (int)kBypassMax;
        enc_put_bin(r, out, (uint32_t)(groups - q * (int)kBypassMax));
        for (int j = 0; j < q; ++j) enc_put_bin(r, out, kBypassMax);
    }
    const CdfTable::EncEntry& e = row[value];
    if (__builtin_expect(e.freq == 0, 0)) throw std::invalid_argument("symbol has zero frequency in its cdf row");
    enc_put(r, out, e);
}

// Sub-stream reader.  p[n] and p[n + 1] are zero bytes: reads past the end
// return 0 without a branch, and pos keeps counting, so pos > n marks a
// truncated or corrupted stream.
struct Reader {
    const uint8_t* p;
    size_t n;
    size_t pos;
    inline uint32_t peek() const { return p[pos < n ? pos : n]; }
    // the next two bytes, as reads at pos and pos + 1 would return them
    inline void peek2(uint32_t& b0, uint32_t& b1) const {
        const size_t q = pos < n ? pos : n;
        b0 = p[q];
        b1 = p[q + 1];
    }
    inline uint32_t next() {
        const uint32_t v = peek();
        ++pos;
        return v;
    }
    bool overrun() const { return pos > n; }
};

// One decoding step (py_rans.cpp:394-442 semantics).  The symbol search
// starts at the row's bucket table entry; the state renormalisation takes at
// most two bytes (the state stays >= 2^23 and loses at most 16 bits per
// symbol), done without branches on bytes loaded before the state is known.
inline int32_t dec_symbol(uint32_t& r, Reader& in, const int32_t* cdf, const uint8_t* lut, int32_t max_value,
                          bool& bad) {
    const uint32_t cum = r & kProbMask;
    int s = lut[cum >> 8];
    while (s < max_value && (uint32_t)cdf[s + 1] <= cum) ++s;
    const uint32_t start = (uint32_t)cdf[s];
    const uint32_t freq = (uint32_t)cdf[s + 1] - start;
    if (freq == 0 || freq > (1u << kProbBits) || start > cum || cum - start >= freq) {
        bad = true;  // corrupted table row or state
        return 0;
    }
    uint32_t n0, n1;
    in.peek2(n0, n1);
    r = freq * (r >> kProbBits) + cum - start;
    const uint32_t b1 = r < kLow;
    r = b1 ? (r << 8) | n0 : r;
    const uint32_t b2 = r < kLow;
    r = b2 ? (r << 8) | (b1 ? n1 : n0) : r;
    in.pos += b1 + b2;
    int32_t value = s;
    if (__builtin_expect(value == max_value, 0)) {
        auto get_bits = [&]() {
            const uint32_t v = r & kBypassMax;
            r >>= kBypassBits;
            if (r < kLow) r = (r << 8) | in.next();
            return v;
        };
        uint32_t v = get_bits();
        int groups = (int)v;
        while (v == kBypassMax) {
            v = get_bits();
            groups += (int)v;
            if (groups > kMaxBypassGroups) {
                bad = true;
                return 0;
            }
        }
        uint32_t raw = 0;
        for (int j = 0; j < groups; ++j) raw |= get_bits() << (j * kBypassBits);
        value = (int32_t)raw + max_value;
    }
    return (value & 1) ? (value + 1) / 2 : -((value + 1) / 2);
}

void build_table(CdfTable& t, const int32_t* cdfs, int n_cdf, int row_len, const int32_t* sizes) {
    if (n_cdf <= 0 || row_len < 3 || !cdfs || !sizes) throw std::invalid_argument("bad cdf table");
    t.n_cdf = n_cdf;
    t.row_len = row_len;
    t.cdf.assign(cdfs, cdfs + (size_t)n_cdf * row_len);
    t.max_value.resize(n_cdf);
    for (int i = 0; i < n_cdf; ++i) {
        // cdf_length - 2, held as int8 (py_rans.cpp:266, 377): lengths of 258..385
        // wrap to 0..127 as there; lengths of 130..257 wrap negative, which the
        // reference would use as a table index, so they are rejected
        const int mv = (int8_t)(sizes[i] - 2);
        if (sizes[i] < 2 || mv < 0 || mv + 2 > row_len) throw std::invalid_argument("bad cdf length");
        t.max_value[i] = mv;
    }
    t.lut.assign((size_t)n_cdf * 256, 0);
    for (int i = 0; i < n_cdf; ++i) {
        const int32_t* c = cdfs + (size_t)i * row_len;
        int s = 0;
        for (int b = 0; b < 256; ++b) {
            while (s < t.max_value[i] && (uint32_t)c[s + 1] <= (uint32_t)b << 8) ++s;
            t.lut[(size_t)i * 256 + b] = (uint8_t)s;  // s <= max_value <= 127
        }
    }
    t.enc.assign((size_t)n_cdf * row_len, CdfTable::EncEntry{0, 0, 0, 0, 0, 0});
    for (int i = 0; i < n_cdf; ++i) {
        const int32_t* c = cdfs + (size_t)i * row_len;
        for (int j = 0; j + 1 < row_len; ++j) {
            CdfTable::EncEntry& e = t.enc[(size_t)i * row_len + j];
            const uint32_t start = (uint32_t)c[j] & 0xFFFF;
            const uint32_t f = (uint32_t)(c[j + 1] - c[j]) & 0xFFFF;
            e.freq = f;
            if (f == 0) continue;  // never coded (padding or an unused row tail)
            e.cmpl = (uint16_t)((1u << kProbBits) - f);
            if (f == 1) {
                // q = r - 1 (r >= 1): r' = r * 2^16 + start
                e.rcp = ~0u;
                e.shift = 0;
                e.bias = start + (1u << kProbBits) - 1;
            } else {
                uint32_t sh = 0;
                while (f > (1u << sh)) ++sh;
                e.rcp = (uint32_t)(((1ull << (sh + 31)) + f - 1) / f);
                e.shift = (uint8_t)(sh - 1);
                e.bias = start;
            }
        }
    }
}

// split [0, n) into `parts` ranges the way the reference does (py_rans.cpp:182-190)
inline void part_range(long n, int parts, int i, long& begin, long& count) {
    const long size0 = n / parts;
    begin = size0 * i;
    count = (i == parts - 1) ? n - size0 * (parts - 1) : size0;
}

int shared_trailing_bytes(const std::vector<uint8_t>& a, const std::vector<uint8_t>& b) {
    // py_rans.cpp:93-111: shared trailing zero bytes, or 1 if the last bytes match
    const int na = (int)a.size(), nb = (int)b.size();
    const int check = std::min({na, nb, 8});
    int same = 0;
    for (int i = 0; i < check; ++i) {
        if (a[na - 1 - i] != 0 || b[nb - 1 - i] != 0) break;
        ++same;
    }
    if (same == 0 && na > 0 && nb > 0 && a[na - 1] == b[nb - 1]) same = 1;
    return same;
}
}  // namespace

int compute_ec_parallel(long symbol_count) {
    const long n = symbol_count / kMinSymbolsPerStream;
    return (int)std::max(1L, std::min((long)kMaxEcParallel, n));
}

std::vector<uint32_t> pmf_to_quantized_cdf(const std::vector<float>& pmf) {
    constexpr uint32_t kTotal = 1u << kProbBits;
    const size_t n = pmf.size();
    std::vector<uint32_t> cdf(n + 1, 0);
    for (size_t i = 0; i < n; ++i) cdf[i + 1] = static_cast<uint32_t>((double)(pmf[i] * (float)kTotal) + 0.5);
    int total = 0;
    for (uint32_t v : cdf) total += (int)v;
    if (total <= 0) throw std::invalid_argument("pmf sums to zero");
    for (auto& v : cdf) v = static_cast<uint32_t>(((uint64_t)kTotal * v) / (uint32_t)total);
    for (size_t i = 1; i < cdf.size(); ++i) cdf[i] += cdf[i - 1];
    cdf.back() = kTotal;
    // give every symbol a nonzero frequency by taking one count from the
    // smallest frequency that can spare it (ryg_rans' scheme)
    for (size_t i = 0; i + 1 < cdf.size(); ++i) {
        if (cdf[i] + 1 > cdf[i + 1]) {
            uint32_t best_freq = ~0u;
            long best = -1;
            for (size_t j = 0; j + 1 < cdf.size(); ++j) {
                const uint32_t f = cdf[j + 1] - cdf[j];
                if (f >= 2 && f < best_freq) {
                    best_freq = f;
                    best = (long)j;
                }
            }
            if (best < 0) throw std::runtime_error("pmf_to_quantized_cdf: no frequency to steal");
            if ((size_t)best < i) {
                for (size_t j = best + 1; j <= i; ++j) cdf[j] -= 1;
            } else {
                for (size_t j = i + 1; j <= (size_t)best; ++j) cdf[j] += 1;
            }
        }
    }
    return cdf;
}

// ---------------------------------------------------------------------------
WorkerPool::WorkerPool(int n) {
    for (int i = 0; i < n; ++i) m_threads.emplace_back([this, i] { loop(i); });
}

WorkerPool::~WorkerPool() {
    {
        std::lock_guard<std::mutex> lk(m_mu);
        m_stop = true;
    }
    m_cv_start.notify_all();
    for (auto& t : m_threads) t.join();
}

void WorkerPool::loop(int id) {
    uint64_t seen = 0;
    for (;;) {
        const std::function<void(int)>* fn;
        {
            std::unique_lock<std::mutex> lk(m_mu);
            m_cv_start.wait(lk, [&] { return m_stop || (m_gen != seen && id < m_active); });
            if (m_stop) return;
            seen = m_gen;
            fn = m_fn;
        }
        (*fn)(id);
        {
            std::lock_guard<std::mutex> lk(m_mu);
            if (--m_pending == 0) m_cv_done.notify_all();
        }
    }
}

void WorkerPool::run(int n, const std::function<void(int)>& fn) {
    if (n <= 1) {
        if (n == 1) fn(0);
        return;
    }
    if (n > (int)m_threads.size()) throw std::invalid_argument("too many parallel jobs");
    std::unique_lock<std::mutex> lk(m_mu);
    m_fn = &fn;
    m_active = n;
    m_pending = n;
    ++m_gen;
    m_cv_start.notify_all();
    m_cv_done.wait(lk, [&] { return m_pending == 0; });
    m_active = 0;
}

// ---------------------------------------------------------------------------
RansEncoder::RansEncoder() : m_pool(kMaxEcParallel) {}

void RansEncoder::set_cdf(const int32_t* cdfs, int n_cdf, int row_len, const int32_t* sizes, int index) {
    if (index < 0 || index > 1) throw std::invalid_argument("cdf index must be 0 (z) or 1 (y)");
    build_table(m_tab[index], cdfs, n_cdf, row_len, sizes);
}

void RansEncoder::reset() {
    for (auto& t : m_tasks) t.clear();
    m_out.clear();
}

void RansEncoder::set_parallel(int n) {
    if (n < 1 || n > kMaxEcParallel) throw std::invalid_argument("entropy coder parallelism must be 1..8");
    m_parallel = n;
}

void RansEncoder::encode_y(const int16_t* symbols, long n) {
    if (!m_tab[1].valid()) throw std::logic_error("y cdf table not set");
    for (int i = 0; i < m_parallel; ++i) {
        Task t;
        t.y = symbols;
        part_range(n, m_parallel, i, t.begin, t.count);
        m_tasks[i].push_back(t);
    }
}

void RansEncoder::encode_z(const int8_t* symbols, long n, int cdf_offset, int ch) {
    if (!m_tab[0].valid()) throw std::logic_error("z cdf table not set");
    if (ch <= 0 || cdf_offset < 0 || cdf_offset + ch > m_tab[0].n_cdf) throw std::invalid_argument("bad z cdf range");
    for (int i = 0; i < m_parallel; ++i) {
        Task t;
        t.z = symbols;
        t.cdf_offset = cdf_offset;
        t.ch = ch;
        part_range(n, m_parallel, i, t.begin, t.count);
        m_tasks[i].push_back(t);
    }
}

void RansEncoder::flush() {
    const int n = m_parallel;
    std::vector<std::string> errors(n);
    m_pool.run(n, [&](int i) {
        try {
            long total = 0;
            for (auto& t : m_tasks[i]) total += t.count;
            // emitted bytes; grown when fewer than kMaxSymbolBytes are left
            std::vector<uint8_t> emitted((size_t)total * 2 + 4 * kMaxSymbolBytes);
            Sink out{emitted.data(), emitted.data() + emitted.size()};
            auto room = [&]() {
                if (out.end - out.p >= kMaxSymbolBytes) return;
                const size_t used = (size_t)(out.p - emitted.data());
                emitted.resize(emitted.size() * 2);
                out = Sink{emitted.data() + used, emitted.data() + emitted.size()};
            };
            uint32_t r = kLow;
            for (const Task& t : m_tasks[i]) {
                if (t.y) {
                    const CdfTable& tab = m_tab[1];
                    const CdfTable::EncEntry* ent = tab.enc.data();
                    const int32_t* max_value = tab.max_value.data();
                    const size_t row_len = (size_t)tab.row_len;
                    const int n_cdf = tab.n_cdf;
                    for (long k = t.begin + t.count - 1; k >= t.begin; --k) {
                        const int16_t c = t.y[k];
                        const int idx = c & 0xFF;
                        if (idx >= n_cdf) throw std::invalid_argument("y cdf index out of range");
                        room();
                        enc_symbol(r, out, static_cast<int8_t>(c >> 8), max_value[idx], ent + idx * row_len);
                    }
                } else if (t.count > 0) {
                    const CdfTable& tab = m_tab[0];
                    const CdfTable::EncEntry* ent = tab.enc.data();
                    const int32_t* max_value = tab.max_value.data();
                    const size_t row_len = (size_t)tab.row_len;
                    // channel of the last symbol, then counting down
                    int c = (int)((t.begin + t.count - 1) % t.ch);
                    for (long k = t.begin + t.count - 1; k >= t.begin; --k) {
                        const int idx = c + t.cdf_offset;
                        room();
                        enc_symbol(r, out, t.z[k], max_value[idx], ent + idx * row_len);
                        c = (c == 0) ? t.ch - 1 : c - 1;
                    }
                }
            }
            const size_t used = (size_t)(out.p - emitted.data());
            std::vector<uint8_t>& s = m_sub[i];
            s.resize(used + 4);
            s[0] = (uint8_t)r;
            s[1] = (uint8_t)(r >> 8);
            s[2] = (uint8_t)(r >> 16);
            s[3] = (uint8_t)(r >> 24);
            std::reverse_copy(emitted.begin(), emitted.begin() + used, s.begin() + 4);
        } catch (const std::exception& e) {
            errors[i] = e.what();
        }
    });
    for (auto& t : m_tasks) t.clear();
    for (auto& e : errors)
        if (!e.empty()) throw std::runtime_error(e);

    // merge (py_rans.cpp:234-327)
    m_out.clear();
    if (n == 1) {
        m_out = m_sub[0];
        return;
    }
    const int pairs = n / 2;
    const bool tail = (n % 2) != 0;
    std::vector<int> same(pairs), group(pairs);
    for (int p = 0; p < pairs; ++p) {
        same[p] = shared_trailing_bytes(m_sub[2 * p], m_sub[2 * p + 1]);
        group[p] = (int)m_sub[2 * p].size() + (int)m_sub[2 * p + 1].size() - same[p];
    }
    const int n_off = pairs - 1 + (tail ? 1 : 0);
    size_t total = (size_t)n_off * 4;
    for (int g : group) total += g;
    if (tail) total += m_sub[n - 1].size();
    m_out.resize(total);
    int cum = group[0];
    for (int k = 0; k < n_off; ++k) {
        const int32_t v = cum;
        std::memcpy(&m_out[(size_t)k * 4], &v, 4);
        if (k + 1 < pairs) cum += group[k + 1];
    }
    size_t pos = (size_t)n_off * 4;
    for (int p = 0; p < pairs; ++p) {
        const auto& a = m_sub[2 * p];
        const auto& b = m_sub[2 * p + 1];
        std::copy(a.begin(), a.end(), m_out.begin() + pos);
        std::reverse_copy(b.begin(), b.end() - same[p], m_out.begin() + pos + a.size());
        pos += group[p];
    }
    if (tail) std::copy(m_sub[n - 1].begin(), m_sub[n - 1].end(), m_out.begin() + pos);
}

// ---------------------------------------------------------------------------
RansDecoder::RansDecoder() : m_pool(kMaxEcParallel) {}

void RansDecoder::set_cdf(const int32_t* cdfs, int n_cdf, int row_len, const int32_t* sizes, int index) {
    if (index < 0 || index > 1) throw std::invalid_argument("cdf index must be 0 (z) or 1 (y)");
    build_table(m_tab[index], cdfs, n_cdf, row_len, sizes);
}

void RansDecoder::set_parallel(int n) {
    if (n < 1 || n > kMaxEcParallel) throw std::invalid_argument("entropy coder parallelism must be 1..8");
    m_parallel = n;
}

void RansDecoder::set_stream(const uint8_t* data, long size) {
    // py_rans.cpp:490-570
    const int n = m_parallel;
    auto init = [](Sub& s) {
        s.size = s.bytes.size();
        s.bytes.resize(s.size + 2, 0);  // reads past the end: zero, and pos > size flags them
        if (s.size < 4) throw std::runtime_error("truncated entropy-coded stream");
        s.pos = 4;
        s.overrun = false;
        s.state = (uint32_t)s.bytes[0] | ((uint32_t)s.bytes[1] << 8) | ((uint32_t)s.bytes[2] << 16) |
                  ((uint32_t)s.bytes[3] << 24);
    };
    if (size < 0 || (size > 0 && !data)) throw std::invalid_argument("bad stream");
    if (n == 1) {
        m_sub[0].bytes.assign(data, data + size);
        init(m_sub[0]);
        return;
    }
    const int pairs = n / 2;
    const bool tail = (n % 2) != 0;
    const int n_off = pairs - 1 + (tail ? 1 : 0);
    const long header = (long)n_off * 4;
    if (size < header) throw std::runtime_error("truncated entropy-coded stream header");
    std::vector<long> off(n_off);
    for (int k = 0; k < n_off; ++k) {
        int32_t v;
        std::memcpy(&v, data + (size_t)k * 4, 4);
        off[k] = v;
    }
    const uint8_t* payload = data + header;
    const long psize = size - header;
    std::vector<long> gstart(pairs), gsize(pairs);
    gstart[0] = 0;
    gsize[0] = (n_off > 0) ? off[0] : psize;
    for (int p = 1; p < pairs; ++p) {
        gstart[p] = off[p - 1];
        const long end = (p < n_off) ? (p == n_off - 1 && tail ? off[n_off - 1] : off[p])
                                     : (tail ? off[n_off - 1] : psize);
        gsize[p] = end - off[p - 1];
    }
    for (int p = 0; p < pairs; ++p) {
        if (gstart[p] < 0 || gsize[p] < 0 || gstart[p] + gsize[p] > psize)
            throw std::runtime_error("corrupted entropy-coded stream header");
        const uint8_t* g = payload + gstart[p];
        m_sub[2 * p].bytes.assign(g, g + gsize[p]);
        m_sub[2 * p + 1].bytes.assign(std::make_reverse_iterator(g + gsize[p]), std::make_reverse_iterator(g));
        init(m_sub[2 * p]);
        init(m_sub[2 * p + 1]);
    }
    if (tail) {
        const long ts = off[n_off - 1];
        if (ts < 0 || ts > psize) throw std::runtime_error("corrupted entropy-coded stream header");
        m_sub[n - 1].bytes.assign(payload + ts, payload + psize);
        init(m_sub[n - 1]);
    }
}

// Decode out[begin, begin + count) from one sub-stream; row_of(k) gives the
// cdf row of symbol k (< 0: index out of range, stop).  The table pointers
// are held in locals: the int8 stores to out may alias anything, and would
// otherwise force them to be reloaded from the table after every symbol.
// Returns false on an out-of-range row.
template <class RowOf>
bool decode_run(RansDecoder::Sub& s, const CdfTable& tab, long begin, long count, RowOf row_of, int8_t* out) {
    const int32_t* cdf = tab.cdf.data();
    const uint8_t* lut = tab.lut.data();
    const int32_t* max_value = tab.max_value.data();
    const size_t row_len = (size_t)tab.row_len;
    Reader rd{s.bytes.data(), s.size, s.pos};
    uint32_t r = s.state;
    bool corrupt = false, in_range = true;
    for (long k = begin; k < begin + count && !corrupt; ++k) {
        const int idx = row_of(k);
        if (idx < 0) {
            in_range = false;
            break;
        }
        out[k] = (int8_t)dec_symbol(r, rd, cdf + idx * row_len, lut + (size_t)idx * 256, max_value[idx], corrupt);
    }
    s.state = r;
    s.pos = rd.pos;
    s.overrun = s.overrun || corrupt || rd.overrun();
    return in_range;
}

void RansDecoder::decode_z(long n, int cdf_offset, int ch, int8_t* out) {
    const CdfTable& tab = m_tab[0];
    if (!tab.valid()) throw std::logic_error("z cdf table not set");
    if (ch <= 0 || cdf_offset < 0 || cdf_offset + ch > tab.n_cdf) throw std::invalid_argument("bad z cdf range");
    m_pool.run(m_parallel, [&](int i) {
        long begin, count;
        part_range(n, m_parallel, i, begin, count);
        if (count <= 0) return;
        int c = (int)(begin % ch);  // channel of symbol k, advanced once per symbol
        auto row_of = [&](long) {
            const int idx = c + cdf_offset;
            c = (c + 1 == ch) ? 0 : c + 1;
            return idx;
        };
        decode_run(m_sub[i], tab, begin, count, row_of, out);
    });
    bool bad = false;
    for (int i = 0; i < m_parallel; ++i) bad |= m_sub[i].overrun;
    if (bad) throw std::runtime_error("entropy-coded stream is truncated or corrupted");
}

void RansDecoder::decode_y(const uint8_t* indexes, long n, int8_t* out) {
    const CdfTable& tab = m_tab[1];
    if (!tab.valid()) throw std::logic_error("y cdf table not set");
    std::vector<char> bad_idx(m_parallel, 0);
    const int n_cdf = tab.n_cdf;
    m_pool.run(m_parallel, [&](int i) {
        long begin, count;
        part_range(n, m_parallel, i, begin, count);
        auto row_of = [&](long k) {
            const int idx = indexes[k];
            return idx < n_cdf ? idx : -1;
        };
        if (!decode_run(m_sub[i], tab, begin, count, row_of, out)) bad_idx[i] = 1;
    });
    for (int i = 0; i < m_parallel; ++i) {
        if (bad_idx[i]) throw std::invalid_argument("y cdf index out of range");
        if (m_sub[i].overrun) throw std::runtime_error("entropy-coded stream is truncated or corrupted");
    }
}

}  // namespace dcvc
