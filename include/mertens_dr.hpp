// Drop-in replacement for the reference's mertens_dr.hpp entry point, backed
// by the MI355X engine (libmdr_hip.so, C ABI in mertens_hip.h).
//
// Same namespace, class and static methods as the reference:
//   mertens_dr::DelégliseRivatEngine::compute_mertens(int64 X, int num_threads = 0)
//       (mertens_dr.hpp:968)
//   mertens_dr::DelégliseRivatEngine::compute_compressed(int64 X, int threads)
//       (mertens_dr.hpp:1442)
//   choose_cx / choose_sieve_limit / choose_compressed_limit /
//   use_compressed_default / use_chmt_env / COMPRESSED_THRESHOLD / DENSE_LIMIT
//       (mertens_dr.hpp:914-962, :1390-1435)
//   isqrt (mertens_dr.hpp:65)
// and the same environment controls, read on every call with the same
// accepted ranges: MERTENS_USE_CHMT, MERTENS_CX, MERTENS_CHMT_U_CAP,
// MERTENS_FX, MERTENS_U_CAP.  Errors raise std::runtime_error (the reference's
// CUDA engine does the same through CUDA_CHECK, mertens_cuda.cuh:20-28).
//
// Parameters: with none of those five variables set, a call takes the GPU's
// autotuned cutoff and split (the result is the same integer; M(X) does not
// depend on them).  Setting any of them selects the reference's parameter
// rules with that control applied, as the reference would use it.
//
// Extra controls of the GPU engine (none changes a result):
//   MERTENS_HIP_DEVICE   GPU ordinal (default 0)
//   MERTENS_HIP_TUNED=1  the autotuned cutoff/split even with reference
//                        controls set; =0 the reference's rules always
//   MERTENS_HIP_LEARNED=1  learned (MLP) chunk partition instead of static
//
// num_threads is accepted for interface parity; the GPU engine does not use
// host threads.  Calls are re-entrant: each takes an engine context (stream,
// buffers, cached table) from a process-wide pool, creating up to
// MERTENS_HIP_CONTEXTS of them (default 4) and waiting for a free one beyond
// that, so concurrent host threads run concurrently on the GPU.  Each context
// caches the last table it built, so concurrent calls at large X hold one
// table each; MERTENS_HIP_CONTEXTS=1 serialises all calls on one context.
//
// Table objects (mertens_dr.hpp:370-457 SieveTable, :662-752
// CompressedMertensTable with BitChunk): built by the GPU sieve, then held in
// host vectors laid out exactly as the reference's (M, mu, mu_odd_store;
// dense_M, mu, chunks, checkpoints), so get_M / get_mu / data / dense_data /
// reconstruct_compressed read them as the reference does.  hip::DeviceTable
// keeps a table resident on the GPU instead (reads and gathers through the C
// ABI).
#ifndef MERTENS_DR_HIP_DROPIN_HPP
#define MERTENS_DR_HIP_DROPIN_HPP

#include <algorithm>
#include <bit>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "mertens_hip.h"

namespace mertens_dr {

using int64 = long long;
using uint64 = unsigned long long;

inline int64 isqrt(int64 n) noexcept {
    if (n <= 0) return 0;
    uint64 s = static_cast<uint64>(std::sqrt(static_cast<double>(n)));
    const uint64 t = static_cast<uint64>(n);
    while ((s + 1) * (s + 1) <= t) ++s;
    while (s * s > t) --s;
    return static_cast<int64>(s);
}

namespace hip_detail {

// Environment parsing with the reference's rules.
inline mdr_opts opts_from_env(int strategy_override = -1) {
    mdr_opts o;
    mdr_default_opts(&o);
    bool reference_control = false;
    for (const char* v : {"MERTENS_USE_CHMT", "MERTENS_CX", "MERTENS_CHMT_U_CAP", "MERTENS_FX", "MERTENS_U_CAP"})
        if (std::getenv(v)) reference_control = true;
    o.params = reference_control ? MDR_PARAMS_REFERENCE : MDR_PARAMS_TUNED;
    if (const char* e = std::getenv("MERTENS_USE_CHMT")) {
        const bool on = e[0] == '1' || e[0] == 'y' || e[0] == 'Y' || e[0] == 't' || e[0] == 'T';
        o.strategy = on ? MDR_STRATEGY_STREAMING : MDR_STRATEGY_DENSE;
    }
    if (const char* e = std::getenv("MERTENS_CX")) {
        const double v = std::atof(e);
        if (v > 0.1 && v < 5.0) o.cx = v;
    }
    if (const char* e = std::getenv("MERTENS_CHMT_U_CAP")) {
        const long long v = std::atoll(e);
        if (v >= 100000000LL && v <= 10000000000LL) o.u_cap = v;
    }
    if (const char* e = std::getenv("MERTENS_FX")) {
        o.fx_set = 1;
        const double v = std::atof(e);
        if (v >= 0.1 && v <= 3.0) o.fx = v;
    }
    if (const char* e = std::getenv("MERTENS_U_CAP")) {
        o.u_cap_dense_set = 1;
        const long long v = std::atoll(e);
        if (v >= 100000000LL && v <= 4000000000LL) o.u_cap_dense = v;
    }
    if (const char* e = std::getenv("MERTENS_HIP_TUNED")) o.params = e[0] == '1' ? MDR_PARAMS_TUNED : MDR_PARAMS_REFERENCE;
    if (const char* e = std::getenv("MERTENS_HIP_LEARNED"))
        o.partition = e[0] == '1' ? MDR_PARTITION_LEARNED : MDR_PARTITION_STATIC;
    if (strategy_override >= 0) o.strategy = strategy_override;
    return o;
}

// Engine contexts shared by the calling threads: a call leases an idle one,
// or creates one while fewer than MERTENS_HIP_CONTEXTS exist, or waits.
class ContextPool {
public:
    ~ContextPool() {
        for (mdr_ctx* c : all_) mdr_shutdown(c);
    }
    mdr_ctx* acquire() {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            if (!idle_.empty()) {
                mdr_ctx* c = idle_.back();
                idle_.pop_back();
                return c;
            }
            if (all_.size() < cap()) break;
            cv_.wait(lk);
        }
        int dev = 0;
        if (const char* e = std::getenv("MERTENS_HIP_DEVICE")) dev = std::atoi(e);
        mdr_ctx* c = nullptr;
        const mdr_status s = mdr_init(&c, dev);
        if (s != MDR_OK)
            throw std::runtime_error("mertens_dr (HIP): mdr_init failed with status " + std::to_string(s));
        all_.push_back(c);
        return c;
    }
    void release(mdr_ctx* c) {
        {
            std::lock_guard<std::mutex> lk(m_);
            idle_.push_back(c);
        }
        cv_.notify_one();
    }

private:
    static size_t cap() {
        long v = 4;
        if (const char* e = std::getenv("MERTENS_HIP_CONTEXTS")) v = std::atol(e);
        return (size_t)std::clamp(v, 1L, 64L);
    }
    std::mutex m_;
    std::condition_variable cv_;
    std::vector<mdr_ctx*> all_, idle_;
};

inline ContextPool& pool() {
    static ContextPool p;
    return p;
}

// A context held for the duration of one call.
class Lease {
public:
    Lease() : c_(pool().acquire()) {}
    ~Lease() { pool().release(c_); }
    Lease(const Lease&) = delete;
    Lease& operator=(const Lease&) = delete;
    mdr_ctx* get() const { return c_; }

private:
    mdr_ctx* c_;
};

inline void check(mdr_status s, mdr_ctx* c, const char* what) {
    if (s != MDR_OK)
        throw std::runtime_error(std::string("mertens_dr (HIP): ") + what + ": " + mdr_last_error(c));
}

}  // namespace hip_detail

class DelégliseRivatEngine {
public:
    static constexpr int64 COMPRESSED_THRESHOLD = 100000000000LL;   // 10^11
    static constexpr int64 DENSE_LIMIT = 128000000LL;

    // The reference's parameter helpers (mertens_dr.hpp:914-923, :931-962,
    // :1401-1419): the environment read with the reference's ranges, the rule
    // applied by the library (mdr_param_*), so the drop-in and the engine
    // share one statement of each rule.
    static inline double choose_cx(int64 X) noexcept {
        const mdr_opts o = hip_detail::opts_from_env();
        return mdr_param_cx(X, &o);
    }

    static inline int64 choose_sieve_limit(int64 X, int threads) {
        (void)threads;
        const mdr_opts o = hip_detail::opts_from_env();
        return mdr_param_dense_cutoff(X, &o);
    }

    static inline int64 choose_compressed_limit(int64 X) {
        const mdr_opts o = hip_detail::opts_from_env();
        return mdr_param_streaming_cutoff(X, &o);
    }

    // mertens_dr.hpp:1421-1428 (x86-64 behaviour)
    static bool use_compressed_default(int64 X) noexcept {
        return X >= COMPRESSED_THRESHOLD && !std::getenv("MERTENS_FX") && !std::getenv("MERTENS_U_CAP");
    }

    // mertens_dr.hpp:1430-1435
    static bool use_chmt_env() noexcept {
        if (const char* e = std::getenv("MERTENS_USE_CHMT"))
            return e[0] == '1' || e[0] == 'y' || e[0] == 'Y' || e[0] == 't' || e[0] == 'T';
        return false;
    }

    // mertens_dr.hpp:968.  Exact M(X) for any X (X < 1 gives 0).
    static int64 compute_mertens(int64 X, int num_threads = 0) {
        mdr_stats st;
        return compute_mertens_stats(X, num_threads, &st);
    }

    // mertens_dr.hpp:1442.  The streaming (compressed-table) strategy.
    static int64 compute_compressed(int64 X, int threads) {
        (void)threads;
        if (X <= 100000LL) return compute_mertens(X, threads);
        const hip_detail::Lease lease;
        mdr_ctx* c = lease.get();
        const mdr_opts o = hip_detail::opts_from_env(MDR_STRATEGY_FORCE_STREAMING);
        long long out = 0;
        hip_detail::check(mdr_mertens(c, X, &o, &out, nullptr), c, "compute_compressed");
        return out;
    }

    // GPU extensions -------------------------------------------------------------
    // compute_mertens with the engine's timing record (TimingStats analogue,
    // mertens_cuda.cuh:410-416).
    static int64 compute_mertens_stats(int64 X, int num_threads, mdr_stats* stats) {
        (void)num_threads;
        const hip_detail::Lease lease;
        mdr_ctx* c = lease.get();
        const mdr_opts o = hip_detail::opts_from_env();
        long long out = 0;
        hip_detail::check(mdr_mertens(c, X, &o, &out, stats), c, "compute_mertens");
        return out;
    }

    // Many X in one call sharing one table.
    static std::vector<int64> compute_mertens_batch(const std::vector<int64>& xs, mdr_stats* stats = nullptr) {
        const hip_detail::Lease lease;
        mdr_ctx* c = lease.get();
        const mdr_opts o = hip_detail::opts_from_env();
        std::vector<int64> out(xs.size());
        std::vector<long long> in(xs.begin(), xs.end());
        std::vector<long long> res(xs.size());
        hip_detail::check(mdr_mertens_batch(c, in.data(), in.size(), &o, res.data(), stats), c,
                          "compute_mertens_batch");
        for (size_t i = 0; i < xs.size(); ++i) out[i] = res[i];
        return out;
    }
};

// ASCII spelling for code bases that avoid non-ASCII identifiers.
using DelegliseRivatEngine = DelégliseRivatEngine;

// mu(n) from an odd-only table mu_odd[k] = mu(2k - 1) (mertens_dr.hpp:130-141):
// odd n read directly, n = 2 (mod 4) is minus the odd half, 4 | n gives 0.
inline int8_t mu_from_odd(int64 n, const int8_t* mu_odd) noexcept {
    if (n & 1) return mu_odd[(n + 1) >> 1];
    if ((n >> 1) & 1) return static_cast<int8_t>(-mu_odd[((n >> 1) + 1) >> 1]);
    return 0;
}

namespace hip {

// A table resident on the GPU (mdr_table_create): M(n) for n <= U, mu(n) for
// n <= mu_limit, read or gathered into host memory on demand.  A table made
// by the constructor has an engine context of its own (created with it, shut
// down with it), so any number of tables can be held while other calls lease
// the pooled contexts.  DeviceTable::leased builds on a pooled context instead
// and holds it until the table is destroyed: for tables that live within one
// call (the table classes below copy one out and drop it).
class DeviceTable {
public:
    DeviceTable(int64 U, int64 u_dense, int64 mu_limit) : own_(new_context()), c_(own_) {
        try {
            create(U, u_dense, mu_limit);
        } catch (...) {
            mdr_shutdown(own_);
            throw;
        }
    }
    static DeviceTable leased(int64 U, int64 u_dense, int64 mu_limit) {
        return DeviceTable(Leased{}, U, u_dense, mu_limit);
    }
    ~DeviceTable() {
        mdr_table_free(t_);
        if (own_) mdr_shutdown(own_);
    }
    DeviceTable(const DeviceTable&) = delete;
    DeviceTable& operator=(const DeviceTable&) = delete;

    int64 limit() const { return info(0); }
    int64 dense_limit() const { return info(1); }
    int64 mu_limit() const { return info(2); }
    // M(first .. first + count - 1), mu(...)
    void copy_M(int64 first, int64 count, int32_t* out) const {
        hip_detail::check(mdr_table_copy_M(t_, first, count, out), c_, "mdr_table_copy_M");
    }
    void copy_mu(int64 first, int64 count, int8_t* out) const {
        hip_detail::check(mdr_table_copy_mu(t_, first, count, out), c_, "mdr_table_copy_mu");
    }
    std::vector<int32_t> get_M(const std::vector<int64>& q) const {
        std::vector<long long> in(q.begin(), q.end());
        std::vector<int32_t> out(q.size());
        hip_detail::check(mdr_table_gather_M(t_, in.data(), in.size(), out.data()), c_, "mdr_table_gather_M");
        return out;
    }
    int32_t get_M(int64 q) const { return get_M(std::vector<int64>{q})[0]; }
    int8_t get_mu(int64 j) const {
        int8_t m = 0;
        copy_mu(j, 1, &m);
        return m;
    }
    // the compressed tier above u_dense_ref in the reference's chunk layout
    void copy_chunks(int64 u_dense_ref, int64 first_chunk, int64 count, void* chunks, int32_t* checkpoints) const {
        hip_detail::check(mdr_table_copy_chunks(t_, u_dense_ref, first_chunk, count, chunks, checkpoints),
                          c_, "mdr_table_copy_chunks");
    }

private:
    struct Leased {};
    DeviceTable(Leased, int64 U, int64 u_dense, int64 mu_limit) {
        c_ = lease_.emplace().get();
        create(U, u_dense, mu_limit);
    }
    static mdr_ctx* new_context() {
        int dev = 0;
        if (const char* e = std::getenv("MERTENS_HIP_DEVICE")) dev = std::atoi(e);
        mdr_ctx* c = nullptr;
        const mdr_status s = mdr_init(&c, dev);
        if (s != MDR_OK)
            throw std::runtime_error("mertens_dr (HIP): mdr_init failed with status " + std::to_string(s));
        return c;
    }
    void create(int64 U, int64 u_dense, int64 mu_limit) {
        hip_detail::check(mdr_table_create(c_, U, u_dense, mu_limit, &t_), c_, "mdr_table_create");
    }
    int64 info(int k) const {
        long long v[3] = {0, 0, 0};
        hip_detail::check(mdr_table_info(t_, &v[0], &v[1], &v[2]), c_, "mdr_table_info");
        return v[k];
    }
    std::optional<hip_detail::Lease> lease_;   // leased tables: the pooled context
    mdr_ctx* own_ = nullptr;                  // constructed tables: their own context
    mdr_ctx* c_ = nullptr;
    mdr_table* t_ = nullptr;
};

}  // namespace hip

namespace hip_detail {
// Host copy of M(0..last) narrowed to int16 (the dense tables' type; |M(n)| <
// 2^15 for every n <= 7.6e9), through a bounded int32 staging buffer.
inline void copy_M16(const hip::DeviceTable& t, int64 last, int16_t* out) {
    constexpr int64 STAGE = int64(1) << 24;
    std::vector<int32_t> buf((size_t)std::min(last + 1, STAGE));
    for (int64 k = 0; k <= last; k += STAGE) {
        const int64 m = std::min(STAGE, last + 1 - k);
        t.copy_M(k, m, buf.data());
        for (int64 i = 0; i < m; ++i) out[k + i] = static_cast<int16_t>(buf[(size_t)i]);
    }
}
}  // namespace hip_detail

// SieveTable(limit, threads, max_mu, retain_mu_odd, max_M) (mertens_dr.hpp:
// 376-457): mu(t) for t <= mu_limit = min(max_mu, u) (u when max_mu is 0),
// M(t) as int16 for t <= m_limit = min(max_M, u) (u when max_M < 0), and with
// retain_mu_odd the odd-only table mu_odd[k] = mu(2k - 1), k <= (u + 1) / 2.
class SieveTable {
public:
    int64 u;
    int64 mu_limit;
    int64 m_limit;
    bool retain_mu_odd;
    std::vector<int8_t> mu;
    std::vector<int16_t> M;
    std::vector<int8_t> mu_odd_store;

    explicit SieveTable(int64 limit, int num_threads = 0, int64 max_mu = 0, bool retain_mu_odd_ = false,
                        int64 max_M_ = -1)
        : u(limit), retain_mu_odd(retain_mu_odd_) {
        (void)num_threads;
        if (u < 0) u = 0;
        mu_limit = max_mu > 0 ? std::min(max_mu, u) : u;
        m_limit = max_M_ >= 0 ? std::min(max_M_, u) : u;
        mu.assign(static_cast<size_t>(mu_limit + 1), 0);
        M.assign(static_cast<size_t>(m_limit + 1), 0);
        const int64 need_mu = retain_mu_odd ? u : mu_limit;
        const int64 U = retain_mu_odd ? u : std::max(m_limit, mu_limit);
        const hip::DeviceTable t = hip::DeviceTable::leased(U, U, need_mu);
        hip_detail::copy_M16(t, m_limit, M.data());
        if (!retain_mu_odd) {
            t.copy_mu(0, mu_limit + 1, mu.data());
            return;
        }
        std::vector<int8_t> all((size_t)u + 1);
        t.copy_mu(0, u + 1, all.data());
        std::memcpy(mu.data(), all.data(), (size_t)mu_limit + 1);
        const int64 half = (u + 1) / 2;
        mu_odd_store.assign((size_t)half + 1, 0);
        for (int64 k = 1; k <= half; ++k) mu_odd_store[(size_t)k] = all[(size_t)(2 * k - 1)];
    }

    int32_t get_M(int64 t) const noexcept { return static_cast<int32_t>(M[(size_t)t]); }
    int16_t operator()(int64 t) const noexcept { return M[(size_t)t]; }
    int8_t get_mu(int64 t) const noexcept { return t <= mu_limit ? mu[(size_t)t] : 0; }
    const int16_t* data() const noexcept { return M.data(); }
    const int8_t* mu_odd_data() const noexcept { return mu_odd_store.empty() ? nullptr : mu_odd_store.data(); }
    int64 mu_odd_size() const noexcept {
        return mu_odd_store.empty() ? 0 : static_cast<int64>(mu_odd_store.size()) - 1;
    }
};

// The reference's compressed-tier record (mertens_dr.hpp:670-673): bit i of
// nonzero / sign is mu != 0 / mu < 0 of the chunk's value i.
struct alignas(64) BitChunk {
    uint8_t nonzero[32];
    uint8_t sign[32];
};

// CompressedMertensTable(total_limit, dense_limit, threads, max_mu)
// (mertens_dr.hpp:675-752): M as int16 for q <= u_dense, 256-value BitChunks
// with four int32 checkpoints (M just before each 64-value word) above it,
// mu(j) for j <= mu_limit = min(u_total, max_mu, or u_dense when max_mu is 0).
class CompressedMertensTable {
public:
    int64 u_dense;
    int64 u_total;
    int64 mu_limit;
    std::vector<int16_t> dense_M;
    std::vector<int8_t> mu;
    int64 compressed_size;
    int64 num_chunks;
    std::vector<BitChunk> chunks;
    std::vector<int32_t> checkpoints;

    CompressedMertensTable(int64 total_limit, int64 dense_limit, int num_threads = 0, int64 max_mu = 0)
        : u_dense(std::min(total_limit, dense_limit)), u_total(total_limit) {
        (void)num_threads;
        mu_limit = std::min(u_total, max_mu > 0 ? max_mu : u_dense);
        compressed_size = std::max<int64>(0, u_total - u_dense);
        num_chunks = compressed_size > 0 ? (compressed_size + 255) / 256 : 0;
        const int64 top = num_chunks ? u_dense : u_total;   // the dense tier's last value
        const hip::DeviceTable t = hip::DeviceTable::leased(u_total, u_dense, std::max<int64>(mu_limit, 0));
        dense_M.assign(static_cast<size_t>(top + 1), 0);
        hip_detail::copy_M16(t, top, dense_M.data());
        mu.assign(static_cast<size_t>(std::max<int64>(mu_limit, 0) + 1), 0);
        t.copy_mu(0, (int64)mu.size(), mu.data());
        chunks.resize(static_cast<size_t>(num_chunks));
        checkpoints.resize(static_cast<size_t>(num_chunks * 4));
        if (num_chunks) t.copy_chunks(u_dense, 0, num_chunks, chunks.data(), checkpoints.data());
    }

    int32_t get_M(int64 q) const noexcept {
        if (q <= u_dense) return static_cast<int32_t>(dense_M[static_cast<size_t>(q)]);
        return reconstruct_compressed(q);
    }
    int8_t get_mu(int64 j) const noexcept { return j <= mu_limit ? mu[static_cast<size_t>(j)] : 0; }
    const int16_t* dense_data() const noexcept { return dense_M.data(); }
    void prefetch_M(int64 q) const noexcept {
        if (q <= u_dense) __builtin_prefetch(&dense_M[static_cast<size_t>(q)], 0, 0);
        else if (q <= u_total) {
            const int64 off = q - u_dense - 1;
            __builtin_prefetch(&chunks[static_cast<size_t>(off >> 8)], 0, 0);
            __builtin_prefetch(&checkpoints[static_cast<size_t>(off >> 6)], 0, 0);
        }
    }
    // M(q) = checkpoint of q's word + popcount of the word's nonzero bits up to
    // q - 2 * popcount of its sign bits up to q (mertens_dr.hpp:741-752)
    int32_t reconstruct_compressed(int64 q) const noexcept {
        const int64 off = q - u_dense - 1;
        const int64 w = off >> 6;
        const BitChunk& ch = chunks[static_cast<size_t>(w >> 2)];
        const size_t byte = static_cast<size_t>((w & 3) * 8);
        uint64_t nz, neg;
        std::memcpy(&nz, ch.nonzero + byte, 8);
        std::memcpy(&neg, ch.sign + byte, 8);
        const uint64_t upto = ~uint64_t(0) >> (63 - (off & 63));
        return checkpoints[static_cast<size_t>(w)] + std::popcount(nz & upto) - 2 * std::popcount(neg & upto);
    }
};

}  // namespace mertens_dr

#endif
