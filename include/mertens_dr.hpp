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
// host threads.  Calls are serialised on one process-wide context.
#ifndef MERTENS_DR_HIP_DROPIN_HPP
#define MERTENS_DR_HIP_DROPIN_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <mutex>
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

class Context {
public:
    ~Context() {
        if (ctx_) mdr_shutdown(ctx_);
    }
    mdr_ctx* get() {
        std::lock_guard<std::mutex> lk(m_);
        if (!ctx_) {
            int dev = 0;
            if (const char* e = std::getenv("MERTENS_HIP_DEVICE")) dev = std::atoi(e);
            const mdr_status s = mdr_init(&ctx_, dev);
            if (s != MDR_OK) {
                ctx_ = nullptr;
                throw std::runtime_error("mertens_dr (HIP): mdr_init failed with status " + std::to_string(s));
            }
        }
        return ctx_;
    }

private:
    std::mutex m_;
    mdr_ctx* ctx_ = nullptr;
};

inline Context& context() {
    static Context c;
    return c;
}

inline void check(mdr_status s, mdr_ctx* c, const char* what) {
    if (s != MDR_OK)
        throw std::runtime_error(std::string("mertens_dr (HIP): ") + what + ": " + mdr_last_error(c));
}

}  // namespace hip_detail

class DelégliseRivatEngine {
public:
    static constexpr int64 COMPRESSED_THRESHOLD = 100000000000LL;   // 10^11
    static constexpr int64 DENSE_LIMIT = 128000000LL;

    // mertens_dr.hpp:914-923
    static inline double choose_cx(int64 X) noexcept {
        if (const char* e = std::getenv("MERTENS_CX")) {
            const double v = std::atof(e);
            if (v > 0.1 && v < 5.0) return v;
        }
        if (X >= 1000000000000000LL) return 1.35;
        if (X >= 10000000000000LL) return 1.15;
        if (X >= 100000000000LL) return 0.95;
        return 0.70;
    }

    // mertens_dr.hpp:931-962
    static inline int64 choose_sieve_limit(int64 X, int threads) {
        (void)threads;
        if (X <= 50000000LL) return X;
        const double ll = std::log(std::max(2.0, std::log(static_cast<double>(X))));
        double fx = 0.32;
        if (const char* e = std::getenv("MERTENS_FX")) {
            const double v = std::atof(e);
            if (v >= 0.1 && v <= 3.0) fx = v;
        }
        int64 u = static_cast<int64>(fx * std::pow(static_cast<double>(X) / ll, 2.0 / 3.0));
        const int64 S = isqrt(X);
        if (u < 3 * S) u = 3 * S;
        int64 cap = 1500000000LL;
        if (const char* e = std::getenv("MERTENS_U_CAP")) {
            const int64 v = static_cast<int64>(std::atoll(e));
            if (v >= 100000000LL && v <= 4000000000LL) cap = v;
        }
        return u < cap ? u : cap;
    }

    // mertens_dr.hpp:1401-1419
    static inline int64 choose_compressed_limit(int64 X) {
        if (X < 50000000LL) return X;
        const double ll = std::log(std::max(2.0, std::log(static_cast<double>(X))));
        double fx = 0.95;
        if (X >= 1000000000000000LL) fx = 1.25;
        if (X >= 10000000000000000LL) fx = 1.50;
        int64 u = static_cast<int64>(fx * std::pow(static_cast<double>(X) / ll, 2.0 / 3.0));
        const int64 S = isqrt(X);
        if (u < 3 * S) u = 3 * S;
        u = ((u + 255) / 256) * 256;
        int64 cap = 8000000000LL;
        if (const char* e = std::getenv("MERTENS_CHMT_U_CAP")) {
            const int64 v = std::atoll(e);
            if (v >= 100000000LL && v <= 10000000000LL) cap = v;
        }
        const int64 ends = static_cast<int64>(choose_cx(X) * std::sqrt(static_cast<double>(X))) + 1000;
        const int64 need = (S + 1000 > ends) ? S + 1000 : ends;
        if (cap < need) cap = need;
        return u < cap ? u : cap;
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
        mdr_ctx* c = hip_detail::context().get();
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
        mdr_ctx* c = hip_detail::context().get();
        const mdr_opts o = hip_detail::opts_from_env();
        long long out = 0;
        hip_detail::check(mdr_mertens(c, X, &o, &out, stats), c, "compute_mertens");
        return out;
    }

    // Many X in one call sharing one table.
    static std::vector<int64> compute_mertens_batch(const std::vector<int64>& xs, mdr_stats* stats = nullptr) {
        mdr_ctx* c = hip_detail::context().get();
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

}  // namespace mertens_dr

#endif
