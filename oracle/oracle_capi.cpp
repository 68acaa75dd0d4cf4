// C entry points of the CPU oracle for the Python tests (ctypes).
// TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke entry point and the
// CPU-baseline leg of bench.py to check results, never by the product.
#include "mdr_oracle.hpp"

#include <chrono>
#include <cstring>

using namespace mdr_oracle;

namespace {
Controls make_controls(const char* use_chmt, const char* cx, const char* chmt_u_cap,
                       const char* fx, const char* u_cap) {
    Controls c;
    c.use_chmt = use_chmt;
    c.cx = cx;
    c.chmt_u_cap = chmt_u_cap;
    c.fx = fx;
    c.u_cap = u_cap;
    return c;
}
void split128(i128 v, i64* lo, i64* hi) {
    *lo = (i64)(u64)v;
    *hi = (i64)(v >> 64);
}
}  // namespace

extern "C" {

i64 orc_isqrt(i64 n) { return isqrt(n); }

i64 orc_mertens(i64 X, int threads, const char* use_chmt, const char* cx, const char* chmt_u_cap,
                const char* fx, const char* u_cap) {
    return mertens(X, threads, make_controls(use_chmt, cx, chmt_u_cap, fx, u_cap));
}

i64 orc_mertens_compressed(i64 X, int threads) { return mertens_compressed(X, threads, Controls{}); }

i64 orc_mertens_small(i64 X) { return mertens_small(X); }

i64 orc_mertens_recurrence(i64 X, i64 base) { return mertens_recurrence(X, base); }

// Wall time of one oracle M(X) call in seconds; the result goes to *out.
double orc_time_mertens(i64 X, int threads, i64* out) {
    const auto t0 = std::chrono::steady_clock::now();
    *out = mertens(X, threads, Controls{});
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

void orc_plan(i64 X, const char* use_chmt, const char* cx, const char* chmt_u_cap, const char* fx,
              const char* u_cap, int* streaming, i64* u, i64* u_dense, i64* mu_limit, i64* N,
              double* cx_out) {
    const Plan p = plan_for(X, make_controls(use_chmt, cx, chmt_u_cap, fx, u_cap));
    *streaming = p.streaming;
    *u = p.u;
    *u_dense = p.u_dense;
    *mu_limit = p.mu_limit;
    *N = p.N;
    *cx_out = p.cx;
}

void orc_mobius_eratosthenes(i64 limit, int8_t* out) {
    const std::vector<int8_t> mu = mobius_eratosthenes(limit);
    std::memcpy(out, mu.data(), mu.size());
}

void orc_mobius_bytelog_block(i64 low, i64 len, i64 sieve_limit, int8_t* out) {
    mobius_bytelog_block(low, len, base_primes(isqrt(sieve_limit)), out);
}

void orc_mobius_bytelog_odd(i64 half_u, i64 u, int threads, int8_t* out) {
    mobius_bytelog_odd(out, half_u, u, threads);
}

int orc_finalize_byte(int acc, i64 n) { return finalize_byte((uint8_t)acc, n); }

i64 orc_piece_closed(int piece, i64 q) { return piece_closed(piece, q); }
i64 orc_piece_lut(int piece, i64 q) { return piece_lut(piece, q); }

void orc_s2_interval(int piece, i64 first, i64 last, i64 y, const int8_t* mu, i64* lo, i64* hi) {
    split128(s2_interval(piece, first, last, y, mu), lo, hi);
}

void orc_dense_table(i64 limit, int threads, i64 max_mu, i64 max_M, int8_t* mu_out, int16_t* M_out) {
    DenseTable t(limit, threads, max_mu, max_M);
    std::memcpy(mu_out, t.mu.data(), t.mu.size());
    std::memcpy(M_out, t.M.data(), t.M.size() * sizeof(int16_t));
}

// Builds the streamed table and answers get_M / get_mu for n = 0..query_limit.
void orc_compressed_table(i64 total, i64 dense, int threads, i64 max_mu, i64 query_limit,
                          int32_t* M_out, int8_t* mu_out, i64* mu_limit_out) {
    CompressedTable t(total, dense, threads, max_mu);
    for (i64 n = 0; n <= query_limit; ++n) {
        M_out[n] = t.get_M(n);
        mu_out[n] = t.get_mu(n);
    }
    *mu_limit_out = t.mu_limit;
}

// G(y) and G(y) - G(y/2) over a caller-provided full int16 prefix table M
// (valid up to u) and mu table.
void orc_g_single(i64 y, double cx, i64 u, const int16_t* M, i64 M_limit, const int8_t* mu, i64* lo,
                  i64* hi) {
    MLookup L{M, M_limit, nullptr};
    split128(g_single(y, cx, u, L, mu), lo, hi);
}

void orc_g_comb2(i64 y, double cx, i64 u, const int16_t* M, i64 M_limit, const int8_t* mu, i64* lo,
                 i64* hi) {
    MLookup L{M, M_limit, nullptr};
    split128(g_comb2(y, cx, u, L, mu), lo, hi);
}

void orc_g_brute(i64 y, i64 u, const int32_t* prefix, i64 prefix_limit, i64* lo, i64* hi) {
    std::vector<int32_t> p(prefix, prefix + prefix_limit + 1);
    split128(g_brute(y, u, p), lo, hi);
}

}  // extern "C"
