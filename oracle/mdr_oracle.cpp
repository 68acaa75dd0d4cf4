// CPU oracle implementation. TEST INFRASTRUCTURE ONLY (see mdr_oracle.hpp).
#include "mdr_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace mdr_oracle {

namespace {
int resolve_threads(int t) {
#ifdef _OPENMP
    return t > 0 ? t : omp_get_max_threads();
#else
    (void)t;
    return 1;
#endif
}
int thread_id() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}
constexpr i64 TWO53 = 9007199254740992LL;
}  // namespace

// mertens_dr.hpp:65-74. Double estimate, then exact correction with
// 128-bit squares so no intermediate can overflow.
i64 isqrt(i64 n) {
    if (n <= 0) return 0;
    i64 s = (i64)std::sqrt((double)n);
    while ((i128)(s + 1) * (s + 1) <= n) ++s;
    while ((i128)s * s > n) --s;
    return s;
}

unsigned bit_width_u64(u64 v) { return v ? 64u - (unsigned)__builtin_clzll(v) : 0u; }

unsigned bytelog_weight(i64 p) { return bit_width_u64((u64)(p - 1)) | 1u; }

// mertens_dr.hpp:143-171 closed form (test_sieve.cpp:62-67): a square flag
// gives 0; otherwise the low bit is the parity of the sieved prime factors and
// an accumulator not above floor(log2 n) means one unsieved prime remains.
int8_t finalize_byte(uint8_t acc, i64 n) {
    if (acc & 0x80) return 0;
    const unsigned threshold = bit_width_u64((u64)n) - 1;
    const int sign = (acc & 1) ? -1 : 1;
    return (int8_t)(acc > threshold ? sign : -sign);
}

std::vector<int8_t> mobius_eratosthenes(i64 limit) {
    std::vector<int8_t> mu((size_t)limit + 1, 1);
    mu[0] = 0;
    std::vector<bool> composite((size_t)limit + 1, false);
    for (i64 p = 2; p <= limit; ++p) {
        if (composite[(size_t)p]) continue;
        for (i64 n = p; n <= limit; n += p) {
            composite[(size_t)n] = true;
            mu[(size_t)n] = (int8_t)-mu[(size_t)n];
        }
        for (i64 n = p * p; p <= limit / p && n <= limit; n += p * p) mu[(size_t)n] = 0;
    }
    return mu;
}

BasePrimes base_primes(i64 r) {
    BasePrimes bp;
    if (r < 2) return bp;
    std::vector<uint8_t> comp((size_t)r + 1, 0);
    for (i64 i = 2; i <= r; ++i) {
        if (comp[(size_t)i]) continue;
        bp.p.push_back(i);
        bp.w.push_back((uint8_t)bytelog_weight(i));
        for (i64 j = i * i; j <= r; j += i) comp[(size_t)j] = 1;
    }
    return bp;
}

// Full-number byte-log block (mertens_dr.hpp:814-839): every base prime adds
// its odd weight to each multiple; every prime square sets the 0x80 flag.
void mobius_bytelog_block(i64 low, i64 len, const BasePrimes& bp, int8_t* out) {
    if (len <= 0) return;
    std::vector<uint8_t> acc((size_t)len, 0);
    const i64 high = low + len - 1;
    for (size_t i = 0; i < bp.p.size(); ++i) {
        const i64 p = bp.p[i];
        const i64 sq = p * p;
        if (sq <= high) {
            for (i64 m = ((low + sq - 1) / sq) * sq; m <= high; m += sq) acc[(size_t)(m - low)] |= 0x80;
        }
        for (i64 m = ((low + p - 1) / p) * p; m <= high; m += p) acc[(size_t)(m - low)] += bp.w[i];
    }
    for (i64 i = 0; i < len; ++i) {
        const i64 n = low + i;
        int8_t v = finalize_byte(acc[(size_t)i], n);
        if (n == 1) v = 1;
        if (n == 2) v = -1;
        out[i] = v;
    }
}

// Odd-only table (mertens_dr.hpp:173-296): index k holds mu(2k-1).
void mobius_bytelog_odd(int8_t* mu_odd, i64 half_u, i64 u, int threads) {
    threads = resolve_threads(threads);
    const i64 r = (i64)std::sqrt((double)u) + 1;
    const BasePrimes bp = base_primes(r);
    const i64 BLOCK = 1 << 17;
    const i64 nblocks = (half_u + BLOCK - 1) / BLOCK;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (i64 b = 0; b < nblocks; ++b) {
        const i64 k0 = b * BLOCK + 1;
        const i64 k1 = std::min(half_u, (b + 1) * BLOCK);
        const i64 len = k1 - k0 + 1;
        std::vector<uint8_t> acc((size_t)len, 0);
        const i64 v0 = 2 * k0 - 1, v1 = 2 * k1 - 1;
        for (size_t i = 0; i < bp.p.size(); ++i) {
            const i64 p = bp.p[i];
            if (p == 2) continue;
            const i64 sq = p * p;
            if (sq <= v1) {
                i64 m = ((v0 + sq - 1) / sq) * sq;
                if (!(m & 1)) m += sq;
                for (; m <= v1; m += 2 * sq) acc[(size_t)((m - v0) / 2)] |= 0x80;
            }
            i64 m = ((v0 + p - 1) / p) * p;
            if (!(m & 1)) m += p;
            for (; m <= v1; m += 2 * p) acc[(size_t)((m - v0) / 2)] += bp.w[i];
        }
        for (i64 i = 0; i < len; ++i) mu_odd[k0 + i] = finalize_byte(acc[(size_t)i], v0 + 2 * i);
    }
    mu_odd[0] = 0;
    if (half_u >= 1) mu_odd[1] = 1;
}

int8_t mu_from_odd(i64 n, const int8_t* mu_odd) {
    if (n & 1) return mu_odd[(n + 1) >> 1];
    if ((n >> 1) & 1) return (int8_t)-mu_odd[((n >> 1) + 1) >> 1];
    return 0;
}

// SieveTable (mertens_dr.hpp:376-457).
DenseTable::DenseTable(i64 limit, int threads, i64 max_mu, i64 max_M) : u(limit) {
    threads = resolve_threads(threads);
    mu_limit = max_mu > 0 ? std::min(max_mu, u) : u;
    m_limit = max_M >= 0 ? std::min(max_M, u) : u;
    mu.assign((size_t)mu_limit + 1, 0);
    M.assign((size_t)m_limit + 1, 0);
    const i64 half_u = (u + 1) / 2;
    std::vector<int8_t> odd((size_t)half_u + 1, 0);
    mobius_bytelog_odd(odd.data(), half_u, u, threads);
    for (i64 i = 1; i <= mu_limit; ++i) mu[(size_t)i] = mu_from_odd(i, odd.data());
    int32_t run = 0;
    for (i64 i = 1; i <= m_limit; ++i) {
        run += mu_from_odd(i, odd.data());
        M[(size_t)i] = (int16_t)run;
    }
}

// CompressedMertensTable (mertens_dr.hpp:675-902). Blocks are sieved in
// parallel; block sums are scanned; local prefixes are then made global.
CompressedTable::CompressedTable(i64 total_limit, i64 dense_limit, int threads, i64 max_mu)
    : u_dense(std::min(total_limit, dense_limit)), u_total(total_limit) {
    threads = resolve_threads(threads);
    mu_limit = std::min(u_total, max_mu > 0 ? max_mu : u_dense);
    compressed_size = std::max<i64>(0, u_total - u_dense);
    num_chunks = compressed_size > 0 ? (compressed_size + 255) / 256 : 0;
    if (num_chunks == 0) {
        DenseTable t(u_total, threads, mu_limit);
        dense_M = std::move(t.M);
        mu = std::move(t.mu);
        return;
    }
    dense_M.assign((size_t)u_dense + 1, 0);
    mu.assign((size_t)mu_limit + 1, 0);
    const i64 words = num_chunks * 4;
    nz.assign((size_t)words, 0);
    neg.assign((size_t)words, 0);
    checkpoints.assign((size_t)words, 0);
    const BasePrimes bp = base_primes(isqrt(u_total));
    const i64 BLOCK = 1 << 17;   // a multiple of 256, so chunks never straddle blocks
    const i64 dblocks = (u_dense + BLOCK - 1) / BLOCK;
    const i64 cblocks = (compressed_size + BLOCK - 1) / BLOCK;
    const i64 nblocks = dblocks + cblocks;
    std::vector<i64> bsum((size_t)nblocks + 1, 0);
#pragma omp parallel num_threads(threads)
    {
        std::vector<int8_t> vals((size_t)BLOCK);
#pragma omp for schedule(dynamic, 1)
        for (i64 b = 0; b < nblocks; ++b) {
            const bool dense = b < dblocks;
            const i64 base = dense ? b * BLOCK : u_dense + (b - dblocks) * BLOCK;
            const i64 end = std::min(base + BLOCK, dense ? u_dense : u_total);
            const i64 len = end - base;
            mobius_bytelog_block(base + 1, len, bp, vals.data());
            for (i64 i = 0; i < len; ++i) {
                const i64 n = base + 1 + i;
                if (n <= mu_limit) mu[(size_t)n] = vals[(size_t)i];
            }
            i64 s = 0;
            if (dense) {
                for (i64 i = 0; i < len; ++i) {
                    s += vals[(size_t)i];
                    dense_M[(size_t)(base + 1 + i)] = (int16_t)s;
                }
            } else {
                const i64 off0 = base - u_dense;   // offset of the first value
                for (i64 i = 0; i < len; ++i) {
                    const i64 off = off0 + i;
                    const i64 w = off >> 6;
                    if ((off & 63) == 0) checkpoints[(size_t)w] = (int32_t)s;
                    const int8_t v = vals[(size_t)i];
                    if (v != 0) nz[(size_t)w] |= 1ull << (off & 63);
                    if (v < 0) neg[(size_t)w] |= 1ull << (off & 63);
                    s += v;
                }
            }
            bsum[(size_t)b + 1] = s;
        }
    }
    for (i64 b = 1; b <= nblocks; ++b) bsum[(size_t)b] += bsum[(size_t)b - 1];
#pragma omp parallel for schedule(static) num_threads(threads)
    for (i64 b = 0; b < nblocks; ++b) {
        const i64 add = bsum[(size_t)b];
        if (b < dblocks) {
            const i64 base = b * BLOCK, end = std::min(base + BLOCK, u_dense);
            for (i64 n = base + 1; n <= end; ++n) dense_M[(size_t)n] = (int16_t)(dense_M[(size_t)n] + add);
        } else {
            const i64 w0 = (b - dblocks) * (BLOCK / 64);
            const i64 w1 = std::min(w0 + BLOCK / 64, (compressed_size + 63) / 64);
            for (i64 w = w0; w < w1; ++w) checkpoints[(size_t)w] = (int32_t)(checkpoints[(size_t)w] + add);
        }
    }
}

// mertens_dr.hpp:741-752: checkpoint plus popcounts of the masked planes.
int32_t CompressedTable::get_M(i64 q) const {
    if (q <= u_dense) return dense_M[(size_t)q];
    const i64 off = q - u_dense - 1;
    const i64 w = off >> 6;
    const u64 mask = ~0ull >> (63 - (off & 63));
    return checkpoints[(size_t)w] + __builtin_popcountll(nz[(size_t)w] & mask) -
           2 * __builtin_popcountll(neg[(size_t)w] & mask);
}

// ---- S2 summands ------------------------------------------------------------
// Closed forms: q = floor(y/j) and floor(q/m) = floor(y/(m j)); each form is
// the inclusion-exclusion over the multiples j, 2j, 3j, 6j that stay within
// the A (and B, for the y/2 half) endpoint (mertens_dr.hpp:470-482).
i64 piece_closed(int piece, i64 q) {
    const i64 h = q / 2, t = q / 3, f = q / 4, s = q / 6, w = q / 12;
    switch (piece) {
        case 1: return (q - h - t + s) - (h - f - s + w);
        case 2: return (q - h - t + s) - (h - f - s);
        case 3: return (q - h - t) - (h - f - s);
        case 4: return (q - h - t) - (h - f);
        case 5: return (q - h) - (h - f);
        case 6: return q - 2 * h;
        case 7: return q - h;
        case 8: return q;
        case 9: return q - h - t + s;
        case 10: return q - h - t;
        case 11: return q - h;
        case 12: return q;
    }
    return 0;
}

// The reference's LUT forms (mertens_dr.hpp:462-482), kept only so the tests
// can show the closed forms above equal them.
i64 piece_lut(int piece, i64 q) {
    static const int8_t P1[12] = {0, 1, 0, 0, 0, 1, 1, 2, 2, 2, 1, 2};
    static const int8_t P3[12] = {0, 1, 0, 0, 0, 1, 0, 1, 1, 1, 0, 1};
    static const int8_t P4[12] = {0, 1, 0, 0, 0, 1, -1, 0, 0, 0, -1, 0};
    static const int8_t S1[12] = {0, 1, 1, 1, 1, 2, 2, 3, 3, 3, 3, 4};
    static const int8_t S2[12] = {0, 1, 1, 1, 1, 2, 1, 2, 2, 2, 2, 3};
    const u64 uq = (u64)q;
    const i64 k = (i64)(uq / 12);
    const int r = (int)(uq % 12);
    switch (piece) {
        case 1: return 2 * k + P1[r];
        case 2: return 3 * k + P1[r];   // LUT_P2 equals LUT_P1
        case 3: return k + P3[r];
        case 4: return -k + P4[r];
        case 5: return (q >> 2) + (q & 1);
        case 6: return q & 1;
        case 7: return (q + 1) >> 1;
        case 8: return q;
        case 9: return 4 * k + S1[r];
        case 10: return 2 * k + S2[r];
        case 11: return (q + 1) >> 1;
        case 12: return q;
    }
    return 0;
}

i128 s2_interval(int piece, i64 first, i64 last, i64 y, const int8_t* mu) {
    i128 sum = 0;
    if (first < 1) first = 1;
    for (i64 j = first; j <= last; ++j) {
        if (j % 2 == 0 || j % 3 == 0) continue;
        const int8_t m = mu[j];
        if (m) sum += (i128)m * piece_closed(piece, y / j);
    }
    return sum;
}

// Breakpoints B/6, A/6, B/3, A/3, B/2, A/2, B, A (mertens_dr.hpp:622-642).
i128 s2_combined(i64 y, i64 A, i64 B, const int8_t* mu) {
    const i64 cut[9] = {0, B / 6, A / 6, B / 3, A / 3, B / 2, A / 2, B, A};
    i128 s = 0;
    for (int p = 1; p <= 8; ++p) s += s2_interval(p, cut[p - 1] + 1, cut[p], y, mu);
    return s;
}

// Breakpoints A/6, A/3, A/2, A (mertens_dr.hpp:647-660).
i128 s2_single(i64 y, i64 A, const int8_t* mu) {
    const i64 cut[5] = {0, A / 6, A / 3, A / 2, A};
    i128 s = 0;
    for (int p = 1; p <= 4; ++p) s += s2_interval(8 + p, cut[p - 1] + 1, cut[p], y, mu);
    return s;
}

// ---- parameters ------------------------------------------------------------
double choose_cx(i64 X, const Controls& c) {   // mertens_dr.hpp:914-923
    if (c.cx) {
        const double v = std::atof(c.cx);
        if (v > 0.1 && v < 5.0) return v;
    }
    if (X >= 1000000000000000LL) return 1.35;
    if (X >= 10000000000000LL) return 1.15;
    if (X >= 100000000000LL) return 0.95;
    return 0.70;
}

i64 choose_sieve_limit(i64 X, const Controls& c) {   // mertens_dr.hpp:931-962
    if (X <= 50000000LL) return X;
    const double ll = std::log(std::max(2.0, std::log((double)X)));
    double fx = 0.32;
    if (c.fx) {
        const double v = std::atof(c.fx);
        if (v >= 0.1 && v <= 3.0) fx = v;
    }
    i64 u = (i64)(fx * std::pow((double)X / ll, 2.0 / 3.0));
    const i64 S = isqrt(X);
    if (u < 3 * S) u = 3 * S;
    i64 cap = 1500000000LL;
    if (c.u_cap) {
        const i64 v = std::atoll(c.u_cap);
        if (v >= 100000000LL && v <= 4000000000LL) cap = v;
    }
    return std::min(u, cap);
}

i64 choose_compressed_limit(i64 X, const Controls& c) {   // mertens_dr.hpp:1401-1419
    if (X < 50000000LL) return X;
    const double ll = std::log(std::max(2.0, std::log((double)X)));
    double fx = 0.95;
    if (X >= 1000000000000000LL) fx = 1.25;
    if (X >= 10000000000000000LL) fx = 1.50;
    i64 u = (i64)(fx * std::pow((double)X / ll, 2.0 / 3.0));
    const i64 S = isqrt(X);
    if (u < 3 * S) u = 3 * S;
    u = ((u + 255) / 256) * 256;
    i64 cap = 8000000000LL;
    if (c.chmt_u_cap) {
        const i64 v = std::atoll(c.chmt_u_cap);
        if (v >= 100000000LL && v <= 10000000000LL) cap = v;
    }
    const i64 ends = (i64)(choose_cx(X, c) * std::sqrt((double)X)) + 1000;
    cap = std::max(cap, std::max(S + 1000, ends));
    return std::min(u, cap);
}

bool use_streaming(i64 X, const Controls& c) {
    bool streaming;
    if (c.use_chmt) {
        const char e = c.use_chmt[0];
        streaming = (e == '1' || e == 'y' || e == 'Y' || e == 't' || e == 'T');
    } else {
        streaming = X >= 100000000000LL && !c.fx && !c.u_cap;   // x86-64 default
    }
    return X >= 100000000000LL && streaming;
}

Plan plan_for(i64 X, const Controls& c) {
    Plan p;
    p.streaming = use_streaming(X, c);
    p.cx = choose_cx(X, c);
    const i64 A_max = (i64)(p.cx * std::sqrt((double)X)) + 1000;
    if (p.streaming) {   // mertens_dr.hpp:1451-1461
        p.u = choose_compressed_limit(X, c);
        i64 ends = A_max;
        ends = (ends + 255) & ~255LL;
        const i64 target = std::min<i64>(128000000LL, (p.u / 4 + 255) & ~255LL);
        p.u_dense = std::min(p.u, std::max(target, ends));
    } else {             // mertens_dr.hpp:1033-1041
        p.u = choose_sieve_limit(X, c);
        p.u_dense = p.u;
    }
    p.N = X / p.u;
    p.mu_limit = std::max(A_max, p.N + 1000);
    return p;
}

// ---- term evaluation ----------------------------------------------------------
namespace {
i64 split_A(i64 y, double cx) {
    i64 A = (i64)(cx * std::sqrt((double)y));
    if (A >= y) A = y - 1;
    if (A < 1) A = 1;
    return A;
}
i128 s1_sum(i64 y, i64 first, i64 last, i64 stride, const MLookup& M) {
    i128 s = 0;
    for (i64 n = first; n <= last; n += stride) s += M(qdiv(y, n));
    return s;
}
}  // namespace

i128 g_single(i64 y, double cx, i64 u, const MLookup& M, const int8_t* mu) {
    const i64 A = split_A(y, cx);
    const i64 kappa = y / (A + 1);
    const i128 S1 = s1_sum(y, y / u + 1, kappa, 1, M);
    const i128 S2 = s2_single(y, A, mu);
    return 1 - S1 + (i128)kappa * M(A) - S2;
}

i128 g_comb2(i64 y, double cx, i64 u, const MLookup& M, const int8_t* mu) {
    const i64 y2 = y / 2;
    const i64 A = split_A(y, cx), B = split_A(y2, cx);
    const i64 ky = y / (A + 1), ky2 = y2 / (B + 1), two = 2 * ky2;
    i64 odd0 = y / u + 1;
    if (!(odd0 & 1)) ++odd0;
    i128 S1 = s1_sum(y, odd0, ky, 2, M);
    if (two > ky) {
        i64 e0 = ky + 1;
        if (e0 & 1) ++e0;
        S1 -= s1_sum(y, e0, two, 2, M);
    } else if (two < ky) {
        i64 e0 = two + 1;
        if (e0 & 1) ++e0;
        S1 += s1_sum(y, e0, ky, 2, M);
    }
    const i128 S2 = s2_combined(y, A, B, mu);
    return -S1 + (i128)ky * M(A) - (i128)ky2 * M(B) - S2;
}

i128 g_brute(i64 y, i64 u, const std::vector<int32_t>& prefix) {
    i128 s = 0;
    for (i64 n = 1; n <= y / u; ++n) s += prefix[(size_t)(y / n)];
    return s;
}

std::vector<Term> outer_terms(i64 N, const int8_t* mu) {
    std::vector<Term> t;
    const i64 n6 = N / 6, n3 = N / 3, n2 = N / 2;
    for (i64 k = 1; k <= N; ++k) {
        if (k % 2 == 0 || k % 3 == 0 || mu[k] == 0) continue;
        const uint8_t r = k <= n6 ? 1 : k <= n3 ? 2 : k <= n2 ? 3 : 4;
        t.push_back({k, mu[k], r});
    }
    return t;
}

// mertens_dr.hpp:1372-1377.
i128 term_value(const Term& t, i64 X, double cx, i64 u, const MLookup& M, const int8_t* mu) {
    const i64 y = X / t.k;
    switch (t.range) {
        case 1: return g_comb2(y, cx, u, M, mu) - g_comb2(y / 3, cx, u, M, mu);
        case 2: return g_comb2(y, cx, u, M, mu) - g_single(y / 3, cx, u, M, mu);
        case 3: return g_comb2(y, cx, u, M, mu);
        default: return g_single(y, cx, u, M, mu);
    }
}

namespace {
i64 sum_terms(i64 X, double cx, i64 u, const MLookup& M, const int8_t* mu, i64 N, int threads) {
    const std::vector<Term> terms = outer_terms(N, mu);
    std::vector<i128> part((size_t)threads, 0);
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads)
    for (i64 i = 0; i < (i64)terms.size(); ++i)
        part[(size_t)thread_id()] += (i128)terms[(size_t)i].mu_k * term_value(terms[(size_t)i], X, cx, u, M, mu);
    i128 total = 0;
    for (i128 v : part) total += v;
    return (i64)total;
}
}  // namespace

// Linear (Euler) sieve over odd numbers, then the plain sum (mertens_dr.hpp:982-1010).
i64 mertens_small(i64 X) {
    if (X < 1) return 0;
    const i64 half = (X + 1) / 2;
    std::vector<int8_t> mo((size_t)half + 1, 0);
    std::vector<uint8_t> comp((size_t)half + 1, 0);
    std::vector<i64> primes;
    if (half >= 1) mo[1] = 1;
    for (i64 k = 2; k <= half; ++k) {
        const i64 v = 2 * k - 1;
        if (!comp[(size_t)k]) { primes.push_back(v); mo[(size_t)k] = -1; }
        for (i64 p : primes) {
            const i64 m = v * p;
            if (m > X) break;
            const i64 idx = (m + 1) / 2;
            comp[(size_t)idx] = 1;
            if (v % p == 0) { mo[(size_t)idx] = 0; break; }
            mo[(size_t)idx] = (int8_t)-mo[(size_t)k];
        }
    }
    i64 s = 0;
    for (i64 n = 1; n <= X; ++n) s += mu_from_odd(n, mo.data());
    return s;
}

i64 mertens_compressed(i64 X, int threads, const Controls& c) {   // :1442-1635
    if (X <= 100000LL) return mertens(X, threads, c);
    threads = resolve_threads(threads);
    const i64 u = choose_compressed_limit(X, c);
    i64 ends = (i64)(choose_cx(X, c) * std::sqrt((double)X)) + 1000;
    ends = (ends + 255) & ~255LL;
    const i64 target = std::min<i64>(128000000LL, (u / 4 + 255) & ~255LL);
    const i64 ud = std::min(u, std::max(target, ends));
    const double cx = choose_cx(X, c);
    const i64 N = X / u;
    const i64 mu_limit = std::max((i64)(cx * std::sqrt((double)X)) + 1000, N + 1000);
    CompressedTable table(u, ud, threads, mu_limit);
    if (X <= u) return table.get_M(X);
    MLookup M{table.dense_M.data(), table.u_dense, &table};
    return sum_terms(X, cx, u, M, table.mu.data(), N, threads);
}

i64 mertens(i64 X, int threads, const Controls& c) {   // mertens_dr.hpp:968-1383
    if (X < 1) return 0;
    if (X == 1) return 1;
    if (X == 2) return 0;
    if (X == 3) return -1;
    threads = resolve_threads(threads);
    if (X <= 100000LL) return mertens_small(X);
    if (X <= 50000000LL) {
        const i64 half = (X + 1) / 2;
        std::vector<int8_t> mo((size_t)half + 1);
        mobius_bytelog_odd(mo.data(), half, X, threads);
        i64 s = 0;
        for (i64 i = (X / 2 + 1) / 2 + 1; i <= half; ++i) s += mo[(size_t)i];
        return s;
    }
    if (use_streaming(X, c)) return mertens_compressed(X, threads, c);
    const i64 u = choose_sieve_limit(X, c);
    const double cx = choose_cx(X, c);
    const i64 N = X / u;
    const i64 mu_limit = std::max((i64)(cx * std::sqrt((double)X)) + 1000, N + 1000);
    DenseTable table(u, threads, X <= u ? 0 : mu_limit);
    if (X <= u) return table.get_M(X);
    MLookup M{table.M.data(), table.m_limit, nullptr};
    return sum_terms(X, cx, u, M, table.mu.data(), N, threads);
}

// test_sieve.cpp:30-41: memoised quotient-grouped recurrence
// M(x) = 1 - sum_{l >= 2} M(floor(x/l)) over an Eratosthenes prefix.
i64 mertens_recurrence(i64 X, i64 base_limit) {
    base_limit = std::max<i64>(base_limit, 16);
    const std::vector<int8_t> mu = mobius_eratosthenes(base_limit);
    std::vector<i64> prefix((size_t)base_limit + 1, 0);
    for (i64 n = 1; n <= base_limit; ++n) prefix[(size_t)n] = prefix[(size_t)n - 1] + mu[(size_t)n];
    std::unordered_map<i64, i64> memo;
    struct Rec {
        const std::vector<i64>& pre;
        std::unordered_map<i64, i64>& memo;
        i64 operator()(i64 x) {
            if (x < (i64)pre.size()) return pre[(size_t)x];
            auto it = memo.find(x);
            if (it != memo.end()) return it->second;
            i64 s = 1;
            for (i64 l = 2, r; l <= x; l = r + 1) {
                const i64 q = x / l;
                r = x / q;
                s -= (r - l + 1) * (*this)(q);
            }
            memo[x] = s;
            return s;
        }
    } rec{prefix, memo};
    if (X < 1) return 0;
    return rec(X);
}

}  // namespace mdr_oracle
