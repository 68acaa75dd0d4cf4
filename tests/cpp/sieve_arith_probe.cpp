// Checks the sieve's 32-bit arithmetic on the host, running the same constexpr
// code as the kernel: the records of large_prime_consts and the functions
// first_multiple, mod_below_square and square_divides against plain
// 64/128-bit remainders, for every prime from 256 up to the largest sieving
// prime a table below SIEVE_MAX_U needs, at segment bases sb = 0, 1, 2^32 - 1
// and seeded random ones; the four-byte finalisation mu_swar and mu_nibbles
// against the per-byte rule mu_byte, and mu_nibbles_direct against
// mu_nibbles of mu_swar; and the compressed lines' kept-value
// packing (kept_index, kept48, expand48) against per-bit rules.  Prints the
// number of checks and "ok", or the first failure.
#include "mdr_common.hpp"

#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using namespace mdr;
typedef unsigned __int128 u128;

static int fail(const char* what, u64 p, u64 a, u64 got, u64 want) {
    std::printf("FAIL %s p=%llu arg=%llu got=%llu want=%llu\n", what, p, a, got, want);
    return 1;
}

int main(int argc, char** argv) {
    const u64 pmax = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 10273905ull;   // isqrt(SIEVE_MAX_U) = 10273904
    const int per_prime = argc > 2 ? std::atoi(argv[2]) : 6;
    std::vector<uint8_t> comp(pmax + 1, 0);
    std::mt19937_64 rng(20261204);
    u64 checks = 0, primes = 0;
    for (u64 p = 2; p <= pmax; ++p) {
        if (comp[p]) continue;
        for (u64 j = p * p; j <= pmax; j += p) comp[j] = 1;
        if (p < 256) continue;
        ++primes;
        const PrimeConsts pc = large_prime_consts((uint32_t)p);
        if ((pc.pw & 0xffffffu) != p) return fail("pw", p, 0, pc.pw & 0xffffffu, p);
        const unsigned w = (64u - (unsigned)__builtin_clzll(p - 1)) | 1u;
        if ((pc.pw >> 24) != w) return fail("weight", p, 0, pc.pw >> 24, w);
        const bool below = p < SQ_SPLIT;
        const u64 m = below ? p * p : p;
        if (m >= (1ull << 31)) return fail("modulus", p, 0, m, 1ull << 31);
        u64 sbs[3 + 16] = {0, 1, 0xffffffffull};
        for (int i = 0; i < per_prime; ++i) sbs[3 + i] = rng() >> 32;
        for (int i = 0; i < 3 + per_prime; ++i) {
            const u64 sb = sbs[i];
            const u128 first = (u128)sb * (u64)SIEVE_UNIT + 1;
            const u64 want = (u64)((m - (u64)(first % m)) % m);
            const uint32_t got = first_multiple((uint32_t)sb, pc.c, pc.s, (uint32_t)m);
            ++checks;
            if (got != want) return fail("first_multiple", p, sb, got, want);
            if (below) {
                ++checks;
                if (mod_below_square(got, (uint32_t)p, pc.aux) != want % p)
                    return fail("mod_below_square", p, got, mod_below_square(got, (uint32_t)p, pc.aux), want % p);
            } else {
                // the hit at first + got and the next ones of the segment
                for (u64 k = 0; k < 3; ++k) {
                    const u128 n = first + got + k * p;
                    if (n / p >= (1ull << 32)) break;
                    const bool sq = (u64)(n % ((u128)p * p)) == 0;
                    ++checks;
                    if (square_divides((uint32_t)(u64)n, (uint32_t)p, pc.aux) != sq)
                        return fail("square_divides", p, (u64)n, !sq, sq);
                }
            }
        }
        if (below) {
            // r = 0, p - 1, p, p^2 - 1 and random r < p^2
            const u64 rs[5] = {0, p - 1, p, m - 1, (rng() >> 1) % m};
            for (u64 r : rs) {
                ++checks;
                if (mod_below_square((uint32_t)r, (uint32_t)p, pc.aux) != r % p)
                    return fail("mod_below_square", p, r, mod_below_square((uint32_t)r, (uint32_t)p, pc.aux), r % p);
            }
        } else {
            // multiples n = t p with n / p = t < 2^32: t = j p (square), j p +- 1, the largest t
            const u64 jmax = 0xffffffffull / p;
            const u64 js[4] = {1, jmax, 1 + (rng() >> 1) % jmax, jmax > 2 ? jmax - 1 : 1};
            for (u64 j : js) {
                for (int d = -1; d <= 1; ++d) {
                    const u64 t = j * p + (u64)(long long)d;
                    if (t == 0 || t > 0xffffffffull) continue;
                    const uint32_t n_lo = (uint32_t)(t * p);
                    ++checks;
                    if (square_divides(n_lo, (uint32_t)p, pc.aux) != (d == 0))
                        return fail("square_divides", p, t, d != 0, d == 0);
                }
            }
        }
    }
    // finalisation: mu_swar against mu_byte for every byte value and every
    // threshold (bytes are independent: check each value in every position,
    // beside seeded other bytes), mu_nibbles against the per-bit gather
    for (unsigned thr = 0; thr < 63; ++thr) {
        const uint32_t add = (127u - thr) * 0x01010101u;
        const u64 n = 1ull << thr;   // floor(log2 n) = thr
        for (unsigned a = 0; a < 256; ++a) {
            for (int pos = 0; pos < 4; ++pos) {
                const uint32_t other = (uint32_t)rng();
                const uint32_t x = (other & ~(0xffu << (8 * pos))) | (a << (8 * pos));
                const uint32_t m = mu_swar(x, add);
                ++checks;
                if (mu_nibbles_direct(x, add) != mu_nibbles(m))
                    return fail("mu_nibbles_direct", thr, x, mu_nibbles_direct(x, add), mu_nibbles(m));
                for (int b = 0; b < 4; ++b) {
                    const uint8_t want = (uint8_t)mu_byte((uint8_t)(x >> (8 * b)), n);
                    ++checks;
                    if ((uint8_t)(m >> (8 * b)) != want) return fail("mu_swar", thr, x, (m >> (8 * b)) & 0xff, want);
                }
            }
        }
    }
    const uint8_t mus[3] = {0x00, 0x01, 0xff};
    for (int i = 0; i < 81; ++i) {
        uint32_t mu = 0, want = 0;
        for (int b = 0, k = i; b < 4; ++b, k /= 3) {
            const uint8_t v = mus[k % 3];
            mu |= (uint32_t)v << (8 * b);
            want |= (uint32_t)(v != 0) << b | (uint32_t)(v == 0xff) << (4 + b);
        }
        ++checks;
        if (mu_nibbles(mu) != want) return fail("mu_nibbles", 0, mu, mu_nibbles(mu), want);
    }
    // compressed lines: kept_index against a count of the kept offsets; kept48
    // of the nibble dwords of 64 seeded mu values (zero at the offsets 3 mod 4)
    // against the kept bits in order, expand48 back to the plane bits
    for (uint32_t r = 0, kept = 0; r < (uint32_t)LINE_VALUES; ++r) {
        if (r % 4 != 3) ++kept;
        ++checks;
        if (kept_index(r) != kept - 1) return fail("kept_index", 0, r, kept_index(r), kept - 1);
    }
    for (int trial = 0; trial < 20000; ++trial) {
        int8_t mu[64];
        for (int r = 0; r < 64; ++r) {
            const u64 x = rng() % 3;
            mu[r] = (int8_t)(r % 4 == 3 ? 0 : (trial < 3 ? (int)trial - 1 : (int)x - 1));
        }
        uint32_t nib[4];
        u64 nz = 0, ng = 0, knz = 0, kng = 0;
        for (int j = 0, kb = 0; j < 4; ++j) {
            nib[j] = 0;
            for (int d = 0; d < 4; ++d) {
                uint32_t bytes = 0;
                for (int bb = 0; bb < 4; ++bb) bytes |= (uint32_t)(uint8_t)mu[16 * j + 4 * d + bb] << (8 * bb);
                nib[j] |= mu_nibbles(bytes) << (8 * d);
            }
            for (int v = 16 * j; v < 16 * j + 16; ++v) {
                nz |= (u64)(mu[v] != 0) << v;
                ng |= (u64)(mu[v] < 0) << v;
                if (v % 4 == 3) continue;
                knz |= (u64)(mu[v] != 0) << kb;
                kng |= (u64)(mu[v] < 0) << kb;
                ++kb;
            }
        }
        const u64 gz = kept48(nib[0], nib[1], nib[2], nib[3]);
        const u64 gg = kept48(nib[0] >> 4, nib[1] >> 4, nib[2] >> 4, nib[3] >> 4);
        checks += 4;
        if (gz != knz) return fail("kept48 nonzero", 0, nz, gz, knz);
        if (gg != kng) return fail("kept48 negative", 0, ng, gg, kng);
        if (expand48(gz) != nz) return fail("expand48 nonzero", 0, gz, expand48(gz), nz);
        if (expand48(gg) != ng) return fail("expand48 negative", 0, gg, expand48(gg), ng);
    }
    std::printf("%llu primes %llu checks ok\n", primes, checks);
    return 0;
}
