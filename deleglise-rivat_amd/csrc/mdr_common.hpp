// Shared host/device definitions for the MI355X Deleglise-Rivat engine.
//
// HBM layout of one table (all arrays device-resident, built by the sieve):
//   Md    int16 [ud + 1]         M(n) for n <= ud (dense tier)
//   cl    CLine [L]              compressed tier above ud: one 64-byte line per
//                                256 values, of which the 192 not divisible by 4
//                                are stored (mu = 0 on multiples of 4) as three
//                                64-value words; word w of a line holds the
//                                nonzero and negative bit planes of its values
//                                and M just before its first value
//   mu6   u32   [ceil(T6/16)]    2-bit codes mu(j)+1 for j coprime to 6, indexed
//                                by t = floor(j/3), so j(t) = 3t + 1 + (t & 1)
// The reference keeps the same two tiers (mertens_dr.hpp:662-752) but stores
// the planes in 256-value chunks and the checkpoints in a separate array; here
// a scattered lookup touches exactly one 64-byte line, which covers 256 values
// in 0.25 B per value.
#ifndef MDR_COMMON_HPP
#define MDR_COMMON_HPP

#include <stdint.h>

namespace mdr {

typedef long long i64;
typedef unsigned long long u64;

// Sieve segment lengths are fixed per kernel instantiation (SIEVE_SHAPES in
// mdr_kernels.hpp): multiples of 256 (compressed lines), 48 (packed mu words)
// and 16 (aligned 16-byte stencil loads).
// Stencil period lcm(16, 4, 9, 25, 7, 11): weights of 2,3,5,7,11 and the square
// flags of 4, 9 and 25 repeat with it; 16 | period keeps phases 16-byte aligned.
constexpr int STENCIL_PERIOD = 277200;
// Second periodic pattern: the weights of the sieving primes 13, 17, 19 and 23
// (the first PATTERN_PRIMES entries of the ascending sieving-prime list), with
// period 13*17*19*23.  It is stored as PATTERN_REPEAT periods (plus one
// segment) so that the read offset, base mod (PATTERN_REPEAT * period), is a
// multiple of 4 for every segment base (a multiple of 16): dword reads.
constexpr int PATTERN_PERIOD = 13 * 17 * 19 * 23;   // 96577
constexpr int PATTERN_REPEAT = 4;
constexpr int PATTERN_PRIMES = 4;
constexpr unsigned pattern_byte(unsigned v) {   // v = n mod period (constexpr: host and device)
    return (v % 13 == 0 ? 5u : 0u) + (v % 17 == 0 ? 5u : 0u) + (v % 19 == 0 ? 5u : 0u) + (v % 23 == 0 ? 5u : 0u);
}
// Third periodic pattern: the weights of the next sieving primes 29, 31 and
// 37 (bit_width(p - 1) | 1 = 5, 5, 7), period 29*31*37, stored the same way
// (PATTERN_REPEAT periods plus one segment: 4 | PATTERN3_PERIOD * 4 keeps the
// reads dword-aligned).
constexpr int PATTERN3_PERIOD = 29 * 31 * 37;   // 33263
constexpr int PATTERN3_PRIMES = 3;
constexpr unsigned pattern3_byte(unsigned v) {   // v = n mod PATTERN3_PERIOD
    return (v % 29 == 0 ? 5u : 0u) + (v % 31 == 0 ? 5u : 0u) + (v % 37 == 0 ? 7u : 0u);
}
// Values per compressed line; u_dense must be a multiple of it below U, and
// every segment length is a multiple of it, so lines never straddle segments.
// A line therefore starts at n = 1 (mod 4): its multiples of 4 (mu = 0) sit
// at the offsets r = 3 (mod 4) and are not stored, so its other LINE_KEPT
// values, the kept values, fill three 64-bit words per plane.
constexpr int LINE_VALUES = 256;
constexpr int LINE_KEPT = 192;
// Largest compressed tier (U - u_dense): line indices stay below 2^32.
constexpr long long MAX_COMPRESSED_VALUES = (long long)LINE_VALUES * 4294967295LL;
// Kept index of the last kept value at or below line offset r (0..255): of the
// r + 1 values up to r, floor((r + 1) / 4) are multiples of 4.  A multiple of
// 4 maps to the kept value before it, whose prefix sum it shares (mu = 0).
constexpr uint32_t kept_index(uint32_t r) { return r - ((r + 1u) >> 2); }
// The kept bits of 16 values from four nibble bytes (bits 0..2 of each byte;
// bit 3, a multiple of 4, is zero and dropped): 12 bits, byte i at 3 i.
constexpr uint32_t pack12(uint32_t v) {
    uint32_t t = v & 0x07070707u;
    t = (t | (t >> 5)) & 0x003f003fu;   // bytes 1 and 3 next to bytes 0 and 2
    return (t | (t >> 10)) & 0xfffu;    // bits 16..21 next to bits 0..5
}
// The 48 kept bits of one 64-value word from its four nibble dwords (16
// values each, byte i of dword j holding the values 16 j + 4 i .. + 3).
constexpr u64 kept48(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3) {
    return (u64)(pack12(d0) | pack12(d1) << 12) | (u64)(pack12(d2) | pack12(d3) << 12) << 24;
}
// Inverse for one 64-value word: 48 kept bits to the 64 plane bits, zero at
// the offsets 3 (mod 4).
constexpr u64 expand48(u64 k) {
    u64 x = (k & 0xfffull) | ((k & 0xfff000ull) << 4) | ((k & 0xfff000000ull) << 8) | ((k & 0xfff000000000ull) << 12);
    x = (x & 0x003f003f003f003full) | ((x & 0x0fc00fc00fc00fc0ull) << 2);
    return (x & 0x0707070707070707ull) | ((x & 0x3838383838383838ull) << 1);
}

// Sieve workgroup shapes (segment length, threads): the table they build is
// the same; the autotuner picks one per decade of X.  Segment lengths are
// 3 * 2^shift, multiples of 256 (lines), 48 (mu words) and 16.
struct SieveShape {
    int seg, threads, shift;
};
constexpr SieveShape SIEVE_SHAPES[] = {{49152, 1024, 14}, {24576, 256, 13}, {98304, 1024, 15},
                                       {49152, 256, 14},  {49152, 512, 14},  {24576, 512, 13}};
constexpr int N_SIEVE_SHAPES = (int)(sizeof(SIEVE_SHAPES) / sizeof(SIEVE_SHAPES[0]));
constexpr bool sieve_shapes_fit() {
    for (const SieveShape& sh : SIEVE_SHAPES)
        if (sh.seg % LINE_VALUES || sh.seg % 48 || sh.seg != 3 << sh.shift) return false;
    return true;
}
static_assert(sieve_shapes_fit(), "segments hold whole lines and mu words");
constexpr int SEG_MAX = 98304;
// Every segment base is a multiple of SIEVE_UNIT (the shortest segment); the
// large-prime stage of the sieve writes a base as sb * SIEVE_UNIT with sb a
// 32-bit integer, so a table ends below SIEVE_UNIT * 2^32 (about 1.05e14).
constexpr long long SIEVE_UNIT = 24576;
constexpr long long SIEVE_MAX_U = SIEVE_UNIT * 4294967296LL - 1;
// Large sieving primes below SQ_SPLIT have p^2 < 2^31: their segment offsets
// come from a remainder mod p^2 (square hit) reduced mod p (first weight
// hit); primes from SQ_SPLIT on take the remainder mod p and test each hit n
// for p^2 | n.
constexpr unsigned SQ_SPLIT = 46341;
// Large primes below COOP_PRIME_LIMIT (about SEG/24 hits or more per segment)
// are marked by a whole wavefront each, lanes on consecutive multiples; the
// larger ones by one thread each.
constexpr unsigned COOP_PRIME_LIMIT = 1024;
static_assert((unsigned long long)COOP_PRIME_LIMIT * COOP_PRIME_LIMIT > SEG_MAX, "one square hit per thread-marked prime");
static_assert(64ull * 256 * 256 > SEG_MAX, "one pass of square hits per wave-marked prime");

// Per-prime constants of the large-prime sieve stage (p >= 256), one 16-byte
// record, so that a segment's offsets take 32-bit multiplies instead of a
// 64-bit division.  With m = p^2 below SQ_SPLIT and m = p from it on:
//   pw = p | w << 24, w = bit_width(p - 1) | 1 (the byte weight);
//   c  = (-SIEVE_UNIT) mod m, so (-b) mod m = sb * c mod m for b = sb * SIEVE_UNIT;
//   s  = floor(c * 2^32 / m): q = mulhi(sb, s) leaves sb * c - q * m in [0, 2m)
//        (Shoup's precomputed-quotient multiplication; m < 2^31);
//   aux = floor(2^32 / p) below SQ_SPLIT (quotients of values below p^2 by p,
//        low by at most 1), p^-1 mod 2^32 from SQ_SPLIT on (exact quotients n / p
//        of multiples of p, and the divisibility test of n / p by p).
struct alignas(16) PrimeConsts {
    uint32_t pw, c, s, aux;
};
// The record array ends with PRIME_RECORD_PAD records of p = 0xffffff (above
// every sieving root), so that a workgroup of up to 1,024 threads reading up
// to 8 * 1,024 records past its last live record stays inside the array
// without clamping its index.
constexpr int PRIME_RECORD_PAD = 8192;
constexpr PrimeConsts PRIME_RECORD_END{0xffffffu, 0u, 0u, 0u};
constexpr PrimeConsts large_prime_consts(uint32_t p) {
    const unsigned bw = 32u - (unsigned)__builtin_clz(p - 1u);
    PrimeConsts r{p | ((bw | 1u) << 24), 0, 0, 0};
    const u64 m = p < SQ_SPLIT ? (u64)p * p : (u64)p;
    const u64 c = (m - (u64)SIEVE_UNIT % m) % m;
    r.c = (uint32_t)c;
    r.s = (uint32_t)((c << 32) / m);
    if (p < SQ_SPLIT) {
        r.aux = (uint32_t)((1ull << 32) / p);
    } else {
        uint32_t x = p;                                 // p * p = 1 mod 8: 3 bits
        for (int i = 0; i < 4; ++i) x *= 2u - p * x;    // 6, 12, 24, 48 bits
        r.aux = x;
    }
    return r;
}

// The arithmetic of the large-prime stage on those records (constexpr: the
// sieve kernel and the host checks run this same code).
constexpr uint32_t umin32(uint32_t x, uint32_t y) { return x < y ? x : y; }
constexpr uint32_t mulhi32(uint32_t x, uint32_t y) { return (uint32_t)(((u64)x * y) >> 32); }
// (-(b + 1)) mod m for the segment base b = sb * SIEVE_UNIT, i.e. the offset
// of the first multiple of m at or after first = b + 1, from c = (-SIEVE_UNIT)
// mod m and its Shoup quotient s (m < 2^31).
constexpr uint32_t first_multiple(uint32_t sb, uint32_t c, uint32_t s, uint32_t m) {
    uint32_t r = sb * c - mulhi32(sb, s) * m;   // (-b) mod m, or that plus m
    r = umin32(r, r - m);                       // x - m wraps above x when x < m
    r = r + m - 1u;
    return umin32(r, r - m);
}
// r mod p for r < p^2 < 2^31, with aux = floor(2^32 / p) (the quotient
// estimate is low by at most 1).
constexpr uint32_t mod_below_square(uint32_t r, uint32_t p, uint32_t aux) {
    const uint32_t x = r - mulhi32(r, aux) * p;
    return umin32(x, x - p);
}
// For a multiple n of an odd p with n / p < 2^32 (n_lo = n mod 2^32,
// pinv = p^-1 mod 2^32): p^2 divides n.  n / p = n_lo * pinv exactly, and a
// 32-bit t is a multiple of p exactly when t * pinv (mod 2^32) times p stays
// below 2^32.
constexpr bool square_divides(uint32_t n_lo, uint32_t p, uint32_t pinv) {
    const uint32_t q = n_lo * pinv;
    return mulhi32(q * pinv, p) == 0u;
}

// Finalisation of the byte-log accumulator (constexpr: the sieve kernel and
// the host checks run this same code).  A byte a holds the odd weights of the
// sieving primes dividing n and 0x80 for a square factor: mu = 0 for a flag,
// otherwise mu = +1 exactly when the parity bit differs from [a > floor(log2 n)]
// (a remaining cofactor above the sieving bound is one more prime).
constexpr int8_t mu_byte(uint8_t a, u64 n) {
    if (a & 0x80) return 0;
    const unsigned thr = 63u - (unsigned)__builtin_clzll(n);
    const int s = (a & 1) ? -1 : 1;
    return (int8_t)(a > thr ? s : -s);
}
// Four bytes at once for values sharing floor(log2 n) = thr, with
// add = (127 - thr) * 0x01010101: (a & 0x7f) + 127 - thr sets bit 7 exactly
// when a & 0x7f > thr and never carries into the next byte.  The byte masks
// come from bits 0, 8, 16, 24 by shifts and a subtraction (b * 0xfe =
// (b << 8) - (b << 1), b * 0xff = (b << 8) - b: nothing carries between
// bytes) rather than by 32-bit multiplies.
constexpr uint32_t mu_swar(uint32_t x, uint32_t add) {
    const uint32_t gt = ((x & 0x7f7f7f7fu) + add) & 0x80808080u;
    const uint32_t t = (x & 0x01010101u) ^ (gt >> 7);
    const uint32_t f = (x >> 7) & 0x01010101u;   // square flags
    return ~((t << 8) - (t << 1)) & ~((f << 8) - f);
}
// The nonzero bits (0, 8, 16, 24) and the negative bits (7, 15, 23, 31, moved
// to 4, 12, 20, 28) of four mu bytes gathered to bits 0..3 and 4..7: a bit at
// 8i + j (j = 0 or 4) reaches i + j through the shift by 7i, and no other
// shifted bit lands in 0..7.
constexpr uint32_t mu_nibbles(uint32_t mu) {
    const uint32_t c = (mu & 0x01010101u) | ((mu >> 3) & 0x10101010u);
    return (c | (c >> 7) | (c >> 14) | (c >> 21)) & 0xffu;
}

// mu_nibbles(mu_swar(x, add)) without forming the mu bytes: a byte is
// nonzero when it has no square flag, and negative (mu = -1) when its parity
// bit equals [a & 0x7f > floor(log2 n)] (mu_byte: mu = s for a > thr, -s
// otherwise, s = -1 for an odd parity).
constexpr uint32_t mu_nibbles_direct(uint32_t x, uint32_t add) {
    const uint32_t gt = ((x & 0x7f7f7f7fu) + add) & 0x80808080u;
    const uint32_t nz = ~(x >> 7) & 0x01010101u;
    const uint32_t ng = ~((x & 0x01010101u) ^ (gt >> 7)) & nz;
    const uint32_t c = nz | (ng << 4);
    return (c | (c >> 7) | (c >> 14) | (c >> 21)) & 0xffu;
}

struct alignas(64) CLine {
    u64 pl[6];        // word w (kept values 64 w .. 64 w + 63): pl[2w] = nonzero
                      // plane, pl[2w+1] = negative plane
    int32_t ck[3];    // word w: M(value before its first kept value)
    int32_t pad;
};
static_assert(sizeof(CLine) == 64 && 3 * 64 == LINE_KEPT && LINE_VALUES - LINE_VALUES / 4 == LINE_KEPT,
              "a line: three kept words and their checkpoints in 64 bytes");

// Segment kinds of one G evaluation.
enum SegType : uint8_t {
    SEG_NONE = 0,
    SEG_S1_DENSE = 1,    // sum M(y/n), q <= ud
    SEG_S1_COMP = 2,     // sum M(y/n), q > ud
    SEG_S2 = 3,          // sum mu(j) piece(y/j); piece id stored separately
};

constexpr int MAX_SEGS = 12;

// One record of the cycle log (training data for the cost model): the
// positions of one cost piece processed by one wave and the shader-clock
// cycles they took.
struct StampRec {
    i64 y;          // argument of the G evaluation
    i64 start;      // first n (S1) or j (S2)
    i64 end;        // last n or j
    uint32_t kind;  // cost-model kind (0..15)
    uint32_t cycles;
    u64 count;      // positions
};

// Per-X parameters (device-resident array).
struct XParam {
    i64 X;
    i64 u;            // sieve cutoff used by this X's outer sum
    i64 N;            // X / u
    i64 T;            // number of k <= N coprime to 6 (term slots)
    i64 slot_base;    // first global G-slot of this X (2 slots per term)
    double cx;
    i64 n6, n3, n2;
};

// Table parameters (kernel argument).
struct TableParam {
    i64 U;            // table covers n = 1..U
    i64 ud;           // dense tier limit (ud % LINE_VALUES == 0 when L > 0)
    i64 L;            // compressed lines
    i64 mu_lim;       // mu6 covers j <= mu_lim
    const int16_t* Md;
    const CLine* cl;
    const uint32_t* mu6;
    // Quotient pass over a windowed table: 0 = the whole table is resident;
    // 1 = the resident part (every position, compressed lookups only for
    // q <= qhi); 2 = one window of the compressed tier (only the compressed
    // lookups with qlo <= q <= qhi, from lines that start at qlo = ud + 1).
    // Passes 0 and 1 have qlo = ud + 1.
    int pass;
    i64 qlo, qhi;
};

}  // namespace mdr

#endif
