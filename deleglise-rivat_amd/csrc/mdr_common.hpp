// Shared host/device definitions for the MI355X Deleglise-Rivat engine.
//
// HBM layout of one table (all arrays device-resident, built by the sieve):
//   Md    int16 [ud + 1]         M(n) for n <= ud (dense tier)
//   cl    CLine [L]              compressed tier above ud: one 64-byte line per
//                                192 values (three 64-value words); word w of a
//                                line holds the nonzero and negative bit planes
//                                of its values and M just before its first value
//   mu6   u32   [ceil(T6/16)]    2-bit codes mu(j)+1 for j coprime to 6, indexed
//                                by t = floor(j/3), so j(t) = 3t + 1 + (t & 1)
// The reference keeps the same two tiers (mertens_dr.hpp:662-752) but stores
// the planes in 256-value chunks and the checkpoints in a separate array; here
// a scattered lookup touches exactly one 64-byte line.
#ifndef MDR_COMMON_HPP
#define MDR_COMMON_HPP

#include <stdint.h>

namespace mdr {

typedef long long i64;
typedef unsigned long long u64;

// Sieve segment lengths are fixed per kernel instantiation (SIEVE_SHAPES in
// mdr_kernels.hpp): multiples of 192 (compressed lines), 48 (packed mu words)
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
// Values per compressed line; u_dense must be a multiple of it below U, and
// every segment length is a multiple of it, so lines never straddle segments.
constexpr int LINE_VALUES = 192;
// Largest compressed tier (U - u_dense): line indices stay below 2^32.
constexpr long long MAX_COMPRESSED_VALUES = (long long)LINE_VALUES * 4294967295LL;

// Sieve workgroup shapes (segment length, threads): the table they build is
// the same; the autotuner picks one per decade of X.  Segment lengths are
// 3 * 2^shift, multiples of 192 (lines), 48 (mu words) and 16.
struct SieveShape {
    int seg, threads, shift;
};
constexpr SieveShape SIEVE_SHAPES[] = {{49152, 1024, 14}, {24576, 256, 13}, {98304, 1024, 15},
                                       {49152, 256, 14},  {49152, 512, 14},  {24576, 512, 13}};
constexpr int N_SIEVE_SHAPES = (int)(sizeof(SIEVE_SHAPES) / sizeof(SIEVE_SHAPES[0]));
constexpr int SEG_MAX = 98304;

struct alignas(64) CLine {
    u64 pl[6];        // word w: pl[2w] = nonzero plane, pl[2w+1] = negative plane
    int32_t ck[3];    // word w: M(value before its first value)
    int32_t pad;
};

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
