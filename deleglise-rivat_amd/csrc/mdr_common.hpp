// Shared host/device definitions for the MI355X Deleglise-Rivat engine.
//
// HBM layout of one table (all arrays device-resident, built by the sieve):
//   Md    int16 [ud + 1]         M(n) for n <= ud (dense tier)
//   cpl   u64x2 [W]              per 64-value word above ud: {nonzero, negative}
//                                bit planes (word w covers n = ud+1+64w .. ud+64+64w)
//   cck   int32 [W]              M(ud + 64w), the value just before word w
//   mu6   u32   [ceil(T6/16)]    2-bit codes mu(j)+1 for j coprime to 6, indexed
//                                by t = floor(j/3), so j(t) = 3t + 1 + (t & 1)
// The reference keeps the same two tiers (mertens_dr.hpp:662-752) with the
// planes in 256-value chunks; here a word's planes are one 16-byte load.
#ifndef MDR_COMMON_HPP
#define MDR_COMMON_HPP

#include <stdint.h>

namespace mdr {

typedef long long i64;
typedef unsigned long long u64;

// Sieve segment: a multiple of 96 (so mu6 words align), of 64 (plane words)
// and of 16 (aligned 16-byte stencil loads).
constexpr int SEG = 49152;
constexpr int SIEVE_THREADS = 512;
// Stencil period lcm(16, 4, 9, 25, 7, 11): weights of 2,3,5,7,11 and the square
// flags of 4, 9 and 25 repeat with it; 16 | period keeps phases 16-byte aligned.
constexpr int STENCIL_PERIOD = 277200;

// Segment kinds of one G evaluation.
enum SegType : uint8_t {
    SEG_NONE = 0,
    SEG_S1_DENSE = 1,    // sum M(y/n), q <= ud
    SEG_S1_COMP = 2,     // sum M(y/n), q > ud
    SEG_S2 = 3,          // sum mu(j) piece(y/j); piece id stored separately
};

constexpr int MAX_SEGS = 12;

// One quotient-phase work chunk: positions [p_lo, p_hi) of one segment.
struct ChunkRec {
    uint32_t x;        // index of X in the batch
    uint32_t g;        // G-evaluation slot within that X
    uint32_t seg;      // segment index (0..MAX_SEGS-1); bit 8 = add the constant
    uint32_t pad;
    u64 p_lo;
    u64 p_hi;
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
    i64 ud;           // dense tier limit (ud % 64 == 0 when W > 0)
    i64 W;            // compressed words
    i64 mu_lim;       // mu6 covers j <= mu_lim
    const int16_t* Md;
    const u64* cpl;   // 2 u64 per word
    const int32_t* cck;
    const uint32_t* mu6;
};

}  // namespace mdr

#endif
