// Device primitives: exact floor division, S2 summands, table lookups,
// 128-bit accumulation and the geometry of one G evaluation.
#ifndef MDR_DEVICE_HPP
#define MDR_DEVICE_HPP

#include <hip/hip_runtime.h>
#include "mdr_common.hpp"

namespace mdr {

typedef __int128 i128;
typedef unsigned __int128 u128;

// ---------------------------------------------------------------------------
// Exact floor division.  The reference truncates an IEEE double quotient for
// y <= 2^53 and falls back to 64-bit integer division above (mertens_dr.hpp:
// 80-86).  Here the quotient estimate is y * (1/n) with the hardware
// reciprocal refined by two Newton steps (relative error < 2^-50 even from a
// 2^-14 seed), and the integer remainder then fixes the estimate exactly.
// The estimate is never trusted on its own, for any y.
// ---------------------------------------------------------------------------
// One Newton step on v_rcp_f64: the hardware estimate is within 2^-24.3
// relative (measured over every n < 2^31 and samples to 2^52), one step
// squares that to <= 2^-48.6 (plus rounding).  Enough for every quotient
// width below: the estimate y*r is within 2^-4 of y/n for q < 2^44 and within
// 2^15 for q < 2^63, and each path ends with an exact integer correction that
// tolerates an estimate off by one (qdiv_big: a second, exact-remainder step).
__device__ __forceinline__ double rcp_fast(double d) {
    const double r = __builtin_amdgcn_rcp(d);
    const double e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Two reciprocals from one v_rcp_f64 (the transcendental costs about four
// plain FP64 operations): 1/(a b) refined once, times b and times a.  Each
// result is within 2^-48.3 relative (a b rounds by 2^-53, the refined
// reciprocal is within 2^-48.6, the products round by 2^-53); a, b < 2^32.
__device__ __forceinline__ void rcp_pair(double a, double b, double& ra, double& rb) {
    const double r = rcp_fast(a * b);
    ra = r * b;
    rb = r * a;
}

// Four reciprocals from one v_rcp_f64: 1/(a b c d) refined once, times c d
// and a b, then times the other factor of each pair.  Each result is within
// 2^-48 relative (the three products round by 2^-53 each, the refined
// reciprocal is within 2^-48.6, the two further products round by 2^-53
// each); a, b, c, d < 2^32, so a b c d < 2^128.
__device__ __forceinline__ void rcp_quad(double a, double b, double c, double d, double& ra, double& rb,
                                         double& rc, double& rd) {
    const double ab = a * b, cd = c * d;
    const double r = rcp_fast(ab * cd);
    const double rab = r * cd, rcd = r * ab;
    ra = rab * b;
    rb = rab * a;
    rc = rcd * d;
    rd = rcd * c;
}

// q = floor(y/n) for q <= 2^32 - 2 and 2n <= 2^32, from an estimate rn of 1/n
// within 2^-47 relative.  2^52 + y*rn (one FMA) rounds the estimate, which is
// within 2^-15 of y/n, to the nearest integer q0 in {q, q + 1} and leaves q0 in
// the low mantissa word.  The true remainder r = y - q0*n lies in [0, n) for
// q0 = q and in [-n, 0) for q0 = q + 1, so t = (q0*n - y - 1) mod 2^32 =
// (-r - 1) mod 2^32 lies in [2^32 - n, 2^32) (at least n, as n <= 2^31) or in
// [0, n): one multiply-add (q0*n + ~y, a single v_mad_u64_u32) and one
// comparison correct the estimate.
__device__ __forceinline__ uint32_t qdiv32n(uint32_t y_lo, double dy, uint32_t n, double rn) {
    const double qm = __builtin_fma(dy, rn, 4503599627370496.0);
    const uint32_t q0 = (uint32_t)__double_as_longlong(qm);
    const uint32_t t = q0 * n + ~y_lo;
    return q0 - (t < n ? 1u : 0u);
}

// Two Newton steps (set-up code outside the hot loops).
__device__ __forceinline__ double rcp_refined(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

// q = floor(y/n) when q < 2^32 and 3n <= 2^32.  The estimate is within 1 of q,
// so the true remainder y - q0*n lies in [-n, 2n), which the 32-bit modular
// remainder identifies unambiguously under 3n <= 2^32.
__device__ __forceinline__ uint32_t qdiv32(uint32_t y_lo, double dy, uint32_t n, double rn) {
    uint32_t q = (uint32_t)(dy * rn);
    const uint32_t r = y_lo - q * n;
    q += (r >= 2u * n) ? 0xffffffffu : (r >= n ? 1u : 0u);
    return q;
}

// q = floor(y/n) when q < 2^44 and n < 2^32 (estimate within 1 of q).
__device__ __forceinline__ u64 qdiv_mid(u64 y, double dy, u64 n, double rn) {
    u64 q = (u64)(dy * rn);
    const i64 r = (i64)(y - q * n);
    q += (r < 0) ? (u64)-1 : (r >= (i64)n ? 1ull : 0ull);
    return q;
}

// q = floor(y/n) for any y < 2^63, 1 <= n < 2^32: the first estimate is
// within 2^15 of q (2^13 with two Newton steps), the remainder is then exact
// in double, a second estimate fixes all but a final +-1.
__device__ __forceinline__ u64 qdiv_big(u64 y, double dy, u64 n, double rn) {
    u64 q = (u64)(dy * rn);
    i64 r = (i64)(y - q * n);
    const i64 e = (i64)__builtin_floor((double)r * rn);
    q += (u64)e;
    r -= e * (i64)n;
    q += (r < 0) ? (u64)-1 : (r >= (i64)n ? 1ull : 0ull);
    return q;
}

// Exact 64-bit division for set-up code (one per chunk, not in hot loops).
__device__ __forceinline__ i64 div_exact(i64 y, i64 n) {
    const double rn = rcp_refined((double)n);
    if (n < (1ll << 32)) return (i64)qdiv_big((u64)y, (double)y, (u64)n, rn);
    return y / n;
}

// ---------------------------------------------------------------------------
// S2 summands.  Closed floor forms of the reference's Piece1..8 and
// SinglePiece1..4 (mertens_dr.hpp:470-482): inclusion-exclusion over the
// multiples j, 2j, 3j, 6j kept by the A (and B) breakpoints, written with
// floor(q/m) = floor(y/(m j)).  Ids 1..8 comb2 pieces, 9..12 single pieces.
// ---------------------------------------------------------------------------
template <int P, typename U>
__device__ __forceinline__ i64 piece_eval(U q) {
    const i64 Q = (i64)q;
    const i64 h = (i64)(q >> 1), f = (i64)(q >> 2);
    if constexpr (P == 6) return Q - 2 * h;
    else if constexpr (P == 7 || P == 11) return Q - h;
    else if constexpr (P == 8 || P == 12) return Q;
    else if constexpr (P == 5) return (Q - h) - (h - f);
    else {
        const U tq = q / 3;
        const i64 t = (i64)tq, s = (i64)(tq >> 1), w = (i64)(tq >> 2);
        if constexpr (P == 1) return (Q - h - t + s) - (h - f - s + w);
        else if constexpr (P == 2) return (Q - h - t + s) - (h - f - s);
        else if constexpr (P == 3) return (Q - h - t) - (h - f - s);
        else if constexpr (P == 4) return (Q - h - t) - (h - f);
        else if constexpr (P == 9) return Q - h - t + s;
        else return Q - h - t;   // P == 10
    }
}

// Number of j in [1, x] with gcd(j, 6) = 1 (x >= 0).
__host__ __device__ __forceinline__ i64 count_coprime6(i64 x) {
    if (x <= 0) return 0;
    const i64 r = x % 6;
    return 2 * (x / 6) + (r >= 1) + (r >= 5);
}
__host__ __device__ __forceinline__ i64 j_of_t(i64 t) { return 3 * t + 1 + (t & 1); }

// mu(j(t)) from its 2-bit two's-complement code (1 = 01, 0 = 00, -1 = 11):
// a signed bit-field extract (one v_bfe_i32; shift is even, at most 30).
__device__ __forceinline__ int mu6_field(uint32_t word, uint32_t shift) {
    return __builtin_amdgcn_sbfe((int)word, shift, 2u);
}
__device__ __forceinline__ int mu6_mu(const uint32_t* mu6, i64 t) {
    return mu6_field(mu6[t >> 4], (uint32_t)(2 * (t & 15)));
}

// M from one compressed line: word sub (0..2), kept bit b (0..63) of the word.
// The prefix popcount over bits 0..b is the popcount of the word shifted left
// by 63 - b (the bits above b leave the word), which needs no mask.
__device__ __forceinline__ int32_t line_M(const CLine* L, uint32_t sub, uint32_t b) {
    const uint4 pw = reinterpret_cast<const uint4*>(L)[sub];          // planes at 16 sub
    const int32_t ck = reinterpret_cast<const int32_t*>(L)[12u + sub];  // checkpoint at 48 + 4 sub
    const uint32_t sh = 63u - (b & 63u);
    const u64 nz = ((u64)pw.y << 32 | pw.x) << sh;
    const u64 ng = ((u64)pw.w << 32 | pw.z) << sh;
    return ck + (int)__popcll(nz) - 2 * (int)__popcll(ng);
}

// M(q) in the compressed tier (q > ud), off = q - ud - 1 < 256 * 2^32: the
// value's line off / 256 and, for the kept index k of its line offset
// (kept_index), word k / 64 and bit k mod 64.
static_assert(LINE_VALUES == 256, "a compressed value's line is off >> 8");
__device__ __forceinline__ int32_t comp_M(const CLine* cl, u64 off) {
    const uint32_t k = kept_index((uint32_t)off & (LINE_VALUES - 1u));
    return line_M(cl + (off >> 8), k >> 6, k & 63u);
}

// Plane bits of the 64 values at tier offsets 64 g .. 64 g + 63 (zero at the
// multiples of 4): the group's 48 kept bits at kept offset 48 (g mod 4) of
// line g / 4, expanded.
__device__ __forceinline__ void raw_planes(const CLine* cl, u64 g, u64& nz, u64& ng) {
    const CLine& L = cl[g >> 2];
    const uint32_t o = 48u * (uint32_t)(g & 3);   // 0, 48, 96, 144
    const uint32_t w = o >> 6, sh = o & 63u;      // the piece from bit sh of word w on
    u64 kz = L.pl[2 * w] >> sh, kg = L.pl[2 * w + 1] >> sh;
    if (sh > 16) {                                // fewer than 48 bits there: the next word
        kz |= L.pl[2 * w + 2] << (64 - sh);
        kg |= L.pl[2 * w + 3] << (64 - sh);
    }
    nz = expand48(kz & 0xffffffffffffull);
    ng = expand48(kg & 0xffffffffffffull);
}

// M(q) from either tier (mertens_dr.hpp:707-712, :741-752).
__device__ __forceinline__ int32_t get_M(const TableParam& tp, i64 q) {
    if (q <= tp.ud) return tp.Md[q];
    return comp_M(tp.cl, (u64)(q - tp.ud - 1));
}

// ---------------------------------------------------------------------------
// Geometry of one G evaluation.
//   slot g of X: term t = g/2 (k = j_of_t(t)), half = g&1.
//   range r1: slot0 = +comb2(y), slot1 = -comb2(y/3)
//         r2: slot0 = +comb2(y), slot1 = -single(y/3)
//         r3: slot0 = +comb2(y);  r4: slot0 = +single(y)   (mertens_dr.hpp:1372-1377)
//   single(y) = 1 - S1 + kappa*M(A) - S2                     (:1081-1151)
//   comb2(y)  = -(S1odd -+ S1even) + ky*M(A) - ky2*M(B) - S2 (:1158-1354)
// ---------------------------------------------------------------------------
struct Frame {
    i64 y;        // argument of G (0 = empty slot)
    i64 A, B;     // S1/S2 split points for y and y/2
    i64 ky, ky2;  // kappa(y) = y/(A+1), kappa(y/2) = (y/2)/(B+1)
    int comb2;    // 1 = G(y) - G(y/2), 0 = G(y)
    int coef;     // mu(k) * slot sign
};

struct Seg {
    int type;     // SegType
    int piece;    // S2 piece id (1..12)
    int sign;     // +-1 inside the G value
    i64 first;    // first n (S1) or first t (S2)
    i64 stride;   // n step (S1)
    i64 count;    // positions
};

__device__ __forceinline__ i64 split_point(i64 y, double cx, i64 u) {
    i64 A = (i64)(cx * __builtin_sqrt((double)y));
    if (A >= y) A = y - 1;
    if (A > u - 1) A = u - 1;
    if (A < 1) A = 1;
    return A;
}

__device__ inline Frame make_frame(const XParam& xp, const uint32_t* mu6, i64 g) {
    Frame f{};
    const i64 t = g >> 1;
    const int half = (int)(g & 1);
    const i64 k = j_of_t(t);
    const int muk = mu6_mu(mu6, t);
    if (muk == 0 || k > xp.N) return f;
    const int range = k <= xp.n6 ? 1 : k <= xp.n3 ? 2 : k <= xp.n2 ? 3 : 4;
    i64 y = div_exact(xp.X, k);
    int sign = 1;
    if (half) {
        if (range >= 3) return f;
        y = y / 3;
        sign = -1;
        f.comb2 = (range == 1);
    } else {
        f.comb2 = (range <= 3);
    }
    f.y = y;
    f.coef = muk * sign;
    f.A = split_point(y, xp.cx, xp.u);
    f.ky = y / (f.A + 1);
    if (f.comb2) {
        const i64 y2 = y >> 1;
        f.B = split_point(y2, xp.cx, xp.u);
        f.ky2 = y2 / (f.B + 1);
    }
    return f;
}

__device__ __forceinline__ int frame_nseg(const Frame& f) {
    return f.y == 0 ? 0 : (f.comb2 ? 12 : 6);
}

// S1 run of stride s over [lo, hi] split at the dense/compressed boundary:
// n <= split gives q > ud (compressed), n > split gives q <= ud (dense).
// Compressed positions also need n >= nmin (q at most the pass's qhi).
__device__ __forceinline__ Seg s1_part(i64 lo, i64 hi, i64 s, i64 split, bool comp, int sign, i64 nmin) {
    Seg r{};
    r.type = comp ? SEG_S1_COMP : SEG_S1_DENSE;
    r.sign = sign;
    r.stride = s;
    i64 a = lo, b = hi;
    if (comp) {
        if (b > split) b = split;
        if (a < nmin) {
            a = nmin;
            if (s == 2 && ((a - lo) & 1)) ++a;
        }
    } else if (a <= split) {
        a = split + 1;
        if (s == 2 && ((a - lo) & 1)) ++a;
    }
    if (a > b) return r;
    r.first = a;
    r.count = (b - a) / s + 1;
    return r;
}

__device__ __forceinline__ Seg s2_part(i64 lo, i64 hi, int piece) {
    Seg r{};
    r.type = SEG_S2;
    r.piece = piece;
    r.sign = -1;
    if (lo < 1) lo = 1;
    if (lo > hi) return r;
    r.first = count_coprime6(lo - 1);
    r.count = count_coprime6(hi) - r.first;
    return r;
}

__device__ inline Seg frame_seg(const Frame& f, const XParam& xp, const TableParam& tp, int idx) {
    const i64 y = f.y;
    // compressed lookups of this pass: qlo <= q <= qhi, i.e. nmin <= n <= split;
    // a window pass (2) has no dense lookups, S2 pieces or constants
    const i64 split = y / tp.qlo;
    const i64 nmin = y / (tp.qhi + 1) + 1;
    const bool win = tp.pass == 2;
    if (win && (idx & 1) && idx < (f.comb2 ? 4 : 2)) return Seg{};   // dense parts
    if (win && idx >= (f.comb2 ? 4 : 2)) return Seg{};               // S2 pieces
    if (!f.comb2) {
        if (idx < 2) return s1_part(y / xp.u + 1, f.ky, 1, split, idx == 0, -1, nmin);
        const i64 A = f.A;
        const i64 cut[5] = {0, A / 6, A / 3, A / 2, A};
        return s2_part(cut[idx - 2] + 1, cut[idx - 1], 9 + (idx - 2));
    }
    if (idx < 2) {
        i64 lo = y / xp.u + 1;
        if (!(lo & 1)) ++lo;
        return s1_part(lo, f.ky, 2, split, idx == 0, -1, nmin);
    }
    if (idx < 4) {
        const i64 two = 2 * f.ky2;
        i64 lo, hi;
        int sign;
        if (two > f.ky) { lo = f.ky + 1; hi = two; sign = 1; }
        else if (two < f.ky) { lo = two + 1; hi = f.ky; sign = -1; }
        else return Seg{};
        if (lo & 1) ++lo;
        return s1_part(lo, hi, 2, split, idx == 2, sign, nmin);
    }
    const i64 A = f.A, B = f.B;
    const i64 cut[9] = {0, B / 6, A / 6, B / 3, A / 3, B / 2, A / 2, B, A};
    const int p = idx - 3;   // 1..8
    return s2_part(cut[p - 1] + 1, cut[p], p);
}

// Constant part of one G value: 1 + kappa*M(A) or ky*M(A) - ky2*M(B).
__device__ inline i128 frame_const(const Frame& f, const TableParam& tp) {
    if (f.comb2) return (i128)f.ky * get_M(tp, f.A) - (i128)f.ky2 * get_M(tp, f.B);
    return 1 + (i128)f.ky * get_M(tp, f.A);
}

// ---------------------------------------------------------------------------
// 128-bit wave reduction and accumulation.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ void add128(u64& lo, i64& hi, u64 blo, i64 bhi) {
    const u64 s = lo + blo;
    hi += bhi + (s < lo ? 1 : 0);
    lo = s;
}

__device__ __forceinline__ void wave_reduce128(u64& lo, i64& hi) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const u64 olo = __shfl_xor(lo, m, 64);
        const i64 ohi = __shfl_xor(hi, m, 64);
        add128(lo, hi, olo, ohi);
    }
}

// Exact: the low word's wrap-arounds are carried into the high word.
__device__ __forceinline__ void atomic_add128(u64* dst, u64 lo, i64 hi) {
    const u64 old = atomicAdd(dst, lo);
    const u64 carry = (old + lo < old) ? 1ull : 0ull;
    atomicAdd(dst + 1, (u64)hi + carry);
}

// Sum of one segment's positions [p_lo, p_hi) for the calling lane
// (defined in mdr_quot.hip).
__device__ i64 run_seg(const TableParam& tp, const Seg& sg, u64 y, u64 p_lo, u64 p_hi, int lane);

}  // namespace mdr

#endif
