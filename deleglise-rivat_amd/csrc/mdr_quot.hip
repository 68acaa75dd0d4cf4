// Quotient phase: partition of the G evaluations into chunks and the
// persistent kernel that evaluates the S1 table sums and the S2 piecewise
// mod-6 sums.
//
// Reference: the term list and outer reduction (mertens_dr.hpp:1050-1071,
// :1361-1380, :1469-1485, :1616-1632), eval_single_S / eval_comb2
// (:1081-1354, :1490-1614), sum_dense_quotients (:88-128),
// reconstruct_compressed (:741-752) and run_s2_fast (:484-617).  The
// reference gives one worker per outer term.  Here every G evaluation of every
// X is laid out in one global position space (position 0 of a G evaluation
// is its constant, then its S1 runs and S2 pieces in order); a boundary array
// cuts that space into P chunks -- at equal positions (static partition) or
// at equal predicted cycles (learned partition) -- and persistent waves sum
// the chunks.
#include <hip/hip_runtime.h>
#include "mdr_device.hpp"
#include "mdr_kernels.hpp"
#include "mdr_mlp.hpp"

namespace mdr {

namespace {

__device__ __forceinline__ int find_x(const XParam* xp, int nx, i64 G) {
    int lo = 0, hi = nx - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (xp[mid].slot_base <= G) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// last index i in [lo, n) with a[i] <= v (a non-decreasing, a[lo] <= v)
__device__ __forceinline__ i64 last_le(const i64* a, i64 n, i64 v, i64 lo = 0) {
    i64 hi = n - 1;
    while (lo < hi) {
        const i64 mid = (lo + hi + 1) >> 1;
        if (a[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace

// Position of a slot's constant term before its segments' positions: passes
// 0 and 1 have one, a window pass (2) none.
__device__ __forceinline__ i64 const_positions(const TableParam& tp) { return tp.pass != 2 ? 1 : 0; }

// Per-slot length (the constant's position, if the pass has one, plus the
// positions of its segments; 0 for an empty slot, which includes every slot
// with no lookups in a window pass) and, for the learned partition, predicted
// cycles (rounded per cost piece).
__global__ void k_slot_len(PartArgs a) {
    const i64 G = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 pos[3] = {0, 0, 0};
    if (G < a.total_slots) {
        const int x = find_x(a.xp, a.nx, G);
        const XParam xp = a.xp[x];
        const Frame f = make_frame(xp, a.tp.mu6, G - xp.slot_base);
        const int ns = frame_nseg(f);
        i64 L = 0, C = 0;
        if (ns > 0) {
            L = const_positions(a.tp);
            for (int i = 0; i < ns; ++i) {
                const Seg sg = frame_seg(f, xp, a.tp, i);
                L += sg.count;
                pos[sg.type == SEG_S1_DENSE ? 0 : sg.type == SEG_S1_COMP ? 1 : 2] += (u64)sg.count;
                if (a.mlp)
                    for (i64 p = 0; p < sg.count;) {
                        const i64 pe = piece_end(sg, p);
                        C += (i64)__builtin_rint(piece_cost(a.mlp, f.y, sg, p, pe));
                        p = pe;
                    }
            }
            if (L > 0) C += SLOT_OVERHEAD_CYCLES;
        }
        a.len[G] = L;
        if (a.mlp) a.cost[G] = C;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        u64 v = pos[k];
        for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(a.type_pos + k, (unsigned long long)v);
    }
}

// Chunk boundaries B[0..P]: with Pe = P (or fewer, see min_chunk) chunks in
// use, static c*T/Pe, or the position where the running predicted cost
// reaches c*C/Pe; B[c] = T for c >= Pe (monotone in c, B[0] = 0, B[P] = T).
__global__ void k_cuts(PartArgs a) {
    const i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > a.P) return;
    const i64 T = a.totals[0];
    i64 Pe = a.P;
    if (a.min_chunk > 0) {
        const i64 want = T / a.min_chunk;
        Pe = want < 1 ? 1 : (want < a.P ? want : a.P);
    }
    if (a.by_slot) {   // whole G evaluations: chunk c starts at slot c * slots / Pe
        const i64 G = (i64)((u128)(u64)c * (u64)a.total_slots / (u64)Pe);
        a.bounds[c] = (c >= Pe || G >= a.total_slots) ? T : a.len[G];
        return;
    }
    if (!a.mlp || c == 0 || c >= Pe) {
        a.bounds[c] = c >= Pe ? T : (i64)((u128)(u64)c * (u64)T / (u64)Pe);
        return;
    }
    const i64 Ctot = a.totals[1];
    const i64 target = (i64)((u128)(u64)c * (u64)Ctot / (u64)Pe);
    const i64 G = last_le(a.cost, a.total_slots, target);
    const i64 base = a.len[G];
    const i64 Lg = (G + 1 < a.total_slots ? a.len[G + 1] : T) - base;
    i64 local = target - a.cost[G];
    i64 pos = 0;
    if (Lg > 0 && local >= SLOT_OVERHEAD_CYCLES) {
        local -= SLOT_OVERHEAD_CYCLES;
        pos = const_positions(a.tp);
        const int x = find_x(a.xp, a.nx, G);
        const XParam xp = a.xp[x];
        const Frame f = make_frame(xp, a.tp.mu6, G - xp.slot_base);
        const int ns = frame_nseg(f);
        bool done = false;
        for (int i = 0; i < ns && !done; ++i) {
            const Seg sg = frame_seg(f, xp, a.tp, i);
            for (i64 p = 0; p < sg.count;) {
                const i64 pe = piece_end(sg, p);
                const i64 pc = (i64)__builtin_rint(piece_cost(a.mlp, f.y, sg, p, pe));
                if (local < pc) {
                    pos += (i64)((u128)(u64)(pe - p) * (u64)local / (u64)(pc > 0 ? pc : 1));
                    done = true;
                    break;
                }
                local -= pc;
                pos += pe - p;
                p = pe;
            }
        }
        if (pos > Lg) pos = Lg;
    }
    a.bounds[c] = base + pos;
}

// ---------------------------------------------------------------------------
// Chunk evaluators.  Lanes take consecutive positions (stride 64), so table
// gathers of neighbouring n coalesce wherever y/n^2 is small.
// ---------------------------------------------------------------------------
// Per-lane iteration count of a run of cnt positions taken with stride 64.
__device__ __forceinline__ uint32_t lane_iters(u64 cnt, int lane) {
    return (u64)lane < cnt ? (uint32_t)((cnt - (u64)lane + 63) >> 6) : 0u;
}

template <int ST>
__device__ i64 s1_dense(const TableParam& tp, u64 y, u64 nlo, u64 cnt, int lane) {
    const double dy = (double)y;
    const u64 nmax = nlo + (u64)ST * (cnt - 1);
    if (nmax < 2147483648ull) {   // 2n <= 2^32: qdiv32n
        // lane positions n, n + 64 ST, ...: two per step share one reciprocal.
        // int32 partial sums of at most 2^15 steps (|M| < 2^15 in the int16
        // tier) are widened before they can overflow.
        const uint32_t ylo = (uint32_t)y;
        const uint32_t iters = lane_iters(cnt, lane);
        uint32_t n = (uint32_t)nlo + (uint32_t)(ST * lane);
        constexpr uint32_t D = 64u * ST;
        i64 s = 0;
        for (uint32_t k0 = 0; k0 < iters; k0 += 32768u) {
            const uint32_t kend = iters - k0 < 32768u ? iters : k0 + 32768u;
            int32_t part = 0;
            uint32_t k = k0;
            for (; k + 2 <= kend; k += 2) {
                double r1, r2;
                rcp_pair((double)n, (double)(n + D), r1, r2);
                part += tp.Md[qdiv32n(ylo, dy, n, r1)];
                part += tp.Md[qdiv32n(ylo, dy, n + D, r2)];
                n += 2 * D;
            }
            if (k < kend) {
                part += tp.Md[qdiv32n(ylo, dy, n, rcp_fast((double)n))];
                n += D;
            }
            s += part;
        }
        return s;
    }
    i64 s = 0;
    for (u64 i = lane; i < cnt; i += 64) {
        const u64 n = nlo + (u64)ST * i;
        s += tp.Md[qdiv_mid(y, dy, n, rcp_fast((double)n))];
    }
    return s;
}

template <int ST>
__device__ i64 s1_comp(const TableParam& tp, u64 y, u64 nlo, u64 cnt, int lane) {
    const double dy = (double)y;
    const u64 nmax = nlo + (u64)ST * (cnt - 1);
    i64 s = 0;
    if (nmax < 2147483648ull) {
        // 2n <= 2^32, q <= U < 2^44.  q0 = nearest integer to the estimate
        // (2^52 + qd leaves it in the low mantissa bits) is q or q + 1, so the
        // true remainder y - q0*n lies in [-n, n) and (q0*n - y - 1) mod 2^32,
        // one multiply-add, is below n exactly when q0 = q + 1 (as in
        // qdiv32n); off = q - ud - 1 is the value's place in the compressed
        // tier.  Two positions per step share one reciprocal.
        const uint32_t nylo = ~(uint32_t)y;
        const u64 base = (u64)tp.ud + 1;
        const uint32_t iters = lane_iters(cnt, lane);
        const uint32_t n0 = (uint32_t)nlo + (uint32_t)(ST * lane);
        constexpr uint32_t D = 64u * ST;
        // int32 partial sums over at most 2^11 positions: the table holds
        // q <= U < 256 * 2^32 + 2^32, where |M(q)| < 0.6 sqrt(q) < 2^19.3
        // (|M(x)| / sqrt(x) stays below 0.6 through x = 10^16), so each
        // partial sum stays below 2^30.3
        auto run = [&](auto lookup) -> i64 {
            uint32_t n = n0;
            i64 acc = 0;
            for (uint32_t k0 = 0; k0 < iters; k0 += 2048u) {
                const uint32_t kend = iters - k0 < 2048u ? iters : k0 + 2048u;
                int32_t part = 0;
                uint32_t k = k0;
                for (; k + 2 <= kend; k += 2) {
                    double r1, r2;
                    rcp_pair((double)n, (double)(n + D), r1, r2);
                    part += lookup(n, r1);
                    part += lookup(n + D, r2);
                    n += 2 * D;
                }
                if (k < kend) {
                    part += lookup(n, rcp_fast((double)n));
                    n += D;
                }
                acc += part;
            }
            return acc;
        };
        // The estimate's bit pattern is 0x433 << 52 | q0, so one 64-bit
        // subtraction gives off = q0 - over - base (over folded into the low
        // word of the subtrahend: base = ud + 1 = 1 mod 256, so that word is
        // never 0xffffffff and cannot carry).  The line off >> 8 fits 32 bits
        // (the tier holds fewer than 256 * 2^32 values), the kept index of
        // off mod 256 gives word and bit.
        const u64 cK = 0x4330000000000000ull + base;
        const uint32_t cK_lo = (uint32_t)cK, cK_hi = (uint32_t)(cK >> 32);
        return run([&](uint32_t nn, double rn) -> int32_t {
            const double qm = __builtin_fma(dy, rn, 4503599627370496.0);
            const u64 bits = (u64)__double_as_longlong(qm);
            const uint32_t tr = (uint32_t)bits * nn + nylo;
            const uint32_t sub_lo = cK_lo + (tr < nn ? 1u : 0u);   // tr < nn: q0 = q + 1
            const u64 off = bits - ((u64)cK_hi << 32 | sub_lo);
            const uint32_t k = kept_index((uint32_t)off & (LINE_VALUES - 1u));
            // line off >> 8 by one funnel shift of the two words (a plain
            // 64-bit shift would be folded with the line size into a 64-bit
            // shift and two masks)
            const uint32_t line = __builtin_amdgcn_alignbit((uint32_t)(off >> 32), (uint32_t)off, 8u);
            return line_M(tp.cl + line, k >> 6, k & 63u);
        });
    }
    for (u64 i = lane; i < cnt; i += 64) {
        const u64 n = nlo + (u64)ST * i;
        const u64 q = qdiv_mid(y, dy, n, rcp_fast((double)n));
        s += comp_M(tp.cl, q - (u64)tp.ud - 1);
    }
    return s;
}

// S2 summand of a quotient q < 2^32 in 32-bit modular arithmetic: every
// value lies in [0, 2^32) (pieces other than 4) or in (-2^29, 2^29) (piece 4),
// so the wrapped 32-bit result is exact.  The floor forms of piece_eval are
// folded with q - 2 floor(q/2) = q mod 2, t - 2 floor(t/2) = t mod 2 and
// t - floor(t/2) = ceil(t/2) for t = floor(q/3), and floor(q/12) =
// floor(t/4) (checked against the unfolded forms on the host).
template <int P>
__device__ __forceinline__ uint32_t piece32(uint32_t q) {
    const uint32_t odd = q & 1u, f = q >> 2;
    if constexpr (P == 6) return odd;
    else if constexpr (P == 7 || P == 11) return q - (q >> 1);
    else if constexpr (P == 8 || P == 12) return q;
    else if constexpr (P == 5) return odd + f;
    else {
        const uint32_t t = __umulhi(q, 0xaaaaaaabu) >> 1;   // q / 3
        if constexpr (P == 1) return odd - (t & 1u) + f - (t >> 2);
        else if constexpr (P == 2) return odd - (t & 1u) + f;
        else if constexpr (P == 3) return odd + f - (t - (t >> 1));
        else if constexpr (P == 4) return odd + f - t;
        else if constexpr (P == 9) return (q - (q >> 1)) - (t - (t >> 1));
        else return (q - (q >> 1)) - t;   // P == 10
    }
}

// mu * v of one position.  Piece 6 (v = q mod 2, so mu * v is -1, 0 or 1)
// adds into an int32 partial sum: a lane of the 32-bit loop takes at most
// 2^31 / 192 positions of a run (j < 2^31, stride 192), far below 2^31.  The
// other pieces take one signed 32 x 32 -> 64-bit multiply-add.
template <int P>
__device__ __forceinline__ void s2_add(i64& s, int32_t& s6, uint32_t word, uint32_t sh, uint32_t q) {
    const int mu = mu6_field(word, sh);
    if constexpr (P == 6) s6 += (q & 1u) ? mu : 0;
    else s += (i64)mu * (i64)(int32_t)piece32<P>(q);
}

template <int P>
__device__ i64 s2_run(const TableParam& tp, u64 y, u64 t_lo, u64 t_hi, int lane) {
    const double dy = (double)y;
    // j >= (y >> 31) + 2 gives q <= 2^31 - 2, so every summand fits int32 and
    // mu * v is one signed multiply-add into 64 bits; the 32-bit remainder
    // test also needs 2j <= 2^32 (qdiv32n).  Smaller j (few positions) take
    // the 64-bit path.
    u64 ts = (u64)count_coprime6((i64)(y >> 31) + 1);
    if (j_of_t((i64)t_hi - 1) >= 2147483648ll) ts = t_hi;
    ts = ts < t_lo ? t_lo : (ts > t_hi ? t_hi : ts);
    i64 s = 0;
    for (u64 t = t_lo + lane; t < ts; t += 64) {
        const u64 j = (u64)j_of_t((i64)t);
        const u64 q = qdiv_big(y, dy, j, rcp_fast((double)j));
        s += (i64)mu6_mu(tp.mu6, (i64)t) * piece_eval<P, u64>(q);
    }
    // A lane's t advances by 64, so j by 192, its mu word by 4 and the
    // field's shift stays; four positions per step share one reciprocal.
    const uint32_t iters = lane_iters(t_hi - ts, lane);
    if (iters) {
        const uint32_t ylo = (uint32_t)y;
        const uint32_t t0 = (uint32_t)ts + (uint32_t)lane;
        const uint32_t* wp = tp.mu6 + (t0 >> 4);
        int32_t s6 = 0;
        const uint32_t sh = 2u * (t0 & 15u);
        uint32_t j = 3u * t0 + 1u + (t0 & 1u);
        // the loop runs on j alone: j_end < 2^31 + 768 * 2^24 fits 32 bits
        const uint32_t j_end = j + 768u * (iters >> 2);
        for (; j != j_end;) {
            double r1, r2, r3, r4;
            rcp_quad((double)j, (double)(j + 192u), (double)(j + 384u), (double)(j + 576u), r1, r2, r3, r4);
            const uint32_t q1 = qdiv32n(ylo, dy, j, r1);
            const uint32_t q2 = qdiv32n(ylo, dy, j + 192u, r2);
            const uint32_t q3 = qdiv32n(ylo, dy, j + 384u, r3);
            const uint32_t q4 = qdiv32n(ylo, dy, j + 576u, r4);
            s2_add<P>(s, s6, wp[0], sh, q1);
            s2_add<P>(s, s6, wp[4], sh, q2);
            s2_add<P>(s, s6, wp[8], sh, q3);
            s2_add<P>(s, s6, wp[12], sh, q4);
            j += 768u;
            wp += 16;
        }
        if (iters & 2u) {
            double r1, r2;
            rcp_pair((double)j, (double)(j + 192u), r1, r2);
            const uint32_t q1 = qdiv32n(ylo, dy, j, r1);
            const uint32_t q2 = qdiv32n(ylo, dy, j + 192u, r2);
            s2_add<P>(s, s6, wp[0], sh, q1);
            s2_add<P>(s, s6, wp[4], sh, q2);
            j += 384u;
            wp += 8;
        }
        if (iters & 1u) s2_add<P>(s, s6, wp[0], sh, qdiv32n(ylo, dy, j, rcp_fast((double)j)));
        s += s6;
    }
    return s;
}

__device__ i64 run_seg(const TableParam& tp, const Seg& sg, u64 y, u64 p_lo, u64 p_hi, int lane) {
    const u64 cnt = p_hi - p_lo;
    if (cnt == 0) return 0;
    if (sg.type == SEG_S1_DENSE) {
        const u64 nlo = (u64)sg.first + (u64)sg.stride * p_lo;
        return sg.stride == 1 ? s1_dense<1>(tp, y, nlo, cnt, lane) : s1_dense<2>(tp, y, nlo, cnt, lane);
    }
    if (sg.type == SEG_S1_COMP) {
        const u64 nlo = (u64)sg.first + (u64)sg.stride * p_lo;
        return sg.stride == 1 ? s1_comp<1>(tp, y, nlo, cnt, lane) : s1_comp<2>(tp, y, nlo, cnt, lane);
    }
    const u64 t_lo = (u64)sg.first + p_lo, t_hi = (u64)sg.first + p_hi;
    switch (sg.piece) {
        case 1: return s2_run<1>(tp, y, t_lo, t_hi, lane);
        case 2: return s2_run<2>(tp, y, t_lo, t_hi, lane);
        case 3: return s2_run<3>(tp, y, t_lo, t_hi, lane);
        case 4: return s2_run<4>(tp, y, t_lo, t_hi, lane);
        case 5: return s2_run<5>(tp, y, t_lo, t_hi, lane);
        case 6: return s2_run<6>(tp, y, t_lo, t_hi, lane);
        case 7: return s2_run<7>(tp, y, t_lo, t_hi, lane);
        case 8: return s2_run<8>(tp, y, t_lo, t_hi, lane);
        case 9: return s2_run<9>(tp, y, t_lo, t_hi, lane);
        case 10: return s2_run<10>(tp, y, t_lo, t_hi, lane);
        case 11: return s2_run<11>(tp, y, t_lo, t_hi, lane);
        default: return s2_run<12>(tp, y, t_lo, t_hi, lane);
    }
}

namespace {
__device__ __forceinline__ i64 uniform64(i64 v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(u64)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((u64)v >> 32));
    return (i64)(((u64)hi << 32) | lo);
}
__device__ __forceinline__ int uniform32(int v) { return (int)__builtin_amdgcn_readfirstlane((uint32_t)v); }
// Frame and segment values are the same in every lane of a wave: reading
// them back from lane 0 lets the compiler keep them in scalar registers
// (fewer vector registers, more resident waves).
__device__ __forceinline__ Frame uniform_frame(Frame f) {
    f.y = uniform64(f.y);
    f.A = uniform64(f.A);
    f.B = uniform64(f.B);
    f.ky = uniform64(f.ky);
    f.ky2 = uniform64(f.ky2);
    f.comb2 = uniform32(f.comb2);
    f.coef = uniform32(f.coef);
    return f;
}
__device__ __forceinline__ XParam uniform_xp(XParam p) {
    p.X = uniform64(p.X);
    p.u = uniform64(p.u);
    p.N = uniform64(p.N);
    p.T = uniform64(p.T);
    p.slot_base = uniform64(p.slot_base);
    p.cx = __longlong_as_double(uniform64(__double_as_longlong(p.cx)));
    p.n6 = uniform64(p.n6);
    p.n3 = uniform64(p.n3);
    p.n2 = uniform64(p.n2);
    return p;
}
__device__ __forceinline__ Seg uniform_seg(Seg s) {
    s.type = uniform32(s.type);
    s.piece = uniform32(s.piece);
    s.sign = uniform32(s.sign);
    s.first = uniform64(s.first);
    s.stride = uniform64(s.stride);
    s.count = uniform64(s.count);
    return s;
}
__device__ __forceinline__ i64 fetch_next(unsigned long long* ctr, int lane) {
    unsigned long long v = 0;
    if (lane == 0) v = atomicAdd(ctr, 1ull);
    return uniform64((i64)__shfl(v, 0, 64));
}
__device__ __forceinline__ void flush(u64* out128, i64 x, u64& lo, i64& hi, int lane) {
    wave_reduce128(lo, hi);
    if (lane == 0 && (lo | (u64)hi)) atomic_add128(out128 + 2 * x, lo, hi);
    lo = 0;
    hi = 0;
}
}  // namespace

__global__ __launch_bounds__(256) void k_quot(QuotArgs a) {
    const int lane = threadIdx.x & 63;
    const i64 wave = ((i64)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = ((i64)gridDim.x * blockDim.x) >> 6;
    const i64 T = a.bounds[a.P];
    u64 acc_lo = 0;
    i64 acc_hi = 0;
    i64 cur_x = -1, x_end = 0;
    XParam xp{};
    const u64 t_start = a.wave_span ? (u64)wall_clock64() : 0;
    for (i64 c = a.dynamic ? fetch_next(a.next, lane) : uniform64(wave); c < a.P;
         c = a.dynamic ? fetch_next(a.next, lane) : c + nwaves) {
        const i64 b0 = uniform64(a.bounds[c]), b1 = uniform64(a.bounds[c + 1]);
        if (b0 >= T) break;   // bounds ascend: every later chunk is empty too
        if (b0 >= b1) continue;
        for (i64 G = uniform64(last_le(a.off, a.nslots, b0)); G < a.nslots; ++G) {
            const i64 og = uniform64(a.off[G]);
            if (og >= b1) break;
            const i64 og1 = uniform64(G + 1 < a.nslots ? a.off[G + 1] : T);
            if (og1 <= og) {
                // an empty slot (a window pass leaves most of them empty): go
                // straight to the slot holding position og, the last one with
                // off <= og (after G, since off[G + 1] = og), instead of
                // walking the empty slots one load pair at a time
                G = uniform64(last_le(a.off, a.nslots, og, G + 1)) - 1;
                continue;
            }
            if (G >= x_end || cur_x < 0) {
                if (cur_x >= 0) flush(a.out128, cur_x, acc_lo, acc_hi, lane);
                cur_x = uniform64(find_x(a.xp, a.nx, G));
                xp = uniform_xp(a.xp[cur_x]);
                x_end = uniform64(cur_x + 1 < a.nx ? a.xp[cur_x + 1].slot_base : a.nslots);
            }
            const Frame f = uniform_frame(make_frame(xp, a.tp.mu6, G - xp.slot_base));
            const i64 lo = (b0 > og ? b0 : og) - og, hi = (b1 < og1 ? b1 : og1) - og;
            if (lo == 0 && lane == 0 && a.tp.pass != 2) {   // a window pass has no constants
                const i128 cst = (i128)f.coef * frame_const(f, a.tp);
                add128(acc_lo, acc_hi, (u64)cst, (i64)(cst >> 64));
            }
            const bool stamp = a.stamps && (((u64)G * 0x9E3779B97F4A7C15ull) >> 40) % (u64)a.stamp_stride == 0;
            const int ns = frame_nseg(f);
            i64 s0 = const_positions(a.tp);
            for (int i = 0; i < ns && s0 < hi; ++i) {
                const Seg sg = uniform_seg(frame_seg(f, xp, a.tp, i));
                const i64 s1 = s0 + sg.count;
                const i64 l = lo > s0 ? lo : s0, h = hi < s1 ? hi : s1;
                if (l < h) {
                    const int sgn = f.coef * sg.sign;
                    if (!stamp) {
                        const i64 v = sgn * run_seg(a.tp, sg, (u64)f.y, (u64)(l - s0), (u64)(h - s0), lane);
                        add128(acc_lo, acc_hi, (u64)v, v >> 63);
                    } else {
                        // time each cost piece separately (training data)
                        for (i64 p = 0; p < sg.count;) {
                            const i64 pe = piece_end(sg, p);
                            const i64 pl = (l - s0) > p ? (l - s0) : p, ph = (h - s0) < pe ? (h - s0) : pe;
                            if (pl < ph) {
                                const long long t0 = clock64();
                                const i64 v = sgn * run_seg(a.tp, sg, (u64)f.y, (u64)pl, (u64)ph, lane);
                                add128(acc_lo, acc_hi, (u64)v, v >> 63);
                                const long long dt = clock64() - t0;
                                if (lane == 0) {
                                    const unsigned long long k = atomicAdd(a.nstamps, 1ull);
                                    if ((i64)k < a.stamp_cap) {
                                        StampRec r;
                                        r.y = f.y;
                                        r.start = seg_value(sg, pl);
                                        r.end = seg_value(sg, ph - 1);
                                        r.kind = (uint32_t)seg_kind(sg);
                                        r.cycles = (uint32_t)(dt > 0xffffffffll ? 0xffffffffll : (dt < 0 ? 0 : dt));
                                        r.count = (u64)(ph - pl);
                                        a.stamps[k] = r;
                                    }
                                }
                            }
                            p = pe;
                        }
                    }
                }
                s0 = s1;
            }
        }
    }
    if (cur_x >= 0) flush(a.out128, cur_x, acc_lo, acc_hi, lane);
    if (a.wave_span && lane == 0) {
        a.wave_span[2 * wave] = t_start;
        a.wave_span[2 * wave + 1] = (u64)wall_clock64();
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan of i64 (three passes: tile sums, scan of tile sums, tiles).
// ---------------------------------------------------------------------------
namespace {
constexpr int SCAN_T = 256;
constexpr int SCAN_PER = 8;
constexpr int SCAN_TILE = SCAN_T * SCAN_PER;

__device__ i64 block_excl_scan(i64 v, i64* total) {
    __shared__ i64 wsum[SCAN_T / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    i64 inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const i64 o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    i64 off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_T / 64; ++w) {
        if (w < wave) off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}
}  // namespace

__global__ void k_scan_reduce(const i64* in, i64 n, i64* tsum) {
    const i64 base = (i64)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER;
    i64 s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) s += (base + i < n) ? in[base + i] : 0;
    i64 tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

__global__ void k_scan_top(i64* tsum, i64 nt, i64* total, i64* total2) {
    i64 carry = 0;
    for (i64 b = 0; b < nt; b += SCAN_TILE) {
        const i64 base = b + threadIdx.x * SCAN_PER;
        i64 v[SCAN_PER];
        i64 s = 0;
#pragma unroll
        for (int i = 0; i < SCAN_PER; ++i) { v[i] = (base + i < nt) ? tsum[base + i] : 0; s += v[i]; }
        i64 tot;
        i64 ex = block_excl_scan(s, &tot) + carry;
#pragma unroll
        for (int i = 0; i < SCAN_PER; ++i) {
            if (base + i < nt) tsum[base + i] = ex;
            ex += v[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        *total = carry;
        if (total2) *total2 = carry;
    }
}

__global__ void k_scan_down(const i64* in, i64* out, i64 n, const i64* tsum) {
    const i64 base = (i64)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER;
    i64 v[SCAN_PER];
    i64 s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) { v[i] = (base + i < n) ? in[base + i] : 0; s += v[i]; }
    i64 tot;
    i64 ex = block_excl_scan(s, &tot) + tsum[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) {
        if (base + i < n) out[base + i] = ex;
        ex += v[i];
    }
}

__global__ void k_clear3(u64* p0, i64 n0, u64* p1, i64 n1, u64* p2, i64 n2) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n0) p0[i] = 0;
    if (i < n1) p1[i] = 0;
    if (i < n2) p2[i] = 0;
}

i64 scan_tmp_len(i64 n) { return (std::max<i64>(n, 1) + SCAN_TILE - 1) / SCAN_TILE + 1; }

// d_tmp must hold scan_tmp_len(n) values; d_tmp[last] receives the total, and
// so does *dev_total when it is not null.  Only kernel launches: capturable.
hipError_t scan_exclusive_i64(const i64* in, i64* out, i64 n, i64* d_tmp, i64 tmp_len, i64* dev_total,
                              hipStream_t st) {
    const i64 nt = (std::max<i64>(n, 1) + SCAN_TILE - 1) / SCAN_TILE;
    if (tmp_len < nt + 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nt), dim3(SCAN_T), 0, st, in, n, d_tmp);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_T), 0, st, d_tmp, nt, d_tmp + nt, dev_total);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nt), dim3(SCAN_T), 0, st, in, out, n, d_tmp);
    return hipGetLastError();
}

}  // namespace mdr
