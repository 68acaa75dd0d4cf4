// Table construction on the GPU: segmented byte-log Moebius sieve in LDS,
// dense int16 prefix, compressed bit planes with checkpoints, packed mu.
//
// Reference: SieveTable / sieve_mu_odd_byte_log / build_dense_prefix
// (mertens_dr.hpp:173-457) and CompressedMertensTable::build (:755-901).
// The byte-log rule is the same: each sieving prime p adds the odd weight
// bit_width(p-1)|1, a prime square sets 0x80, and the final byte gives mu
// by parity plus a comparison with floor(log2 n).  The GPU keeps one segment
// per workgroup in LDS, adds weights with LDS atomics (no carry can leave a
// byte: weights stay below 0x80), writes group planes from wavefront ballots
// and scans group sums per segment; a second pass adds segment offsets.
#include <hip/hip_runtime.h>
#include "mdr_device.hpp"
#include "mdr_kernels.hpp"

namespace mdr {

namespace {


__device__ __forceinline__ uint32_t mod_u64(u64 x, uint32_t p, double rp) {
    // x < 2^52; estimate within 1 of floor(x/p), then fix.
    i64 q = (i64)((double)x * rp);
    i64 r = (i64)x - q * (i64)p;
    if (r < 0) r += p;
    else if (r >= (i64)p) r -= p;
    return (uint32_t)r;
}


}  // namespace

template <int SEGT, int THREADS>
__global__ __launch_bounds__(THREADS) void k_sieve_t(SieveArgs a) {
    constexpr int GROUPS = SEGT / 64;      // 64-value words per segment
    __shared__ __align__(16) uint32_t acc32[SEGT / 4];
    __shared__ __align__(16) uint32_t nib32[SEGT / 16];   // per dword: nonzero nibble | negative nibble << 4
    __shared__ int32_t gsum[GROUPS];
    __shared__ int32_t gpre[GROUPS];
    // kept planes of the segment's lines: per line 12 16-bit pieces of the
    // nonzero plane (192 bits), then 12 of the negative plane
    __shared__ __align__(8) uint16_t stg[SEGT / LINE_VALUES][2 * LINE_KEPT / 16];
    __shared__ uint32_t small_off[64];
    __shared__ uint32_t sq_off[64];
    uint8_t* acc = reinterpret_cast<uint8_t*>(acc32);

    const i64 s = (i64)blockIdx.x + a.seg_first;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const u64 base = (u64)s * SEGT;
    const u64 lastn = (u64)(base + SEGT <= (u64)a.U ? base + SEGT : (u64)a.U);
    const int len = (int)(lastn - base);

    // 1. periodic stencils: weights of 2,3,5,7,11 and flags of 4, 9, 25, plus
    //    the weights of 13, 17, 19, 23 and of 29, 31, 37 from two more
    //    patterns (byte sums stay below 0x80: no carries)
    {
        const uint32_t phase = (uint32_t)(base % STENCIL_PERIOD);
        const uint4* src = reinterpret_cast<const uint4*>(a.stencil + phase);
        const uint32_t* pat =
            reinterpret_cast<const uint32_t*>(a.pattern + base % ((u64)PATTERN_REPEAT * PATTERN_PERIOD));
        const uint32_t* pat3 =
            reinterpret_cast<const uint32_t*>(a.pattern3 + base % ((u64)PATTERN_REPEAT * PATTERN3_PERIOD));
        uint4* dst = reinterpret_cast<uint4*>(acc32);
        for (int i = tid; i < SEGT / 16; i += THREADS) {
            uint4 v = src[i];
            v.x += pat[4 * i] + pat3[4 * i];
            v.y += pat[4 * i + 1] + pat3[4 * i + 1];
            v.z += pat[4 * i + 2] + pat3[4 * i + 2];
            v.w += pat[4 * i + 3] + pat3[4 * i + 3];
            dst[i] = v;
        }
    }
    // offsets of the block-cooperative primes
    const u64 first = base + 1;
    if (tid < a.n_small) {
        const uint32_t p = a.wprimes[tid];
        const uint32_t r = mod_u64(first, p, a.wrcp[tid]);
        small_off[tid] = r ? p - r : 0;
    } else if (tid >= 64 && tid - 64 < a.n_sq_small) {
        const int i = tid - 64;
        const uint32_t p = a.sqprimes[i];
        const uint32_t p2 = p * p;
        const uint32_t r = mod_u64(first, p2, 1.0 / (double)p2);
        sq_off[i] = r ? p2 - r : 0;
    }
    __syncthreads();

    // 2. square flags for 7 <= p < 256: wave w takes primes i = w (mod 8), its
    //    lanes the consecutive multiples (a block-wide loop per prime costs
    //    more in per-thread set-up than the few hits it has).
    constexpr int NWAVES = THREADS / 64;
    for (int i = wave; i < a.n_sq_small; i += NWAVES) {
        const uint32_t p = a.sqprimes[i];
        const uint32_t p2 = p * p;
        if ((u64)p2 > lastn) break;
        for (uint32_t x = sq_off[i] + (uint32_t)lane * p2; x < (uint32_t)len; x += 64u * p2)
            atomicOr(&acc32[x >> 2], 0x80u << (8 * (x & 3)));
    }
    // 3. weights for 41 <= p < 256 (13..37 came with the patterns): the
    //    block-wide loop for p < 64 (many hits per thread), the wave loop above it
    constexpr int NPAT = PATTERN_PRIMES + PATTERN3_PRIMES;
    int i3 = a.n_small < NPAT ? a.n_small : NPAT;
    for (; i3 < a.n_small; ++i3) {
        const uint32_t p = a.wprimes[i3];
        if (p >= 64) break;
        const uint32_t w = a.wweights[i3];
        for (uint32_t x = small_off[i3] + (uint32_t)tid * p; x < (uint32_t)len; x += THREADS * p)
            atomicAdd(&acc32[x >> 2], w << (8 * (x & 3)));
    }
    for (int i = i3 + wave; i < a.n_small; i += NWAVES) {
        const uint32_t p = a.wprimes[i];
        const uint32_t w = a.wweights[i];
        for (uint32_t x = small_off[i] + (uint32_t)lane * p; x < (uint32_t)len; x += 64u * p)
            atomicAdd(&acc32[x >> 2], w << (8 * (x & 3)));
    }
    // 4. primes 256 <= p <= sqrt(last n), one per thread, four in flight.
    //    With the base b = sb * SIEVE_UNIT, each prime's record gives
    //    (-b) mod m by one Shoup multiplication (m = p^2 below SQ_SPLIT, p
    //    from it on), and (-first) mod m = ((-b) mod m - 1) mod m is an
    //    offset in the segment.  Below SQ_SPLIT it is the first square hit
    //    and its residue mod p the first weight hit; from SQ_SPLIT on it is
    //    the first weight hit (at most a few per segment), and a hit at n is
    //    a square hit when p divides n / p = n * p^-1 (mod 2^32; n / p < 2^32).
    //    Primes ascend, so a thread stops at the first p above sqrt(last n).
    {
        constexpr int UNR = 4;
        // the walks read up to (2 UNR - 1) THREADS records past a live one,
        // the wave-per-prime loop up to 64 NWAVES past the cooperative primes
        static_assert((2 * UNR - 1) * THREADS <= PRIME_RECORD_PAD && 64 * NWAVES <= PRIME_RECORD_PAD,
                      "record reads past the end stay inside the end records");
        const uint32_t sb = (uint32_t)(base / (u64)SIEVE_UNIT);
        const uint32_t ulen = (uint32_t)len;
        const uint32_t f_lo = (uint32_t)first;
        uint32_t proot = (uint32_t)__builtin_sqrt((double)lastn);
        while ((u64)proot * proot > lastn) --proot;
        while ((u64)(proot + 1) * (proot + 1) <= lastn) ++proot;
        // walks the records from index i_begin (relative to the first large
        // prime) while their primes stay <= bound: each thread takes UNR
        // records per step, one from each run of THREADS records, in
        // alternating directions (thread t takes t, 2 THREADS - 1 - t, ...,
        // so that the small primes' many hits spread over the threads), and
        // loads its next step's records before marking these, so the loads
        // overlap the marking.  Primes ascend and the array ends with
        // PRIME_RECORD_PAD records of p above every bound, so reading ahead
        // needs no index check.
        auto walk = [&](int i_begin, uint32_t bound, auto&& mark) {
            const PrimeConsts* rec = a.wlarge + i_begin;
            auto at = [&](int u) { return rec[u * THREADS + ((u & 1) ? THREADS - 1 - tid : tid)]; };
            PrimeConsts cur[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) cur[u] = at(u);
            while ((cur[0].pw & 0xffffffu) <= bound) {   // cur[0] has the step's lowest index
                rec += UNR * THREADS;
                PrimeConsts nxt[UNR];
#pragma unroll
                for (int u = 0; u < UNR; ++u) nxt[u] = at(u);
#pragma unroll
                for (int u = 0; u < UNR; ++u)
                    if ((cur[u].pw & 0xffffffu) <= bound) mark(cur[u]);
#pragma unroll
                for (int u = 0; u < UNR; ++u) cur[u] = nxt[u];
            }
        };
        // p < COOP_PRIME_LIMIT (many hits each): one wavefront per prime, its
        // lanes on consecutive multiples.  Wave v takes the primes v, v + NWAVES,
        // ... (a mix of sizes for every wave): lane l loads the record of its
        // l-th prime and computes its offsets, then the wave walks its primes
        // one by one with those values read from the lanes (scalars).
        {
            const int wv = (int)__builtin_amdgcn_readfirstlane((uint32_t)wave);
            const int ncoop = a.n_coop - a.n_small;
            for (int i0 = 0; i0 < ncoop; i0 += 64 * NWAVES) {
                const int i = i0 + wv + NWAVES * lane;          // < n_coop + 64 NWAVES: inside the padded array
                const PrimeConsts pc = a.wlarge[i];
                const uint32_t p = pc.pw & 0xffffffu;
                const bool live = i < ncoop && p <= proot;      // a prefix of the lanes
                const uint32_t m = p * p;
                uint32_t x2 = 0, x1 = 0;
                if (live) {
                    x2 = first_multiple(sb, pc.c, pc.s, m);
                    x1 = mod_below_square(x2, p, pc.aux);
                }
                const int cnt = __popcll(__ballot(live));
                for (int j = 0; j < cnt; ++j) {
                    const uint32_t pj = (uint32_t)__builtin_amdgcn_readlane((int)p, j);
                    const uint32_t mj = pj * pj;
                    const uint32_t wj = (uint32_t)__builtin_amdgcn_readlane((int)pc.pw, j) >> 24;
                    const uint32_t x2j = (uint32_t)__builtin_amdgcn_readlane((int)x2, j);
                    const uint32_t x1j = (uint32_t)__builtin_amdgcn_readlane((int)x1, j);
                    const uint32_t xs = x2j + (uint32_t)lane * mj;   // 64 p^2 > SEG_MAX: one pass
                    if (xs < ulen) atomicOr(&acc32[xs >> 2], 0x80u << (8 * (xs & 3)));
                    for (uint32_t x = x1j + (uint32_t)lane * pj; x < ulen; x += 64u * pj)
                        atomicAdd(&acc32[x >> 2], wj << (8 * (x & 3)));
                }
            }
        }
        // COOP_PRIME_LIMIT <= p < SQ_SPLIT: square hits from the remainder mod p^2
        walk(a.n_coop - a.n_small, proot < SQ_SPLIT ? proot : SQ_SPLIT - 1u, [&](const PrimeConsts& pc) {
            const uint32_t p = pc.pw & 0xffffffu;
            const uint32_t m = p * p;
            const uint32_t r = first_multiple(sb, pc.c, pc.s, m);   // p^2 > SEG_MAX: at most one square hit
            if (r < ulen) atomicOr(&acc32[r >> 2], 0x80u << (8 * (r & 3)));
            const uint32_t w = pc.pw >> 24;
            for (uint32_t x = mod_below_square(r, p, pc.aux); x < ulen; x += p)
                atomicAdd(&acc32[x >> 2], w << (8 * (x & 3)));
        });
        // p >= SQ_SPLIT: each hit tested for p^2 | n; from the segment length
        // on a prime hits a segment at most once
        walk(a.n_split - a.n_small, proot, [&](const PrimeConsts& pc) {
            const uint32_t p = pc.pw & 0xffffffu;
            const uint32_t w = pc.pw >> 24, pinv = pc.aux;
            uint32_t x = first_multiple(sb, pc.c, pc.s, p);
            if (p >= ulen) {
                if (x < ulen) {
                    const uint32_t sh = 8u * (x & 3u);
                    atomicAdd(&acc32[x >> 2], w << sh);
                    if (square_divides(f_lo + x, p, pinv)) atomicOr(&acc32[x >> 2], 0x80u << sh);
                }
                return;
            }
            for (; x < ulen; x += p) {
                const uint32_t sh = 8u * (x & 3u);
                atomicAdd(&acc32[x >> 2], w << sh);
                if (square_divides(f_lo + x, p, pinv)) atomicOr(&acc32[x >> 2], 0x80u << sh);
            }
        });
    }
    __syncthreads();

    // 5. finalise to mu in {-1, 0, 1}, 16 values per lane (four dwords, SWAR):
    //    a flag gives 0; otherwise mu = +1 exactly when the parity bit differs
    //    from [acc > floor(log2 n)].  floor(log2 n) is tseg below the next power
    //    of two b and tseg + 1 from b on (segments from 2^17 on are shorter than
    //    b); the group holding b - 1 and b, groups reaching past U, n = 1, 2 and
    //    the first segments take the per-byte path.  Each dword also leaves a
    //    byte of its nonzero and negative bits (low/high nibble) for the plane
    //    builder.
    {
        const unsigned tseg = 63u - (unsigned)__clzll(first);
        const u64 db = (2ull << tseg) - first;                  // b - first >= 1
        const uint32_t gb = db >= (u64)SEGT ? 0xffffffffu : (uint32_t)(db >> 4);
        const bool straddle = (db & 15) != 0;
        const uint32_t gend = (uint32_t)len >> 4;               // groups wholly <= U
        const bool general = first < (1ull << 17);
        const uint32_t add_lo = (127u - tseg) * 0x01010101u, add_hi = (126u - tseg) * 0x01010101u;
        // the mu bytes are read again only for the packed mu codes and the
        // dense tier; elsewhere the SWAR groups go straight to their nibbles
        const bool keep_mu = s < a.mu_segs || first <= (u64)a.ud;
        uint4* acc4 = reinterpret_cast<uint4*>(acc32);
        for (uint32_t g = tid; g < SEGT / 16; g += THREADS) {
            uint4 v = acc4[g];
            if (!general && g < gend && !(straddle && g == gb)) {
                const uint32_t add = g < gb ? add_lo : add_hi;
                if (!keep_mu) {
                    nib32[g] = mu_nibbles_direct(v.x, add) | (mu_nibbles_direct(v.y, add) << 8) |
                               (mu_nibbles_direct(v.z, add) << 16) | (mu_nibbles_direct(v.w, add) << 24);
                    continue;
                }
                v.x = mu_swar(v.x, add);
                v.y = mu_swar(v.y, add);
                v.z = mu_swar(v.z, add);
                v.w = mu_swar(v.w, add);
            } else {
                uint32_t* w = &v.x;
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    uint32_t mu = 0;
#pragma unroll
                    for (int bb = 0; bb < 4; ++bb) {
                        const u64 n = first + 16 * (u64)g + 4 * d + bb;
                        int8_t m = 0;
                        if (n <= (u64)a.U) {
                            m = mu_byte((uint8_t)(w[d] >> (8 * bb)), n);
                            if (n == 1) m = 1;
                            if (n == 2) m = -1;
                        }
                        mu |= (uint32_t)(uint8_t)m << (8 * bb);
                    }
                    w[d] = mu;
                }
            }
            acc4[g] = v;
            nib32[g] = mu_nibbles(v.x) | (mu_nibbles(v.y) << 8) | (mu_nibbles(v.z) << 16) | (mu_nibbles(v.w) << 24);
        }
    }
    __syncthreads();

    // 6. one lane per 64-value word: word sums from 16 nibble bytes and, for
    //    words of a compressed line, the word's 48 kept bits per plane
    //    (pack12 drops the multiples of 4: the word starts at n = 1 mod 4) as
    //    three 16-bit pieces at kept offset 48 (word mod 4) of its line's
    //    staging.  Words past U are zero, so a line's staging is whole.
    const i64 gl0 = ((i64)base - a.ud) / LINE_VALUES;   // line of the segment's first value (< 0: dense)
    for (int k = tid; k < GROUPS; k += THREADS) {
        const uint4 b = reinterpret_cast<const uint4*>(nib32)[k];
        const u64 nz = kept48(b.x, b.y, b.z, b.w), ng = kept48(b.x >> 4, b.y >> 4, b.z >> 4, b.w >> 4);
        gsum[k] = (int)__popcll(nz) - 2 * (int)__popcll(ng);
        const int ll = k >> 2;
        if (gl0 + ll >= 0 && gl0 + ll < a.L) {
            uint16_t* pz = &stg[ll][3 * (k & 3)];
            uint16_t* pn = pz + LINE_KEPT / 16;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                pz[i] = (uint16_t)(nz >> (16 * i));
                pn[i] = (uint16_t)(ng >> (16 * i));
            }
        }
    }
    __syncthreads();
    // exclusive scan of the 768 word sums by wave 0 (12 per lane)
    if (wave == 0) {
        int32_t loc[GROUPS / 64];
        int32_t t = 0;
#pragma unroll
        for (int i = 0; i < GROUPS / 64; ++i) { loc[i] = t; t += gsum[lane * (GROUPS / 64) + i]; }
        int32_t inc = t;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int32_t o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        const int32_t excl = inc - t;
#pragma unroll
        for (int i = 0; i < GROUPS / 64; ++i) gpre[lane * (GROUPS / 64) + i] = excl + loc[i];
        if (lane == 63) a.segsum[s - a.seg0] = inc;
    }
    __syncthreads();

    // 7. compressed lines, written whole by four lanes each: lane w < 3 the
    //    planes of kept word w, lane 3 the checkpoints M(before the word) of
    //    the three words, segment-local (the fix-up adds the segment offset):
    //    the line's first value starts word 0, so its checkpoint is the
    //    exclusive prefix of the line's first 64-value group, and each next
    //    one adds the word sum of the kept word before it.
    for (int t = tid; t < GROUPS; t += THREADS) {
        const int ll = t >> 2, part = t & 3;
        if (gl0 + ll < 0 || gl0 + ll >= a.L) continue;
        const u64* sw = reinterpret_cast<const u64*>(stg[ll]);   // nonzero words 0..2, negative words 0..2
        uint4 v;
        if (part < 3) {
            const u64 nz = sw[part], ng = sw[3 + part];
            v = make_uint4((uint32_t)nz, (uint32_t)(nz >> 32), (uint32_t)ng, (uint32_t)(ng >> 32));
        } else {
            const int32_t c0 = gpre[4 * ll];
            const int32_t c1 = c0 + (int)__popcll(sw[0]) - 2 * (int)__popcll(sw[3]);
            const int32_t c2 = c1 + (int)__popcll(sw[1]) - 2 * (int)__popcll(sw[4]);
            v = make_uint4((uint32_t)c0, (uint32_t)c1, (uint32_t)c2, 0u);
        }
        reinterpret_cast<uint4*>(a.cl + (gl0 + ll))[part] = v;
    }
    if (first <= (u64)a.ud) {
        for (int i = tid; i < SEGT / 4; i += THREADS) {
            const uint32_t x = acc32[i];
            const int b0 = (int8_t)x, b1 = (int8_t)(x >> 8), b2 = (int8_t)(x >> 16), b3 = (int8_t)(x >> 24);
            const int s4 = b0 + b1 + b2 + b3;
            // exclusive prefix over the word's 16 dwords (aligned groups of 16 lanes)
            int inc = s4;
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                const int o = __shfl_up(inc, d, 16);
                if ((lane & 15) >= d) inc += o;
            }
            int v = gpre[i >> 4] + inc - s4;
            const u64 n = first + 4 * (u64)i;
            const int bb[4] = {b0, b1, b2, b3};
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                v += bb[t];
                if (n + t <= (u64)a.ud) a.Md[n + t] = (int16_t)v;
            }
        }
    }

    // 8. packed mu for j coprime to 6 (t = j/3; 16 two's-complement 2-bit
    //    codes per word)
    if (s < a.mu_segs) {
        const i64 w0 = s * (SEGT / 48);
        for (int wi = tid; wi < SEGT / 48; wi += THREADS) {
            uint32_t word = 0;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int tl = wi * 16 + e;                // local t
                const int jl = 3 * tl + (tl & 1);          // j - base - 1
                const int m = (int8_t)acc[jl];
                word |= ((uint32_t)m & 3u) << (2 * e);
            }
            a.mu6[w0 + wi] = word;
        }
    }
}

// Adds the exclusive segment offsets to the segment-local values and sets
// M(0) = 0 (a whole table's fix-up; the sieve writes M(n) from n = 1, so
// M(0) is set here even when the dense tier holds nothing else).
// Grid-stride: a dispatch holds fewer than 2^32 work-items, the words can
// be more.
__global__ void k_table_fixup(FixupArgs a) {
    const i64 i0 = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 stride = (i64)gridDim.x * blockDim.x;
    if (i0 == 0 && !a.carry) a.Md[0] = 0;
    for (i64 i = i0; i < a.nd; i += stride) {
        const i64 n = i + 1;
        const i64 s = ((n - 1) >> a.seg_shift) / 3;
        a.Md[n] = (int16_t)(a.Md[n] + (int32_t)a.segoff[s - a.seg0]);
    }
    const int32_t carry = a.carry ? (int32_t)a.carry[0] : 0;
    for (i64 i = i0; i < a.L; i += stride) {   // line i: its three checkpoints
        const int32_t add = carry + (int32_t)a.segoff[((a.ud + (i64)LINE_VALUES * i) >> a.seg_shift) / 3 - a.seg0];
        int32_t* ck = a.cl[i].ck;
        ck[0] += add;
        ck[1] += add;
        ck[2] += add;
    }
}

hipError_t launch_sieve(int shape, unsigned blocks, hipStream_t st, const SieveArgs& a) {
    switch (shape) {
        case 0: hipLaunchKernelGGL((k_sieve_t<49152, 1024>), dim3(blocks), dim3(1024), 0, st, a); break;
        case 1: hipLaunchKernelGGL((k_sieve_t<24576, 256>), dim3(blocks), dim3(256), 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_sieve_t<98304, 1024>), dim3(blocks), dim3(1024), 0, st, a); break;
        case 3: hipLaunchKernelGGL((k_sieve_t<49152, 256>), dim3(blocks), dim3(256), 0, st, a); break;
        case 4: hipLaunchKernelGGL((k_sieve_t<49152, 512>), dim3(blocks), dim3(512), 0, st, a); break;
        case 5: hipLaunchKernelGGL((k_sieve_t<24576, 512>), dim3(blocks), dim3(512), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mdr
