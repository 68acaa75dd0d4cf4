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

constexpr int GROUPS = SEG / 64;          // 768 groups of 64 values

__device__ __forceinline__ uint32_t mod_u64(u64 x, uint32_t p, double rp) {
    // x < 2^52; estimate within 1 of floor(x/p), then fix.
    i64 q = (i64)((double)x * rp);
    i64 r = (i64)x - q * (i64)p;
    if (r < 0) r += p;
    else if (r >= (i64)p) r -= p;
    return (uint32_t)r;
}

__device__ __forceinline__ int8_t finalize(uint8_t a, u64 n) {
    if (a & 0x80) return 0;
    const unsigned thr = 63u - (unsigned)__clzll(n);
    const int s = (a & 1) ? -1 : 1;
    return (int8_t)(a > thr ? s : -s);
}

}  // namespace

__global__ __launch_bounds__(SIEVE_THREADS) void k_sieve(SieveArgs a) {
    __shared__ uint32_t acc32[SEG / 4];
    __shared__ int32_t gsum[GROUPS];
    __shared__ int32_t gpre[GROUPS];
    __shared__ uint32_t small_off[64];
    __shared__ uint32_t sq_off[64];
    uint8_t* acc = reinterpret_cast<uint8_t*>(acc32);

    const i64 s = (i64)blockIdx.x + a.seg_first;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const u64 base = (u64)s * SEG;
    const u64 lastn = (u64)(base + SEG <= (u64)a.U ? base + SEG : (u64)a.U);
    const int len = (int)(lastn - base);

    // 1. periodic stencil: weights of 2,3,5,7,11 and flags of 4, 9, 25.
    {
        const uint32_t phase = (uint32_t)(base % STENCIL_PERIOD);
        const uint4* src = reinterpret_cast<const uint4*>(a.stencil + phase);
        uint4* dst = reinterpret_cast<uint4*>(acc32);
        for (int i = tid; i < SEG / 16; i += SIEVE_THREADS) dst[i] = src[i];
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

    // 2. square flags for 7 <= p < 256, block-cooperative over the multiples
    for (int i = 0; i < a.n_sq_small; ++i) {
        const uint32_t p = a.sqprimes[i];
        const uint32_t p2 = p * p;
        if ((u64)p2 > lastn) break;
        for (uint32_t x = sq_off[i] + (uint32_t)tid * p2; x < (uint32_t)len; x += SIEVE_THREADS * p2)
            atomicOr(&acc32[x >> 2], 0x80u << (8 * (x & 3)));
    }
    // 3. weights for 13 <= p < 256, block-cooperative over the multiples
    for (int i = 0; i < a.n_small; ++i) {
        const uint32_t p = a.wprimes[i];
        const uint32_t w = a.wweights[i];
        for (uint32_t x = small_off[i] + (uint32_t)tid * p; x < (uint32_t)len; x += SIEVE_THREADS * p)
            atomicAdd(&acc32[x >> 2], w << (8 * (x & 3)));
    }
    // 4. primes 256 <= p <= sqrt(last n), one per thread, four in flight:
    //    r2 = first mod p^2 gives the (at most one) square hit, r2 mod p the
    //    first weight hit.  Primes ascend, so a thread stops at the first
    //    p with p^2 beyond the segment.
    constexpr int UNR = 4;
    for (int i0 = a.n_small + tid; i0 < a.n_w; i0 += UNR * SIEVE_THREADS) {
        uint32_t pp[UNR];
        double rr[UNR];
        uint32_t ww[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int i = i0 + u * SIEVE_THREADS;
            const bool ok = i < a.n_w;
            pp[u] = ok ? a.wprimes[i] : 0xffffffffu;
            rr[u] = ok ? a.wrcp[i] : 0.0;
            ww[u] = ok ? a.wweights[i] : 0u;
        }
        if ((u64)pp[0] * pp[0] > lastn) break;
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const u64 p = pp[u];
            const u64 p2 = p * p;
            if (p2 > lastn) continue;
            const double rp = rr[u];
            i64 q2 = (i64)((double)first * (rp * rp));
            i64 r2 = (i64)first - q2 * (i64)p2;
            while (r2 < 0) r2 += (i64)p2;
            while (r2 >= (i64)p2) r2 -= (i64)p2;
            const u64 xs = r2 ? p2 - (u64)r2 : 0;
            if (xs < (u64)len) atomicOr(&acc32[xs >> 2], 0x80u << (8 * (xs & 3)));
            const i64 q1 = (i64)((double)r2 * rp);
            i64 r = r2 - q1 * (i64)p;
            if (r < 0) r += (i64)p;
            else if (r >= (i64)p) r -= (i64)p;
            const uint32_t w = ww[u];
            for (uint32_t x = r ? (uint32_t)(p - r) : 0; x < (uint32_t)len; x += (uint32_t)p)
                atomicAdd(&acc32[x >> 2], w << (8 * (x & 3)));
        }
    }
    __syncthreads();

    // 5. finalise bytes to mu in {-1, 0, 1}
    for (int i = tid; i < SEG / 4; i += SIEVE_THREADS) {
        uint32_t v = acc32[i];
        uint32_t out = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const u64 n = first + 4 * (u64)i + b;
            int8_t m = 0;
            if (n <= (u64)a.U) {
                m = finalize((uint8_t)(v >> (8 * b)), n);
                if (n == 1) m = 1;
                if (n == 2) m = -1;
            }
            out |= (uint32_t)(uint8_t)m << (8 * b);
        }
        acc32[i] = out;
    }
    __syncthreads();

    // 6. group sums and compressed planes (ballots over 64 values)
    for (int g = wave; g < GROUPS; g += SIEVE_THREADS / 64) {
        const int8_t v = (int8_t)acc[g * 64 + lane];
        const u64 nz = __ballot(v != 0);
        const u64 ng = __ballot(v < 0);
        if (lane == 0) gsum[g] = (int)__popcll(nz) - 2 * (int)__popcll(ng);
        const u64 n0 = first + (u64)g * 64;
        if (n0 > (u64)a.ud && n0 <= (u64)a.U && lane < 2) {
            const i64 w = (i64)(n0 - (u64)a.ud - 1) >> 6;
            a.cpl[2 * w + lane] = lane ? ng : nz;
        }
    }
    __syncthreads();
    // exclusive scan of the 768 group sums by wave 0 (12 per lane)
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
        if (lane == 63) a.segsum[s] = inc;
    }
    __syncthreads();

    // 7. dense prefix values and compressed checkpoints (segment-local)
    for (int g = wave; g < GROUPS; g += SIEVE_THREADS / 64) {
        const u64 n0 = first + (u64)g * 64;
        if (n0 > (u64)a.U) break;
        if (n0 <= (u64)a.ud) {
            int32_t v = (int8_t)acc[g * 64 + lane];
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int32_t o = __shfl_up(v, d, 64);
                if (lane >= d) v += o;
            }
            const u64 n = n0 + lane;
            if (n <= (u64)a.ud) a.Md[n] = (int16_t)(gpre[g] + v);
        } else if (lane == 0) {
            const i64 w = (i64)(n0 - (u64)a.ud - 1) >> 6;
            a.cck[w] = gpre[g];
        }
    }

    // 8. packed mu for j coprime to 6 (t = j/3; 16 codes per word)
    if (s < a.mu_segs) {
        const i64 w0 = s * (SEG / 48);
        for (int wi = tid; wi < SEG / 48; wi += SIEVE_THREADS) {
            uint32_t word = 0;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int tl = wi * 16 + e;                // local t
                const int jl = 3 * tl + (tl & 1);          // j - base - 1
                const int m = (int8_t)acc[jl];
                word |= (uint32_t)(m + 1) << (2 * e);
            }
            a.mu6[w0 + wi] = word;
        }
    }
}

// Adds the exclusive segment offsets to the segment-local values.
__global__ void k_table_fixup(FixupArgs a) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.ud) {
        const i64 n = i + 1;
        const i64 s = (n - 1) / SEG;
        a.Md[n] = (int16_t)(a.Md[n] + (int32_t)a.segoff[s]);
    }
    if (i < a.W) {
        const i64 s = (a.ud + 64 * i) / SEG;
        a.cck[i] = a.cck[i] + (int32_t)a.segoff[s];
    }
}

}  // namespace mdr
