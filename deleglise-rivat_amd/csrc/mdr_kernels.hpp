// Kernel argument blocks and kernel declarations shared by the .hip units.
#ifndef MDR_KERNELS_HPP
#define MDR_KERNELS_HPP

#include <hip/hip_runtime.h>
#include "mdr_common.hpp"

namespace mdr {

struct SieveArgs {
    i64 U, ud, L;
    i64 seg_first;
    i64 mu_segs;               // segments whose packed mu is written
    const uint8_t* stencil;    // STENCIL_PERIOD + SEG bytes
    const uint8_t* pattern;    // PATTERN_REPEAT * PATTERN_PERIOD + SEG_MAX bytes (weights of 13..23)
    const uint8_t* pattern3;   // PATTERN_REPEAT * PATTERN3_PERIOD + SEG_MAX bytes (weights of 29, 31, 37)
    const uint32_t* wprimes;   // sieving primes >= 13, ascending
    const double* wrcp;        // 1/p
    const uint8_t* wweights;   // bit_width(p-1)|1
    int n_w, n_small;          // total, and those below 256
    const PrimeConsts* wlarge; // wlarge[i - n_small]: constants of prime i >= 256
    int n_split;               // first index with p >= SQ_SPLIT (n_small <= n_split <= n_w)
    int n_coop;                // first index with p >= COOP_PRIME_LIMIT (n_small <= n_coop <= n_split)
    const uint32_t* sqprimes;  // primes >= 7 whose squares are flagged
    int n_sq, n_sq_small;      // total, and those below 256 (larger ones go with the weights)
    int16_t* Md;
    CLine* cl;
    uint32_t* mu6;
    i64* segsum;               // segsum[s - seg0]: segment s's sum
    i64 seg0;
};

struct FixupArgs {
    i64 ud, L;                 // lines hold the values above ud
    i64 nd;                    // dense values 1..nd to fix (0 for a window of lines)
    int seg_shift;             // segment length 3 * 2^seg_shift
    int16_t* Md;
    CLine* cl;
    const i64* segoff;         // segoff[s - seg0]: exclusive offset of segment s
    i64 seg0;
    const i64* carry;          // nullptr, or M(ud) added to every checkpoint
};

// Fixed per-slot cost (cycles) charged by the learned partition for a G
// evaluation's set-up (frame, segment geometry, constant).
constexpr long long SLOT_OVERHEAD_CYCLES = 1500;
// The learned partition cuts chunks only for calls whose largest summed X is
// at least this (the low end of the cost model's training range, 512 X in
// [1e14, 1e16]): below it the model extrapolates, and its partition phase
// (the network per cost piece, a second scan) costs about 0.35 ms against a
// whole call of 0.4-3 ms; those calls take the static partition.
constexpr long long LEARNED_MIN_X = 100000000000000LL;
// Smallest automatic chunk (quotient positions): a chunk's set-up (locating its
// slots, building their frames) costs about as much as a few thousand positions.
constexpr long long MIN_CHUNK_POSITIONS = 1LL << 15;

// Partition of the quotient phase.
struct PartArgs {
    const XParam* xp;
    int nx;
    i64 total_slots;
    TableParam tp;
    const float* mlp;              // learned cost model (nullptr = static partition)
    i64* len;                      // per slot positions; position offsets after the scan
    i64* cost;                     // per slot predicted cycles; cost offsets after the scan
    unsigned long long* type_pos;  // [3]: S1 dense, S1 compressed, S2 positions
    const i64* totals;             // [0] total positions, [1] total predicted cycles
    i64 P;                         // chunks (bounds has P + 1 entries)
    i64* bounds;                   // [P + 1]
    i64 min_chunk;                 // > 0: use only max(1, min(P, T / min_chunk)) chunks,
                                   // the rest stay empty (small calls, little set-up)
    int by_slot;                   // 1: cut at slot boundaries, equal slot counts per chunk
};

struct QuotArgs {
    const XParam* xp;
    int nx;
    TableParam tp;
    const i64* off;                // per slot first global position
    i64 nslots;
    const i64* bounds;             // [P + 1]
    i64 P;
    unsigned long long* next;      // dynamic work counter
    u64* out128;                   // 2 u64 per X
    int dynamic;                   // 1: atomic chunk queue, 0: static round robin
    StampRec* stamps;              // optional cycle log
    unsigned long long* nstamps;
    i64 stamp_cap;
    int stamp_stride;              // log slots whose hash is 0 mod stride
    u64* wave_span;                // optional: [2 w] first and [2 w + 1] last wall clock of wave w
};

hipError_t launch_sieve(int shape, unsigned blocks, hipStream_t st, const SieveArgs& a);
__global__ void k_table_fixup(FixupArgs a);
__global__ void k_slot_len(PartArgs a);
__global__ void k_cuts(PartArgs a);
__global__ void k_quot(QuotArgs a);

// Exclusive scan of an i64 array (in place allowed), three kernel launches
// for every n (n = 0 included), so that captured graphs keep one topology.
// The total lands in d_tmp[max(1, (n + 2047) / 2048)] and, when dev_total is
// not null, in *dev_total (device memory).
hipError_t scan_exclusive_i64(const i64* in, i64* out, i64 n, i64* d_tmp, i64 tmp_len, i64* dev_total,
                              hipStream_t st);

// Sets n0 words at p0, n1 at p1 and n2 at p2 to zero (a kernel, not a memset
// node, so that it replays correctly inside captured graphs).
__global__ void k_clear3(u64* p0, i64 n0, u64* p1, i64 n1, u64* p2, i64 n2);
i64 scan_tmp_len(i64 n);

// Cost-model network (mdr_mlp.hip).
hipError_t mlp_predict(const float* w, const float* X, i64 n, float* out, hipStream_t st);
hipError_t mlp_grad(const float* w, const float* X, const float* T, const int* idx, i64 n, float* buf, float* err,
                    float* grad, float* loss, hipStream_t st);
// one Adam step over the batch idx[0..n): forward/backward, then gradient and
// update of every parameter in one launch
hipError_t mlp_train_step(float* w, const float* X, const float* T, const int* idx, i64 n, float* buf, float* m,
                          float* v, float lr, int step, hipStream_t st);
i64 mlp_row_floats();

}  // namespace mdr

#endif
