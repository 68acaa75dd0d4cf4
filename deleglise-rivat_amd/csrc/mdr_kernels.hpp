// Kernel argument blocks and kernel declarations shared by the .hip units.
#ifndef MDR_KERNELS_HPP
#define MDR_KERNELS_HPP

#include <hip/hip_runtime.h>
#include "mdr_common.hpp"

namespace mdr {

struct SieveArgs {
    i64 U, ud, W;
    i64 seg_first;
    i64 mu_segs;               // segments whose packed mu is written
    const uint8_t* stencil;    // STENCIL_PERIOD + SEG bytes
    const uint32_t* wprimes;   // sieving primes >= 13, ascending
    const double* wrcp;        // 1/p
    const uint8_t* wweights;   // bit_width(p-1)|1
    int n_w, n_small;          // total, and those below 256
    const uint32_t* sqprimes;  // primes >= 7 whose squares are flagged
    int n_sq, n_sq_small;      // total, and those below 256 (larger ones go with the weights)
    int16_t* Md;
    u64* cpl;
    int32_t* cck;
    uint32_t* mu6;
    i64* segsum;
};

struct FixupArgs {
    i64 ud, W;
    int16_t* Md;
    int32_t* cck;
    const i64* segoff;
};

// Partition of the quotient phase.
struct PartArgs {
    const XParam* xp;
    int nx;
    i64 total_slots;
    TableParam tp;
    i64 chunk;                 // static chunk length in positions
    const float* cost_rate;    // learned: predicted cycles per position per segment class (nullptr = static)
    double cost_target;        // learned: target cycles per chunk
    i64* counts;               // per slot chunk count (k_count output)
    unsigned long long* type_pos;  // [3]: S1 dense, S1 compressed, S2 positions
    const i64* offsets;        // per slot first record (k_expand input)
    i64 nrec;                  // total records
    ChunkRec* recs;
};

struct QuotArgs {
    const XParam* xp;
    TableParam tp;
    const ChunkRec* recs;
    i64 nrec;
    unsigned long long* next;  // dynamic work counter
    u64* out128;               // 2 u64 per X
    uint32_t* stamps;          // optional per-record cycle counts
    int dynamic;               // 1: atomic work queue, 0: static round robin
};

__global__ void k_sieve(SieveArgs a);
__global__ void k_table_fixup(FixupArgs a);
__global__ void k_count(PartArgs a);
__global__ void k_expand(PartArgs a);
__global__ void k_quot(QuotArgs a);

// Exclusive scan of an i64 array (in place allowed); returns total via *total.
hipError_t scan_exclusive_i64(const i64* in, i64* out, i64 n, i64* d_tmp, i64 tmp_len, i64* total,
                              hipStream_t st);
i64 scan_tmp_len(i64 n);

}  // namespace mdr

#endif
