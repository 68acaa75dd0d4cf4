// Tuned GPU parameters: the cx/fx split and the chunk length.  Values come
// from the autotuner's cache when present, else from the built-in defaults.
#ifndef MDR_TUNE_HPP
#define MDR_TUNE_HPP

#include <stddef.h>

#include "mdr_common.hpp"

namespace mdr {

struct TuneChoice {
    i64 u;      // sieve cutoff
    i64 ud;     // dense tier limit (multiple of 64)
    double cx;  // S1/S2 split factor
    i64 window = 0;   // > 0: the cutoff exceeds what one table may hold; stream the
                      // compressed tier through windows of this many values
};

TuneChoice tune_lookup(i64 X, bool streaming);
i64 tune_chunk();
i64 tune_chunks(i64 X);   // tuned quotient-phase chunk count for X (0 = automatic)
// shared cutoff of a batch whose X values sum to xsum (factor 0 = tuned factor)
i64 tune_batch_u(double xsum, double factor = 0);
int tune_sieve_shape(i64 X);   // tuned sieve workgroup shape for X (index into SIEVE_SHAPES)
// Lowers the largest tuned cutoff to what a device of this memory holds.
void tune_set_device_memory(size_t bytes);
i64 tune_u_cap();   // largest one-table cutoff for this device

}  // namespace mdr

#endif
