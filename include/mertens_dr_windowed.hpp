// Drop-in replacement for the reference's mertens_dr_windowed.hpp on the
// MI355X engine: the same namespace and entry point,
//   mertens_windowed::compute_mertens_windowed(int64 X, int num_threads = 0,
//                                              int64 block_M = 1 << 19,
//                                              bool verbose = false)
// (mertens_dr_windowed.hpp:49), returning the exact M(X).
//
// The reference keeps mu and M up to mu_limit = max(cx sqrt(X) + 1000,
// X/u + 1000) resident and streams the M table above it in cache-sized
// windows of block_M values with the dense-strategy cutoff
// (choose_sieve_limit).  Here the resident part is the GPU table up to that
// limit and the compressed tier above it is sieved and summed window by window
// (mdr_opts.window).  A GPU window pass repartitions the whole quotient phase,
// so block_M is raised to at least 2^30 values (a few hundred MB of lines);
// the result does not depend on the window size.  num_threads is accepted and
// ignored; verbose prints the phase times to stderr as the reference's
// "[win]" line does (:388-394).
#ifndef MERTENS_DR_WINDOWED_HPP
#define MERTENS_DR_WINDOWED_HPP

#include <algorithm>
#include <iostream>

#include "mertens_dr.hpp"

namespace mertens_windowed {
using mertens_dr::int64;

// Smallest window the GPU path uses (values).
constexpr int64 MIN_GPU_WINDOW = int64(1) << 30;

static inline int64 compute_mertens_windowed(int64 X, int num_threads = 0, int64 block_M = (1 << 19),
                                             bool verbose = false) {
    (void)num_threads;
    if (X < 1) return 0;
    const mertens_dr::hip_detail::Lease lease;
    mdr_ctx* c = lease.get();
    mdr_opts o = mertens_dr::hip_detail::opts_from_env(MDR_STRATEGY_DENSE);
    // the windowed engine keeps the reference's dense-strategy cutoff unless
    // MERTENS_HIP_TUNED=1 asks for the tuned one
    if (!std::getenv("MERTENS_HIP_TUNED")) o.params = MDR_PARAMS_REFERENCE;
    o.window = std::max(block_M, MIN_GPU_WINDOW);
    mdr_stats st;
    long long out = 0;
    mertens_dr::hip_detail::check(mdr_mertens(c, X, &o, &out, &st), c, "compute_mertens_windowed");
    if (verbose)
        std::cerr << "[win] table=" << st.table_ms / 1000.0 << "s partition=" << st.partition_ms / 1000.0
                  << "s quotient=" << st.quotient_ms / 1000.0 << "s windows=" << st.windows << " u=" << st.u
                  << " mu_limit=" << st.mu_limit << std::endl;
    return out;
}

}  // namespace mertens_windowed

#endif
