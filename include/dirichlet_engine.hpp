// Drop-in for the optional dirichlet_engine.hpp the reference's tests and
// benchmark include when present (bench_dr.cpp:9-12, test_dr.cpp:8-11):
//   dirichlet::DirichletEngine::compute_mertens(int64 x, int threads = 0)
// (called at bench_dr.cpp:98 and test_dr.cpp:80).  The reference does not
// ship the engine; this one computes the exact M(x) on the GPU from
// sum_{d <= v} M(floor(v/d)) = 1 over the quotients v = floor(x/k) with a table
// to about x^(2/3) (mdr_mertens_dirichlet), independently of the
// Deleglise-Rivat split, so the reference's cross-validation (DR against
// DP, test_dr.cpp:71-88) runs as written.  threads is accepted and ignored.
#ifndef DIRICHLET_ENGINE_HPP
#define DIRICHLET_ENGINE_HPP

#include "mertens_dr.hpp"

namespace dirichlet {
using mertens_dr::int64;

class DirichletEngine {
public:
    static int64 compute_mertens(int64 x, int threads = 0) {
        (void)threads;
        if (x < 1) return 0;
        const mertens_dr::hip_detail::Lease lease;
        mdr_ctx* c = lease.get();
        long long out = 0;
        mertens_dr::hip_detail::check(mdr_mertens_dirichlet(c, x, &out, nullptr), c, "DirichletEngine::compute_mertens");
        return out;
    }
};

}  // namespace dirichlet

#endif
