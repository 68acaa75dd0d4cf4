// Prints the drop-in header's parameter helpers for the X values on the
// command line (host logic only; no GPU call is made).
#include "mertens_dr.hpp"

#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv) {
    using E = mertens_dr::DelégliseRivatEngine;
    for (int i = 1; i < argc; ++i) {
        const long long X = std::atoll(argv[i]);
        std::printf("%lld %.6f %lld %lld %d %d %lld %d\n", X, E::choose_cx(X), E::choose_sieve_limit(X, 0),
                    E::choose_compressed_limit(X), (int)E::use_compressed_default(X), (int)E::use_chmt_env(),
                    mertens_dr::isqrt(X), mertens_dr::hip_detail::opts_from_env().params);
    }
    return 0;
}
