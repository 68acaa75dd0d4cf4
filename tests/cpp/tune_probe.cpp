// Prints the tuned choices the engine's tuning-cache reader makes for the X
// values on the command line: decade-tuned sieve workgroup shape, chunk count,
// cutoff, dense tier and cx.  Host logic only (mdr_tune.cpp compiled on its
// own, no GPU code); the cache is the file named by MDR_TUNE_FILE.
#include "mdr_tune.hpp"

#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        const long long X = std::atoll(argv[i]);
        const mdr::TuneChoice t = mdr::tune_lookup(X, true);
        std::printf("%lld %d %lld %lld %lld %.6f\n", X, mdr::tune_sieve_shape(X), mdr::tune_chunks(X), t.u, t.ud,
                    t.cx);
    }
    return 0;
}
