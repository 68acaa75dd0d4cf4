// Drop-in check of the windowed entry point: the reference's known-value list
// (test_dr.cpp:26-47) through mertens_windowed::compute_mertens_windowed, as
// a caller of mertens_dr_windowed.hpp would write it, with the default and
// with explicit window sizes.  Exit status 0 when every value matches.
#include "mertens_dr_windowed.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char** argv) {
    using mertens_windowed::int64;
    const long long max_x = argc > 1 ? std::atoll(argv[1]) : 1000000000000000LL;
    struct Case { int64 X, M; };
    const std::vector<Case> cases = {
        {1, 1}, {2, 0}, {3, -1}, {10, -1}, {100000, -48}, {1000000, 212}, {100000000, 1928},
        {1000000000, -222}, {10000000000LL, -33722}, {100000000000LL, -87856}, {1000000000000LL, 62366},
        {10000000000000LL, 599582}, {100000000000000LL, -875575}, {1000000000000000LL, -3216373},
        {10000000000000000LL, -3195437}};
    int failed = 0;
    for (const Case& c : cases) {
        if (c.X > max_x) continue;
        for (int64 block : {int64(1) << 19, int64(3) << 30}) {
            const int64 got = mertens_windowed::compute_mertens_windowed(c.X, 0, block, c.X == max_x);
            std::printf("%s M(%lld) = %lld (expected %lld), block_M %lld\n", got == c.M ? "[PASS]" : "[FAIL]",
                        c.X, got, c.M, block);
            failed += got != c.M;
        }
    }
    std::printf("%d failed\n", failed);
    return failed ? 1 : 0;
}
