// Drop-in check: the reference's own known-value list (test_dr.cpp:26-47,
// test_cuda.cu:38-60) through mertens_dr::DelégliseRivatEngine, exactly as a
// reference caller would write it, now backed by the MI355X engine.
// Exit status 0 when every value matches.
#include "mertens_dr.hpp"

#include <cstdio>
#include <vector>

int main(int argc, char** argv) {
    using namespace mertens_dr;
    const long long max_x = argc > 1 ? std::atoll(argv[1]) : 10000000000000000LL;
    struct Case { int64 X, M; };
    const std::vector<Case> cases = {
        {1, 1}, {2, 0}, {3, -1}, {4, -1}, {5, -2}, {10, -1}, {100, 1}, {1000, 2}, {10000, -23},
        {100000, -48}, {1000000, 212}, {10000000, 1037}, {100000000, 1928}, {1000000000, -222},
        {10000000000LL, -33722}, {100000000000LL, -87856}, {1000000000000LL, 62366},
        {10000000000000LL, 599582}, {100000000000000LL, -875575}, {1000000000000000LL, -3216373},
        {10000000000000000LL, -3195437}};
    int failed = 0;
    for (const Case& c : cases) {
        if (c.X > max_x) continue;
        mdr_stats st{};
        const int64 got = DelégliseRivatEngine::compute_mertens_stats(c.X, 0, &st);
        std::printf("%s M(%lld) = %lld (expected %lld)  total %.4f s, kernel %.4f s\n",
                    got == c.M ? "[PASS]" : "[FAIL]", c.X, got, c.M, st.total_time, st.kernel_time);
        failed += got != c.M;
    }
    // compute_compressed boundary cases of test_sieve.cpp:205-214
    const int64 b1 = DelegliseRivatEngine::compute_compressed(0, 0);
    const int64 b2 = DelegliseRivatEngine::compute_compressed(100000000LL, 3);
    const int64 b3 = DelegliseRivatEngine::compute_compressed(10000000000LL, 3);
    std::printf("%s compute_compressed boundary cases: %lld %lld %lld\n",
                (b1 == 0 && b2 == 1928 && b3 == -33722) ? "[PASS]" : "[FAIL]", b1, b2, b3);
    failed += !(b1 == 0 && b2 == 1928 && b3 == -33722);
    // ASCII alias and the parameter helpers compile and agree with the reference rules
    std::printf("choose_cx(1e17) = %.2f, choose_compressed_limit(1e17) = %lld\n",
                DelegliseRivatEngine::choose_cx(100000000000000000LL),
                DelegliseRivatEngine::choose_compressed_limit(100000000000000000LL));
    std::printf("%d failed\n", failed);
    return failed ? 1 : 0;
}
