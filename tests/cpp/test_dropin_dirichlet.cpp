// The reference's cross-validation of its Deleglise-Rivat engine against the
// Dirichlet DP engine (test_dr.cpp:71-88), written as the reference writes it,
// on this repository's drop-in headers.  Exit status 0 when every value matches.
#include "dirichlet_engine.hpp"
#include "mertens_dr.hpp"

#include <cstdio>
#include <vector>

int main() {
    using namespace mertens_dr;
    const std::vector<int64> cross_checks = {54321LL,      987654LL,        12345678LL,     987654321LL,
                                             5000000000LL, 123456789012LL, 1000000000000LL};
    int failed = 0;
    for (int64 x : cross_checks) {
        const int64 dr_val = DelégliseRivatEngine::compute_mertens(x);
        const int64 dp_val = dirichlet::DirichletEngine::compute_mertens(x);
        std::printf("%s M(%lld) DR %lld, DP %lld\n", dr_val == dp_val ? "[MATCH]" : "[MISMATCH]", x, dr_val, dp_val);
        failed += dr_val != dp_val;
    }
    std::printf("%d failed\n", failed);
    return failed ? 1 : 0;
}
