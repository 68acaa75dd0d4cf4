// Drop-in check of the table objects, written as a reference caller would
// use them (the checks of test_sieve.cpp:79-96 and :165-182, and its
// reduced-team block :189-204 as concurrent host threads): SieveTable and
// CompressedMertensTable from mertens_dr.hpp, now built by the MI355X sieve,
// against mu and its prefix sums from a plain sieve of Eratosthenes written
// here.  The compressed tables' chunk records are also checked bit by bit
// against that mu: the layout is the reference's (BitChunk + checkpoints).
// Exit status 0 when every check passes.
#include "mertens_dr.hpp"

#include <cstdio>
#include <memory>
#include <thread>
#include <vector>

namespace {

int failures = 0;

void check(bool ok, const char* what) {
    if (!ok && failures++ < 20) std::printf("[FAIL] %s\n", what);
}

// mu(n) and M(n) for n <= limit: every prime's multiples flip the sign, every
// square of a prime zeroes.
struct Truth {
    std::vector<int8_t> mu;
    std::vector<int32_t> prefix;
    explicit Truth(int limit) : mu(limit + 1, 1), prefix(limit + 1, 0) {
        mu[0] = 0;
        std::vector<bool> composite(limit + 1, false);
        for (long long p = 2; p <= limit; ++p) {
            if (composite[p]) continue;
            for (long long m = p; m <= limit; m += p) {
                if (m > p) composite[m] = true;
                mu[m] = static_cast<int8_t>(-mu[m]);
            }
            for (long long m = p * p; m <= limit; m += p * p) mu[m] = 0;
        }
        for (int n = 1; n <= limit; ++n) prefix[n] = prefix[n - 1] + mu[n];
    }
};

void check_chunks(const mertens_dr::CompressedMertensTable& t, const Truth& ref) {
    for (long long c = 0; c < t.num_chunks; ++c) {
        for (int k = 0; k < 4; ++k) {
            const long long n0 = t.u_dense + 1 + 256 * c + 64 * k;
            if (n0 <= t.u_total) check(t.checkpoints[4 * c + k] == ref.prefix[n0 - 1], "chunk checkpoint");
            for (int i = 0; i < 64; ++i) {
                const long long n = n0 + i;
                const int m = n <= t.u_total ? ref.mu[n] : 0;
                const int bit = 64 * k + i;
                const bool nz = (t.chunks[c].nonzero[bit / 8] >> (bit % 8)) & 1;
                const bool ng = (t.chunks[c].sign[bit / 8] >> (bit % 8)) & 1;
                check(nz == (m != 0) && ng == (m < 0), "chunk plane bit");
            }
        }
    }
}

}  // namespace

int main() {
    using namespace mertens_dr;
    const Truth ref(1048576 + 64);

    // test_sieve.cpp:79-96: streamed tables over the reference's size lists
    int built = 0;
    for (int total : {0, 1, 2, 3, 63, 64, 65, 127, 128, 129, 255, 256, 257, 13859, 13860, 13861, 17324, 17325,
                      17326, 131071, 131072, 131073, 262147, 1000033}) {
        for (int dense : {0, 1, 7, 63, 64, 65, 255, 256, 257, 1000, 131071, total}) {
            if (dense > total) continue;
            CompressedMertensTable table(total, dense, 3, total + 17);
            ++built;
            for (int n = 0; n <= total; ++n) {
                check(table.get_M(n) == ref.prefix[n], "streamed prefix");
                check(table.get_mu(n) == ref.mu[n], "streamed mu");
            }
            check(table.mu_limit == total, "streamed mu limit clamped to the total");
            check(table.get_mu(total + 1) == 0, "streamed mu beyond its limit");
            if (total <= 262147) check_chunks(table, ref);
        }
    }
    std::printf("%s %d streamed tables\n", failures ? "[FAIL]" : "[PASS]", built);

    // test_sieve.cpp:165-182
    for (int limit : {0, 1, 2, 3, 4, 7, 31, 256, 257, 262145, 1000033}) {
        SieveTable table(limit, 3);
        for (int n = 0; n <= limit; ++n) {
            check(table.get_mu(n) == ref.mu[n], "sieve table mu");
            check(table.get_M(n) == ref.prefix[n], "sieve table prefix");
            check(table(n) == ref.prefix[n] && table.data()[n] == ref.prefix[n], "sieve table operator() / data()");
        }
    }
    SieveTable partial(262145, 3, 20000, true, 12345);
    check(partial.mu_limit == 20000 && partial.m_limit == 12345, "partial table limits");
    check(partial.mu_odd_size() == (262145 + 1) / 2, "retained odd table size");
    for (int n = 0; n <= 12345; ++n) check(partial.get_M(n) == ref.prefix[n], "partial prefix");
    for (int n = 0; n <= 20000; ++n) check(partial.get_mu(n) == ref.mu[n], "partial mu");
    check(partial.get_mu(20001) == 0, "partial mu beyond its limit");
    for (int n = 1; n <= 262145; ++n) check(mu_from_odd(n, partial.mu_odd_data()) == ref.mu[n], "retained odd table");
    CompressedMertensTable compressed(1048576, 1024, 3, 1048576);
    for (int n = 0; n <= 1048576; ++n) {
        check(compressed.get_M(n) == ref.prefix[n], "compressed prefix");
        check(compressed.get_mu(n) == ref.mu[n], "compressed mu");
    }
    check_chunks(compressed, ref);
    std::printf("%s sieve, partial, retained-odd and compressed tables\n", failures ? "[FAIL]" : "[PASS]");

    // test_sieve.cpp:189-204 runs two table builds in a reduced OpenMP team;
    // here two host threads build and read tables at once (each call leases
    // its own engine context)
    int thread_failures[2] = {0, 0};
    std::vector<std::thread> ts;
    for (int w = 0; w < 2; ++w)
        ts.emplace_back([&, w] {
            SieveTable table(262145, 8);
            CompressedMertensTable comp(1048576, 1024, 8, 1024);
            for (int n = 0; n <= 262145; ++n) {
                if (table.get_M(n) != ref.prefix[n]) ++thread_failures[w];
                if (comp.get_M(n) != ref.prefix[n]) ++thread_failures[w];
            }
            if (DelégliseRivatEngine::compute_mertens(987654321LL, 8) != 1697) ++thread_failures[w];
        });
    for (auto& t : ts) t.join();
    check(thread_failures[0] == 0 && thread_failures[1] == 0, "concurrent table builds");
    std::printf("%s concurrent threads\n", (thread_failures[0] + thread_failures[1]) ? "[FAIL]" : "[PASS]");

    // a table kept on the GPU: gathers against the same truth
    {
        const hip::DeviceTable d(1048576, 4096, 1048576);
        std::vector<int64> qs;
        for (int64 q = 0; q <= 1048576; q += 997) qs.push_back(q);
        const std::vector<int32_t> got = d.get_M(qs);
        for (size_t i = 0; i < qs.size(); ++i) check(got[i] == ref.prefix[qs[i]], "device table gather");
        check(d.get_mu(30030) == ref.mu[30030] && d.get_mu(1048577) == 0, "device table mu");
        check(d.limit() == 1048576 && d.mu_limit() == 1048576 && d.dense_limit() <= 4096, "device table info");
    }
    // resident tables take no pooled context: more of them than the pool's
    // four contexts, held while the same thread computes M(X)
    {
        std::vector<std::unique_ptr<hip::DeviceTable>> held;
        for (int k = 0; k < 6; ++k) held.push_back(std::make_unique<hip::DeviceTable>(100000 + 50000 * k, 256, 1000));
        check(DelégliseRivatEngine::compute_mertens(987654321LL) == 1697, "compute beside resident tables");
        for (int k = 0; k < 6; ++k) {
            const int64 q = 99999 + 50000 * k;
            check(held[k]->get_M(q) == ref.prefix[q], "resident table after the call");
        }
    }
    std::printf("%d failed\n", failures);
    return failures ? 1 : 0;
}
