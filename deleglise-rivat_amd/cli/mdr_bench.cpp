// Command-line front end of the MI355X engine with the reference CLIs'
// modes and report formats (bench_dr.cpp:14-135, bench_cuda.cu:40-122):
//
//   mdr_bench <X> [threads]                      M(X) with the timing breakdown
//   mdr_bench --benchmark [max_exp] [threads] [repeats]
//                                                sweep 10^1..10^max_exp (median time)
//   mdr_bench --compare [max_exp] [threads]      Dirichlet DP engine vs Deleglise-Rivat,
//                                                10^10..10^max_exp (bench_dr.cpp:76-122)
//   mdr_bench --compare-split [max_exp]          reference parameter rules vs the tuned
//                                                GPU split (warm: median of 3 after a warm-up)
//
// It calls the drop-in header, so the reference's environment controls
// (MERTENS_USE_CHMT, MERTENS_CX, MERTENS_CHMT_U_CAP, MERTENS_FX,
// MERTENS_U_CAP) and MERTENS_HIP_TUNED / MERTENS_HIP_LEARNED apply.
// --compare times the Dirichlet cross-check engine (dirichlet_engine.hpp, on
// the GPU here) against the Deleglise-Rivat engine and stops if they differ,
// as the reference's assert does.  threads is accepted and ignored.
#include "dirichlet_engine.hpp"
#include "mertens_dr.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

namespace {

void usage(const char* prog) {
    std::cout << "Usage:\n"
              << "  " << prog << " <X> [threads]                                M(X) at a single value\n"
              << "  " << prog << " --benchmark [max_exp] [threads] [repeats]   sweep (median time)\n"
              << "  " << prog << " --compare [max_exp] [threads]               Dirichlet DP vs Deleglise-Rivat\n"
              << "  " << prog << " --compare-split [max_exp]                   reference vs tuned GPU split\n";
}

using mertens_dr::int64;

int64 eval(int64 X, mdr_stats* st, double* wall) {
    const auto t0 = std::chrono::steady_clock::now();
    const int64 r = mertens_dr::DelégliseRivatEngine::compute_mertens_stats(X, 0, st);
    *wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

int64 eval_opts(mdr_ctx* c, int64 X, const mdr_opts& o, mdr_stats* st) {
    long long out = 0;
    const mdr_status s = mdr_mertens(c, X, &o, &out, st);
    if (s != MDR_OK) throw std::runtime_error(std::string("mdr_mertens: ") + mdr_last_error(c));
    return out;
}

// Warm time of one configuration, as the reference benchmark times (one
// warm-up call, then the median of three; bench_dr.cpp:48-73): the call
// before sizes the context's buffers for this configuration, so allocation
// of the other configuration's table sizes is not in the figure.
int64 eval_warm(mdr_ctx* c, int64 X, const mdr_opts& o, mdr_stats* st) {
    const int64 v = eval_opts(c, X, o, st);
    std::vector<mdr_stats> runs(3);
    for (mdr_stats& r : runs)
        if (eval_opts(c, X, o, &r) != v) throw std::runtime_error("M(X) changed between repetitions");
    std::sort(runs.begin(), runs.end(), [](const mdr_stats& a, const mdr_stats& b) { return a.total_time < b.total_time; });
    *st = runs[1];
    return v;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        usage(argv[0]);
        return 1;
    }
    const std::string a1 = argv[1];
    try {
        if (a1 == "--benchmark") {
            const int max_exp = argc >= 3 ? std::stoi(argv[2]) : 16;
            const int repeats = argc >= 5 ? std::stoi(argv[4]) : 1;
            if (repeats < 1 || max_exp < 1 || max_exp > 18) {
                std::cerr << "max_exp must be 1..18 and repeats positive." << std::endl;
                return 1;
            }
            std::cout << "==========================================================================================================\n"
                      << "      Deleglise-Rivat MI355X Mertens Benchmark Sweep (10^1..10^" << max_exp << ")\n"
                      << "==========================================================================================================\n"
                      << "Median of " << repeats << " run(s) per input.\n"
                      << "| Power | Target X            | Exact M(X)         | Table (s) | H2D (s)   | GPU Kernel (s) | Total Time (s) |\n"
                      << "| :---  | :---                | :---               | :---      | :---      | :---           | :---           |\n";
            int64 X = 10;
            for (int e = 1; e <= max_exp; ++e) {
                std::vector<double> wall;
                mdr_stats med{};
                int64 res = 0;
                std::vector<mdr_stats> sts;
                for (int r = 0; r < repeats; ++r) {
                    mdr_stats st{};
                    double w = 0;
                    const int64 v = eval(X, &st, &w);
                    if (r > 0 && v != res) {
                        std::cerr << "Inconsistent M(" << X << ") across benchmark runs." << std::endl;
                        return 2;
                    }
                    res = v;
                    wall.push_back(w);
                    sts.push_back(st);
                }
                std::vector<size_t> idx(wall.size());
                for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
                std::sort(idx.begin(), idx.end(), [&](size_t p, size_t q) { return wall[p] < wall[q]; });
                med = sts[idx[(idx.size() - 1) / 2]];
                std::cout << "| 10^" << std::setw(2) << std::left << e << " | " << std::setw(19) << std::right << X
                          << " | " << std::setw(18) << res << " | " << std::setw(9) << std::fixed
                          << std::setprecision(4) << med.table_ms / 1000.0 << " | " << std::setw(9)
                          << std::setprecision(4) << med.h2d_copy_time << " | " << std::setw(14) << std::setprecision(5)
                          << med.kernel_time << " | " << std::setw(14) << std::setprecision(5)
                          << wall[idx[(idx.size() - 1) / 2]] << " |\n";
                if (e < 18) X *= 10;
            }
            return 0;
        }
        if (a1 == "--compare") {
            const int max_exp = argc >= 3 ? std::stoi(argv[2]) : 14;
            if (max_exp < 10 || max_exp > 18) {
                std::cerr << "max_exp must be 10..18." << std::endl;
                return 1;
            }
            std::cout << "==========================================================================================================\n"
                      << "  Comparative Performance Benchmark: O(X^(3/4)) Dirichlet DP vs. O(X^(2/3)) Deleglise-Rivat Engine\n"
                      << "==========================================================================================================\n"
                      << "| Target X    | Exact M(X)        | Dirichlet DP Time | Deleglise-Rivat Time | Speedup (vs DP) |\n"
                      << "| :---        | :---              | :---              | :---                 | :---            |\n";
            int64 X = 10000000000LL;
            for (int e = 10; e <= max_exp; ++e) {
                mdr_stats st{};
                double dr_time = 0;
                const int64 dr_val = eval(X, &st, &dr_time);
                const auto t2 = std::chrono::steady_clock::now();
                const int64 dp_val = dirichlet::DirichletEngine::compute_mertens(X, 0);
                const double dp_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t2).count();
                if (dr_val != dp_val) {
                    std::cerr << "M(" << X << "): Deleglise-Rivat " << dr_val << " vs Dirichlet " << dp_val << std::endl;
                    return 2;
                }
                std::cout << "| 10^" << std::setw(2) << std::left << e << "      | " << std::setw(17) << std::right
                          << dr_val << " | " << std::setw(15) << std::fixed << std::setprecision(4) << dp_time
                          << " s | " << std::setw(18) << dr_time << " s | " << std::setw(14) << std::setprecision(2)
                          << dp_time / dr_time << "x |\n";
                if (e < 18) X *= 10;
            }
            return 0;
        }
        if (a1 == "--compare-split") {
            const int max_exp = argc >= 3 ? std::stoi(argv[2]) : 16;
            if (max_exp < 10 || max_exp > 18) {
                std::cerr << "max_exp must be 10..18." << std::endl;
                return 1;
            }
            mdr_ctx* c = nullptr;
            if (mdr_init(&c, 0) != MDR_OK) throw std::runtime_error("mdr_init failed");
            mdr_opts ref;
            mdr_default_opts(&ref);
            ref.no_table_reuse = 1;
            mdr_opts tuned = ref;   // the same (static, the drop-in's default) partition: parameters only
            tuned.params = MDR_PARAMS_TUNED;
            std::cout << "==========================================================================================================\n"
                      << "  Reference parameter rules vs the autotuned GPU split (both exact on the MI355X engine)\n"
                      << "==========================================================================================================\n"
                      << "| Target X    | Exact M(X)        | Reference u       | Reference (s)     | Tuned u           | Tuned (s)         | Speedup |\n"
                      << "| :---        | :---              | :---              | :---              | :---              | :---              | :---    |\n";
            int64 X = 10000000000LL;
            for (int e = 10; e <= max_exp; ++e) {
                mdr_stats a{}, b{};
                const int64 va = eval_warm(c, X, ref, &a);
                const int64 vb = eval_warm(c, X, tuned, &b);
                if (va != vb) {
                    std::cerr << "M(" << X << ") differs between the two splits." << std::endl;
                    mdr_shutdown(c);
                    return 2;
                }
                std::cout << "| 10^" << std::setw(2) << std::left << e << "      | " << std::setw(17) << std::right << va
                          << " | " << std::setw(17) << a.u << " | " << std::setw(17) << std::fixed
                          << std::setprecision(4) << a.total_time << " | " << std::setw(17) << b.u << " | "
                          << std::setw(17) << b.total_time << " | " << std::setw(6) << std::setprecision(2)
                          << a.total_time / b.total_time << "x |\n";
                if (e < 18) X *= 10;
            }
            mdr_shutdown(c);
            return 0;
        }
        const int64 X = std::stoll(a1);
        std::cout << "Computing M(" << X << ") with the Deleglise-Rivat MI355X engine..." << std::endl;
        mdr_stats st{};
        double wall = 0;
        const int64 r = eval(X, &st, &wall);
        std::cout << "------------------------------------------------------------\n"
                  << "M(" << X << ") = " << r << "\n"
                  << "------------------------------------------------------------\n"
                  << "Timing Breakdown:\n"
                  << std::fixed << std::setprecision(6)
                  << "  Host Sieve Precomputation: " << st.host_sieve_time << " s (the table is built on the GPU)\n"
                  << "  Host-to-Device Memory Copy:" << st.h2d_copy_time << " s\n"
                  << "  GPU Kernel Execution:      " << st.kernel_time << " s (table " << st.table_ms / 1000.0
                  << ", partition " << st.partition_ms / 1000.0 << ", quotient " << st.quotient_ms / 1000.0 << ")\n"
                  << "  Device-to-Host Memory Copy:" << st.d2h_copy_time << " s\n"
                  << "  Total Execution Time:      " << st.total_time << " s\n"
                  << "Elapsed time: " << wall << " seconds" << std::endl;
        return 0;
    } catch (const std::exception& ex) {
        std::cerr << ex.what() << std::endl;
        return 3;
    }
}
