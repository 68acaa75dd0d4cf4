// Tuned GPU split (cutoff u, dense tier, cx) and chunk length.
//
// The cache is a text file written by the project's autotuner
// (deleglise-rivat_amd/autotune.py): one line per decade of X,
//     decade  fx  cx  chunks  u_dense  [sieve shape]
// and optionally one line "batch fb" (the batch cutoff factor),
// where u = fx * (X / ln ln X)^(2/3) (the reference's cutoff form,
// mertens_dr.hpp:1401-1419, with a GPU-specific factor), chunks is the
// quotient-phase chunk count (0 = automatic) and u_dense the dense-tier
// limit.  The file named by MDR_TUNE_FILE, else tune_cache.txt beside the
// library, is read once; without it the built-in defaults apply.
#include "mdr_tune.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include <dlfcn.h>

namespace mdr {

namespace {

struct Entry {
    int decade;
    double fx, cx;
    i64 chunk;
    i64 ud;
    int sieve;   // sieve workgroup shape (0 = the default)
};

std::once_flag g_once;
std::vector<Entry> g_entries;
double g_batch_factor = 1.0;
// Largest tuned cutoff: 5e11 values (125 GB of compressed lines), lowered to
// what the device's memory holds (tune_set_device_memory).
constexpr long long U_CAP_MAX = 500000000000LL;
constexpr long long WINDOW_U_MAX = 1000000000000LL;
std::atomic<long long> g_u_cap{U_CAP_MAX};

void load() {
    std::string path;
    if (const char* e = std::getenv("MDR_TUNE_FILE")) path = e;
    else {
        Dl_info info;
        if (dladdr((void*)&tune_chunk, &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tune_cache.txt";
        }
    }
    if (path.empty()) return;
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return;
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        Entry e{};
        long long ch = 0, ud = 0;
        if (line[0] == '#') continue;
        double fb = 0;
        if (std::sscanf(line, "batch %lf", &fb) == 1) {
            if (fb >= 0.05 && fb <= 20) g_batch_factor = fb;
            continue;
        }
        int sv = 0;
        const int k = std::sscanf(line, "%d %lf %lf %lld %lld %d", &e.decade, &e.fx, &e.cx, &ch, &ud, &sv);
        if (k >= 4 && e.fx > 0.05 && e.fx < 100 && e.cx > 0.1 && e.cx < 5.0 && ch >= 0) {
            e.chunk = ch;
            e.ud = k >= 5 && ud >= 0 ? ud : 0;
            e.sieve = k >= 6 && sv >= 0 && sv < N_SIEVE_SHAPES ? sv : 0;
            g_entries.push_back(e);
        }
    }
    std::fclose(f);
}

i64 isqrt_t(i64 n) {
    i64 s = (i64)std::sqrt((double)n);
    while ((__int128)(s + 1) * (s + 1) <= n) ++s;
    while ((__int128)s * s > n) --s;
    return s;
}

}  // namespace

// Built-in defaults, measured on MI355X at X = 1e17 (sweep of u, the dense
// tier and cx): u = 0.7 (X / ln ln X)^(2/3) (6.4e10 at 1e17), a dense int16
// tier of about 2^20 entries (L2-resident; everything above it in 64-byte
// compressed lines, which keeps the hot part of the table in the last-level
// cache) and cx = 2.0, which moves work from table gathers to S2.
TuneChoice tune_lookup(i64 X, bool streaming) {
    (void)streaming;
    std::call_once(g_once, load);
    const int dec = (int)std::floor(std::log10((double)X));
    double fx = 0.7;
    double cx = X >= 10000000000000LL ? 2.0 : X >= 100000000000LL ? 0.95 : 0.70;
    i64 ud = 1048576;
    for (const Entry& e : g_entries)
        if (e.decade == dec) {
            fx = e.fx;
            cx = e.cx;
            if (e.ud > 0) ud = e.ud;
        }
    const double ll = std::log(std::max(2.0, std::log((double)X)));
    i64 u = (i64)(fx * std::pow((double)X / ll, 2.0 / 3.0));
    const i64 S = isqrt_t(X);
    u = std::max(u, 3 * S);
    const i64 ends = (i64)(cx * std::sqrt((double)X)) + 1000;
    u = std::max(u, ends + 64);
    TuneChoice t;
    // above the one-table cap the tuned cutoff is kept (up to 1e12, where the
    // quotient kernels' int32 partial sums stay exact) and the table is
    // windowed: measured at X = 2^63 - 1 on MI355X, u = 1e12 in windows of
    // 4e11 takes 19 % less GPU time than one 4e11 table
    const i64 cap = g_u_cap.load();
    if (u > cap) {
        t.window = cap;
        u = std::min<i64>(u, WINDOW_U_MAX);
    }
    u = ((u + LINE_VALUES - 1) / LINE_VALUES) * LINE_VALUES;
    t.u = u;
    t.cx = cx;
    t.ud = std::min(u, (ud / LINE_VALUES) * LINE_VALUES);
    return t;
}

// Shared cutoff of a batch.  The table is built once per call, so its cost
// (linear in the cutoff) is paid once while the quotient work of every X
// (proportional to X / sqrt(u)) falls as the cutoff grows: the optimum of
// a U + sum_i b X_i / sqrt(U) is the single-X optimum at X = sum_i X_i.
// The tuned batch factor (or factor > 0) scales it; the result is capped at
// 1e12 like tune_lookup (above the one-table cap the caller windows the
// table).
i64 tune_batch_u(double xsum, double factor) {
    std::call_once(g_once, load);
    const double xe = std::min(xsum, 9.0e18);
    const TuneChoice t = tune_lookup((i64)xe, true);
    double u = (double)t.u * (factor > 0 ? factor : g_batch_factor);
    u = std::min(u, (double)WINDOW_U_MAX);
    const i64 ui = ((i64)u / LINE_VALUES) * LINE_VALUES;
    return ui > 0 ? ui : LINE_VALUES;
}

i64 tune_chunks(i64 X) {
    std::call_once(g_once, load);
    const int dec = (int)std::floor(std::log10((double)(X > 1 ? X : 1)));
    for (const Entry& e : g_entries)
        if (e.decade == dec) return e.chunk;
    return 0;
}

i64 tune_chunk() { return tune_chunks(1); }

int tune_sieve_shape(i64 X) {
    std::call_once(g_once, load);
    const int dec = (int)std::floor(std::log10((double)(X > 1 ? X : 1)));
    for (const Entry& e : g_entries)
        if (e.decade == dec) return e.sieve;
    return 0;
}

i64 tune_u_cap() { return g_u_cap.load(); }

void tune_set_device_memory(size_t bytes) {
    // compressed lines take 1/4 byte per value: 45 % of the memory at most
    const double cap = std::min((double)U_CAP_MAX, (double)bytes * 0.45 * 4.0);
    g_u_cap.store(std::max<long long>(20000000000LL, (long long)cap));
}

}  // namespace mdr
