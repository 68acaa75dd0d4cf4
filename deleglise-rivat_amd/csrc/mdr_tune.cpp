// Tuned GPU split (cutoff u, dense tier, cx) and chunk length.
//
// The cache is a text file written by the project's autotuner
// (deleglise-rivat_amd/autotune.py): one line per decade of X,
//     decade  fx  cx  chunk
// where u = fx * (X / ln ln X)^(2/3) (the reference's cutoff form,
// mertens_dr.hpp:1401-1419, with a GPU-specific factor).  The file named by
// MDR_TUNE_FILE is read once; without it the built-in defaults apply.
#include "mdr_tune.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace mdr {

namespace {

struct Entry {
    int decade;
    double fx, cx;
    i64 chunk;
};

std::once_flag g_once;
std::vector<Entry> g_entries;
i64 g_default_chunk = 32768;

void load() {
    const char* path = std::getenv("MDR_TUNE_FILE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        Entry e{};
        long long ch = 0;
        if (line[0] == '#') continue;
        if (std::sscanf(line, "%d %lf %lf %lld", &e.decade, &e.fx, &e.cx, &ch) == 4) {
            if (e.fx > 0.05 && e.fx < 100 && e.cx > 0.1 && e.cx < 5.0 && ch >= 1024) {
                e.chunk = ch;
                g_entries.push_back(e);
            }
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

TuneChoice tune_lookup(i64 X, bool streaming) {
    std::call_once(g_once, load);
    const int dec = (int)std::floor(std::log10((double)X));
    double fx = streaming ? 4.0 : 1.0;
    double cx = X >= 1000000000000000LL ? 1.35 : X >= 10000000000000LL ? 1.15 : 0.95;
    for (const Entry& e : g_entries)
        if (e.decade == dec) { fx = e.fx; cx = e.cx; }
    const double ll = std::log(std::max(2.0, std::log((double)X)));
    i64 u = (i64)(fx * std::pow((double)X / ll, 2.0 / 3.0));
    const i64 S = isqrt_t(X);
    u = std::max(u, 3 * S);
    const i64 ends = (i64)(cx * std::sqrt((double)X)) + 1000;
    u = std::max(u, ends + 64);
    u = std::min<i64>(u, 64000000000LL);
    u = ((u + 255) / 256) * 256;
    TuneChoice t;
    t.u = u;
    t.cx = cx;
    t.ud = std::min(u, ((std::max<i64>(ends, 128000000LL) + 255) / 256) * 256);
    return t;
}

i64 tune_chunk() {
    std::call_once(g_once, load);
    return g_default_chunk;
}

}  // namespace mdr
