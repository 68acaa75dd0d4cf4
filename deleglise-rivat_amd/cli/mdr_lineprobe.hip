// Random-line read probe: the rate at which the device serves scattered
// 16-byte reads, one per 64-byte line (the access pattern of the compressed
// table's sparse lookups, one CLine each), by buffer size.  It gives the
// memory-side ceiling the quotient kernel's table lookups run against: each
// L2 miss fetches a 128-byte line, so lines/s x 128 B is the L2 fill rate.
//
//   mdr_lineprobe [max GiB]     (default 24)
//
// Every thread walks its own hash sequence, so the loads of a wave are
// independent and many are in flight; the line index is reduced modulo the
// buffer's line count, so every read stays inside the buffer.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

__global__ void k_random_lines(const uint4* __restrict__ buf, unsigned long long nlines, int iters,
                               unsigned* sink) {
    const unsigned long long tid = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long h = tid * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    unsigned acc = 0;
    for (int i = 0; i < iters; ++i) {
        h ^= h >> 29;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 32;
        const uint4 v = buf[(h % nlines) * 4];   // 4 uint4 per 64-byte line
        acc += v.x ^ v.w;
    }
    if (acc == 0x9E3779B9u) sink[0] = acc;   // keeps the loads
}

bool check(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return false;
}

}  // namespace

int main(int argc, char** argv) {
    const double max_gib = argc > 1 ? std::atof(argv[1]) : 24.0;
    const size_t maxbytes = (size_t)(max_gib * (double)(1ull << 30)) / 64 * 64;
    uint4* buf = nullptr;
    unsigned* sink = nullptr;
    hipEvent_t e0, e1;
    if (!check(hipMalloc(&buf, maxbytes), "hipMalloc") || !check(hipMalloc(&sink, 64), "hipMalloc") ||
        !check(hipMemset(buf, 1, maxbytes), "hipMemset") || !check(hipEventCreate(&e0), "event") ||
        !check(hipEventCreate(&e1), "event"))
        return 1;
    const int blocks = 8192, threads = 256, iters = 512;
    const double reads = (double)blocks * threads * iters;
    std::printf("| buffer | ms per launch (best of 3) | G lines/s | TB/s at 128 B per line |\n|---|---|---|---|\n");
    for (size_t S = (size_t)16 << 20; S <= maxbytes; S *= 2) {
        float best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            if (!check(hipEventRecord(e0), "event")) return 1;
            hipLaunchKernelGGL(k_random_lines, dim3(blocks), dim3(threads), 0, 0, buf, (unsigned long long)(S / 64),
                               iters, sink);
            if (!check(hipGetLastError(), "launch") || !check(hipEventRecord(e1), "event") ||
                !check(hipEventSynchronize(e1), "sync"))
                return 1;
            float ms = 0;
            if (!check(hipEventElapsedTime(&ms, e0, e1), "elapsed")) return 1;
            if (ms < best) best = ms;
        }
        std::printf("| %.0f MiB | %.3f | %.2f | %.2f |\n", S / 1048576.0, best, reads / best / 1e6,
                    reads * 128 / best / 1e9);
        if (S * 2 > maxbytes && S != maxbytes) S = maxbytes / 2;   // the largest size last
    }
    (void)hipFree(buf);
    (void)hipFree(sink);
    return 0;
}
