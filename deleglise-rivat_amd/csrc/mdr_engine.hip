// Host engine and C ABI of the MI355X Deleglise-Rivat Mertens engine.
//
// One context = one GPU, one stream, cached device buffers.  A call plans the
// parameters of every X (reference rules or the tuned GPU split), builds or
// reuses the table, partitions the quotient phase into chunks and runs the
// persistent kernel; results come back as 128-bit sums per X.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mertens_hip.h"
#include "mdr_device.hpp"
#include "mdr_kernels.hpp"
#include "mdr_mlp.hpp"
#include "mdr_tune.hpp"

#include <dlfcn.h>
#include <random>

using namespace mdr;

static_assert(sizeof(mdr_stamp) == sizeof(StampRec), "cycle-log record layout");
static_assert(MDR_MLP_NPARAM == MLP_NPARAM && MDR_MLP_NFEAT == MLP_IN, "cost-model size");

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return e;
        }
        size_t want = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    // ensure, after giving back an allocation far larger than needed (a big
    // table must not hold the memory the next one needs; calls that alternate
    // between nearby sizes keep their buffer)
    hipError_t fit(size_t bytes) {
        if (p && cap > 2 * bytes && cap - bytes > ((size_t)64 << 30)) release();
        return ensure(bytes);
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// One device-resident table: M(n) for n <= U (int16 dense tier up to ud,
// 64-byte compressed lines above), packed mu codes up to mu, and the segment
// sums of its build.
struct TableStore {
    bool valid = false;
    i64 U = 0, ud = 0, mu = 0, L = 0, nseg = 0, musegs = 0;
    DevBuf Md, cl, mu6, segsum;
    TableParam tp() const {
        TableParam t;
        t.U = U;
        t.ud = ud;
        t.L = L;
        t.mu_lim = mu;
        t.Md = Md.as<int16_t>();
        t.cl = cl.as<CLine>();
        t.mu6 = mu6.as<uint32_t>();
        t.pass = 0;
        t.qlo = ud + 1;
        t.qhi = U;
        return t;
    }
    void release() {
        Md.release();
        cl.release();
        mu6.release();
        segsum.release();
        valid = false;
    }
};

i64 isqrt_host(i64 n) {
    if (n <= 0) return 0;
    i64 s = (i64)std::sqrt((double)n);
    while ((__int128)(s + 1) * (s + 1) <= n) ++s;
    while ((__int128)s * s > n) --s;
    return s;
}

}  // namespace

struct mdr_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    int ncu = 0;
    int quot_blocks_per_cu = 1;
    int clock_khz = 0;
    std::mutex mu;
    std::string err;
    hipEvent_t ev[8] = {};

    // the cached table (reused by later calls that it covers)
    TableStore tab;
    DevBuf scan_tmp;
    // sieving primes
    i64 primes_root = -1;
    int n_w = 0, n_small = 0, n_split = 0, n_coop = 0, n_sq = 0, n_sq_small = 0;
    DevBuf stencil, pattern, pattern3, wprimes, wrcp, wweights, wlarge, sqprimes;
    // windows of the compressed tier (windowed tables): lines, segment sums,
    // scan scratch, [carry M(window start - 1), window total]
    DevBuf wcl, wseg, wscan, wcarry;
    // Dirichlet cross-check engine: M(x/k) for the large quotients
    DevBuf big;
    // work
    DevBuf xparam, slot_len, slot_cost, bounds, out128, ctr, stamps, comp_out, lookup, wspan;
    int wall_khz = 0;   // rate of the wall clock the quotient waves stamp
    void* na_host = nullptr;   // pinned result words of mdr_mertens_noalloc
    std::vector<StampRec> last_stamps;
    // instantiated graphs, one per launch sequence (see launch_graph)
    hipGraphExec_t graph_exec[16] = {};
    // learned cost model
    DevBuf mlp;
    bool mlp_ok = false;
    std::vector<float> mlp_host;

};

namespace {

mdr_status fail(mdr_ctx* c, mdr_status code, const std::string& m) {
    c->err = m;
    return code;
}
mdr_status hip_fail(mdr_ctx* c, hipError_t e, const char* what) {
    const mdr_status code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? MDR_ERR_NOMEM : MDR_ERR_HIP;
    return fail(c, code, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPT(call, what)                                  \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return hip_fail(c, e_, what); \
    } while (0)

// Copies and clears go through the context's own (non-blocking) stream and
// wait for it: nothing touches the legacy null stream, which must not depend
// on a stream that another thread is capturing into a graph.
hipError_t copy_sync(mdr_ctx* c, void* dst, const void* src, size_t n, hipMemcpyKind kind) {
    if (n == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, n, kind, c->st);
    return e == hipSuccess ? hipStreamSynchronize(c->st) : e;
}
hipError_t h2d(mdr_ctx* c, void* dst, const void* src, size_t n) { return copy_sync(c, dst, src, n, hipMemcpyHostToDevice); }
hipError_t d2h(mdr_ctx* c, void* dst, const void* src, size_t n) { return copy_sync(c, dst, src, n, hipMemcpyDeviceToHost); }
hipError_t clear_sync(mdr_ctx* c, void* dst, size_t n) {
    hipError_t e = hipMemsetAsync(dst, 0, n, c->st);
    return e == hipSuccess ? hipStreamSynchronize(c->st) : e;
}

// ---- parameters -------------------------------------------------------------
// Reference rules: mertens_dr.hpp:914-923 (cx), :931-962 (dense cutoff),
// :1401-1419 (streaming cutoff), :1451-1461 (dense tier and mu limit),
// :1027-1041 and :1421-1435 (strategy).
double ref_cx(i64 X, const mdr_opts& o) {
    if (o.cx > 0.1 && o.cx < 5.0) return o.cx;
    if (X >= 1000000000000000LL) return 1.35;
    if (X >= 10000000000000LL) return 1.15;
    if (X >= 100000000000LL) return 0.95;
    return 0.70;
}

i64 ref_dense_u(i64 X, const mdr_opts& o) {
    if (X <= 50000000LL) return X;
    const double ll = std::log(std::max(2.0, std::log((double)X)));
    const double fx = (o.fx >= 0.1 && o.fx <= 3.0) ? o.fx : 0.32;
    i64 u = (i64)(fx * std::pow((double)X / ll, 2.0 / 3.0));
    const i64 S = isqrt_host(X);
    if (u < 3 * S) u = 3 * S;
    const i64 cap = (o.u_cap_dense >= 100000000LL && o.u_cap_dense <= 4000000000LL) ? o.u_cap_dense : 1500000000LL;
    return std::min(u, cap);
}

i64 ref_stream_u(i64 X, const mdr_opts& o, double cx) {
    if (X < 50000000LL) return X;
    const double ll = std::log(std::max(2.0, std::log((double)X)));
    double fx = 0.95;
    if (X >= 1000000000000000LL) fx = 1.25;
    if (X >= 10000000000000000LL) fx = 1.50;
    i64 u = (i64)(fx * std::pow((double)X / ll, 2.0 / 3.0));
    const i64 S = isqrt_host(X);
    if (u < 3 * S) u = 3 * S;
    u = ((u + 255) / 256) * 256;
    i64 cap = (o.u_cap >= 100000000LL && o.u_cap <= 10000000000LL) ? o.u_cap : 8000000000LL;
    const i64 ends = (i64)(cx * std::sqrt((double)X)) + 1000;
    cap = std::max(cap, std::max(S + 1000, ends));
    return std::min(u, cap);
}

struct XPlan {
    i64 X = 0;
    int kind = 0;      // 0 constant, 1 table lookup, 2 Deleglise-Rivat sum
    i64 value = 0;     // kind 0
    bool streaming = false;
    i64 u = 0, ud = 0, mu_lim = 0, N = 0;
    double cx = 0;
    i64 window = 0;    // tuned plans whose cutoff exceeds one table: window size
};

mdr_status validate_opts(mdr_ctx* c, const mdr_opts& o) {
    if (o.strategy < 0 || o.strategy > 3) return fail(c, MDR_ERR_INVALID, "strategy out of range");
    if (o.cx != 0 && !(o.cx > 0.1 && o.cx < 5.0)) return fail(c, MDR_ERR_INVALID, "cx outside (0.1, 5.0)");
    if (o.u_cap != 0 && (o.u_cap < 100000000LL || o.u_cap > 10000000000LL))
        return fail(c, MDR_ERR_INVALID, "u_cap outside [1e8, 1e10]");
    if (o.fx != 0 && !(o.fx >= 0.1 && o.fx <= 3.0)) return fail(c, MDR_ERR_INVALID, "fx outside [0.1, 3.0]");
    if (o.u_cap_dense != 0 && (o.u_cap_dense < 100000000LL || o.u_cap_dense > 4000000000LL))
        return fail(c, MDR_ERR_INVALID, "u_cap_dense outside [1e8, 4e9]");
    if (o.partition < 0 || o.partition > 2) return fail(c, MDR_ERR_INVALID, "partition out of range");
    if (o.params < 0 || o.params > 1) return fail(c, MDR_ERR_INVALID, "params out of range");
    if (o.sched < 0 || o.sched > 1) return fail(c, MDR_ERR_INVALID, "sched out of range");
    if (o.chunk < 0) return fail(c, MDR_ERR_INVALID, "chunk must be >= 0");
    if (o.tune_u < 0 || o.tune_ud < 0 || o.tune_cx < 0 || (o.tune_cx != 0 && !(o.tune_cx > 0.1 && o.tune_cx < 5.0)))
        return fail(c, MDR_ERR_INVALID, "tuning override out of range");
    if (o.tune_u > 1000000000000LL) return fail(c, MDR_ERR_INVALID, "tune_u above 1e12");
    if (o.use_graph < 0 || o.use_graph > 1) return fail(c, MDR_ERR_INVALID, "use_graph must be 0 or 1");
    if (o.sieve_shape < 0 || o.sieve_shape > N_SIEVE_SHAPES)
        return fail(c, MDR_ERR_INVALID, "sieve_shape out of range");
    if (o.tune_batch != 0 && !(o.tune_batch >= 0.05 && o.tune_batch <= 20))
        return fail(c, MDR_ERR_INVALID, "tune_batch outside [0.05, 20]");
    if (o.window < 0) return fail(c, MDR_ERR_INVALID, "window must be >= 0");
    if (o.wave_span < 0 || o.wave_span > 1) return fail(c, MDR_ERR_INVALID, "wave_span must be 0 or 1");
    return MDR_OK;
}

XPlan plan_x(i64 X, const mdr_opts& o) {
    XPlan p;
    p.X = X;
    if (X < 1) { p.kind = 0; p.value = 0; return p; }
    if (X <= 3) { p.kind = 0; p.value = X == 1 ? 1 : X == 2 ? 0 : -1; return p; }
    const double cx = ref_cx(X, o);
    bool streaming;
    if (o.strategy == MDR_STRATEGY_FORCE_STREAMING) streaming = X > 100000LL;
    else if (o.strategy == MDR_STRATEGY_STREAMING) streaming = X >= 100000000000LL;
    else if (o.strategy == MDR_STRATEGY_DENSE) streaming = false;
    else streaming = X >= 100000000000LL && !o.fx_set && !o.u_cap_dense_set;
    if (X <= 50000000LL && !(streaming && o.strategy == MDR_STRATEGY_FORCE_STREAMING)) {
        p.kind = 1;   // sieve to X and read M(X)
        p.u = p.ud = p.mu_lim = X;
        return p;
    }
    p.streaming = streaming;
    p.cx = cx;
    const i64 A_max = (i64)(cx * std::sqrt((double)X)) + 1000;
    if (o.params == MDR_PARAMS_TUNED && X >= 100000000LL) {
        TuneChoice t = tune_lookup(X, streaming);
        if (o.tune_cx > 0) t.cx = o.tune_cx;
        if (o.tune_u > 0) t.u = std::max<i64>(((o.tune_u + LINE_VALUES - 1) / LINE_VALUES) * LINE_VALUES,
                                              ((i64)(t.cx * std::sqrt((double)X)) + 1000 + LINE_VALUES) / LINE_VALUES * LINE_VALUES);
        if (o.tune_ud > 0) t.ud = std::min(t.u, (o.tune_ud / LINE_VALUES) * LINE_VALUES);
        p.cx = t.cx;
        if (o.tune_u == 0) p.window = t.window;
        const i64 A2 = (i64)(p.cx * std::sqrt((double)X)) + 1000;
        p.u = t.u;
        p.ud = std::min(p.u, t.ud);
        p.N = X / p.u;
        p.mu_lim = std::max(A2, p.N + 1000);
    } else if (streaming) {
        p.u = ref_stream_u(X, o, cx);
        const i64 ends = (A_max + 255) & ~255LL;
        const i64 target = std::min<i64>(128000000LL, (p.u / 4 + 255) & ~255LL);
        p.ud = std::min(p.u, std::max(target, ends));
        p.N = X / p.u;
        p.mu_lim = std::max(A_max, p.N + 1000);
    } else {
        p.u = ref_dense_u(X, o);
        p.ud = p.u;
        p.N = X / p.u;
        p.mu_lim = std::max(A_max, p.N + 1000);
    }
    p.kind = X <= p.u ? 1 : 2;
    return p;
}

// Tuned batches share one cutoff: every X above it is summed with u = that
// cutoff (M(X) does not depend on u), every X at or below it becomes a table
// lookup.  Calls with one summed X, explicit overrides or the reference
// parameters keep their per-X plans.
void share_batch_cutoff(std::vector<XPlan>& plans, const mdr_opts& o) {
    if (o.params != MDR_PARAMS_TUNED || o.tune_u > 0) return;
    double xsum = 0;
    size_t nsum = 0;
    i64 umax = 0;
    for (const XPlan& p : plans)
        if (p.kind == 2) {
            xsum += (double)p.X;
            ++nsum;
            umax = std::max(umax, p.u);
        }
    if (nsum < 2) return;
    i64 ub = tune_batch_u(xsum, o.tune_batch);
    // above the one-table cap the compressed tier is windowed, which needs
    // every lookup inside the resident part: a batch with an X at or below
    // the shared cutoff (it would become a lookup) keeps the capped cutoff
    i64 window = 0;
    if (ub > tune_u_cap()) {
        bool inside = false;
        for (const XPlan& p : plans)
            if (p.kind == 2 && p.X <= ub) inside = true;
        if (inside) ub = tune_u_cap() / LINE_VALUES * LINE_VALUES;
        else window = tune_u_cap();
    }
    if (ub <= umax) return;
    for (XPlan& p : plans) {
        if (p.kind != 2) continue;
        p.window = std::max(p.window, window);
        if (p.u >= ub) continue;
        if (p.X <= ub) {
            p.kind = 1;
            p.u = p.X;
            p.mu_lim = 1;
            continue;
        }
        p.u = ub;
        p.N = p.X / p.u;
        const i64 A2 = (i64)(p.cx * std::sqrt((double)p.X)) + 1000;
        p.mu_lim = std::max(A2, p.N + 1000);
    }
}

// ---- sieving primes and stencil ------------------------------------------------
mdr_status ensure_primes(mdr_ctx* c, i64 U) {
    const i64 root = isqrt_host(U) + 1;
    if (root <= c->primes_root && c->stencil.p && c->pattern.p && c->pattern3.p) return MDR_OK;
    std::vector<uint8_t> comp((size_t)root + 1, 0);
    std::vector<uint32_t> wp, sq;
    std::vector<double> wr;
    std::vector<uint8_t> ww;
    std::vector<PrimeConsts> wl;
    for (i64 i = 2; i <= root; ++i) {
        if (comp[(size_t)i]) continue;
        for (i64 j = i * i; j <= root; j += i) comp[(size_t)j] = 1;
        if (i >= 13) {
            wp.push_back((uint32_t)i);
            wr.push_back(1.0 / (double)i);
            const unsigned bw = 64u - (unsigned)__builtin_clzll((unsigned long long)(i - 1));
            ww.push_back((uint8_t)(bw | 1u));
            if (i >= 256) wl.push_back(large_prime_consts((uint32_t)i));
        }
        if (i >= 7) sq.push_back((uint32_t)i);
    }
    c->n_w = (int)wp.size();
    c->n_small = (int)(std::lower_bound(wp.begin(), wp.end(), 256u) - wp.begin());
    c->n_split = (int)(std::lower_bound(wp.begin(), wp.end(), SQ_SPLIT) - wp.begin());
    c->n_coop = (int)(std::lower_bound(wp.begin(), wp.end(), COOP_PRIME_LIMIT) - wp.begin());
    wl.insert(wl.end(), PRIME_RECORD_PAD, PRIME_RECORD_END);
    c->n_sq = (int)sq.size();
    // squares of p < 256 are flagged block-cooperatively; larger ones with the weights
    c->n_sq_small = (int)(std::lower_bound(sq.begin(), sq.end(), 256u) - sq.begin());
    HIPT(c->wprimes.ensure(wp.size() * 4 + 4), "alloc primes");
    HIPT(c->wrcp.ensure(wr.size() * 8 + 8), "alloc primes");
    HIPT(c->wweights.ensure(ww.size() + 4), "alloc primes");
    HIPT(c->wlarge.ensure(wl.size() * sizeof(PrimeConsts)), "alloc primes");
    HIPT(c->sqprimes.ensure(sq.size() * 4 + 4), "alloc primes");
    if (!wp.empty()) {
        HIPT(h2d(c, c->wprimes.p, wp.data(), wp.size() * 4), "upload primes");
        HIPT(h2d(c, c->wrcp.p, wr.data(), wr.size() * 8), "upload primes");
        HIPT(h2d(c, c->wweights.p, ww.data(), ww.size()), "upload primes");
    }
    HIPT(h2d(c, c->wlarge.p, wl.data(), wl.size() * sizeof(PrimeConsts)), "upload primes");
    if (!sq.empty()) HIPT(h2d(c, c->sqprimes.p, sq.data(), sq.size() * 4), "upload primes");
    if (!c->stencil.p) {
        // value for n with n = x+1 (mod period): weights of 2,3,5,7,11 and the
        // square flags of 4, 9 and 25 (mertens_dr.hpp:784-793 uses 2..11, 4, 9).
        std::vector<uint8_t> st((size_t)STENCIL_PERIOD + SEG_MAX);
        for (size_t x = 0; x < st.size(); ++x) {
            const unsigned v = (unsigned)(x % STENCIL_PERIOD) + 1;
            unsigned a = 0;
            for (unsigned p : {2u, 3u, 5u, 7u, 11u})
                if (v % p == 0) a += (32u - (unsigned)__builtin_clz(p - 1)) | 1u;
            if (v % 4 == 0 || v % 9 == 0 || v % 25 == 0) a |= 0x80u;
            st[x] = (uint8_t)a;
        }
        HIPT(c->stencil.ensure(st.size()), "alloc stencil");
        HIPT(h2d(c, c->stencil.p, st.data(), st.size()), "upload stencil");
    }
    if (!c->pattern.p) {
        // weights of 13, 17, 19, 23 for n = x+1, over PATTERN_REPEAT periods and a segment
        std::vector<uint8_t> pt((size_t)PATTERN_REPEAT * PATTERN_PERIOD + SEG_MAX);
        for (size_t x = 0; x < pt.size(); ++x) pt[x] = (uint8_t)pattern_byte((unsigned)(x % PATTERN_PERIOD) + 1);
        HIPT(c->pattern.ensure(pt.size()), "alloc pattern");
        HIPT(h2d(c, c->pattern.p, pt.data(), pt.size()), "upload pattern");
    }
    if (!c->pattern3.p) {
        // weights of 29, 31, 37 for n = x+1, over PATTERN_REPEAT periods and a segment
        std::vector<uint8_t> pt((size_t)PATTERN_REPEAT * PATTERN3_PERIOD + SEG_MAX);
        for (size_t x = 0; x < pt.size(); ++x) pt[x] = (uint8_t)pattern3_byte((unsigned)(x % PATTERN3_PERIOD) + 1);
        HIPT(c->pattern3.ensure(pt.size()), "alloc pattern");
        HIPT(h2d(c, c->pattern3.p, pt.data(), pt.size()), "upload pattern");
    }
    c->primes_root = root;
    return MDR_OK;
}

// ---- table -----------------------------------------------------------------------
// A table request after validation: the sizes every kernel of the build sees.
struct TableJob {
    TableStore* t = nullptr;   // the table being built
    bool build = false;   // false: the cached table already covers the request
    int shape = 0;        // sieve workgroup shape (SIEVE_SHAPES)
    i64 U = 0, ud = 0, mu_lim = 0, NL = 0, nseg = 0, musegs = 0;
};

// Validates a table request and allocates what its build needs (host side:
// everything that may allocate happens here, before any kernel is enqueued).
mdr_status prepare_table(mdr_ctx* c, TableStore& t, i64 U, i64 ud, i64 mu_lim, bool reuse, int shape,
                         TableJob& j) {
    if (U < 1) U = 1;
    if (ud > U) ud = U;
    if (ud < U && (ud % LINE_VALUES) != 0)
        return fail(c, MDR_ERR_INVALID, "u_dense must be a multiple of 256 below U");
    if (mu_lim > U) return fail(c, MDR_ERR_INVALID, "mu_limit above U");
    if (mu_lim < 1) mu_lim = 1;
    if (U - ud > MAX_COMPRESSED_VALUES)
        return fail(c, MDR_ERR_RANGE, "table above the compressed-tier limit (U - u_dense > 256 * (2^32 - 1))");
    if (U > SIEVE_MAX_U) return fail(c, MDR_ERR_RANGE, "table above the sieve's limit (U >= 24576 * 2^32)");
    if (shape < 0 || shape >= N_SIEVE_SHAPES) return fail(c, MDR_ERR_INVALID, "sieve shape out of range");
    const i64 seg = SIEVE_SHAPES[shape].seg;
    j = TableJob{};
    j.t = &t;
    j.shape = shape;
    j.U = U;
    j.ud = ud;
    j.mu_lim = mu_lim;
    j.nseg = (U + seg - 1) / seg;
    j.musegs = (mu_lim + seg - 1) / seg;
    j.NL = U > ud ? (U - ud + LINE_VALUES - 1) / LINE_VALUES : 0;
    if (reuse && t.valid && t.U == U && t.ud == ud && t.mu == mu_lim) return MDR_OK;
    j.build = true;
    t.valid = false;
    mdr_status s = ensure_primes(c, U);
    if (s != MDR_OK) return s;
    HIPT(t.Md.fit((size_t)(ud + 1) * 2), "alloc dense table");
    HIPT(t.cl.fit((size_t)j.NL * sizeof(CLine)), "alloc compressed lines");
    HIPT(t.mu6.ensure((size_t)j.musegs * (seg / 48) * 4), "alloc mu");
    HIPT(t.segsum.ensure((size_t)j.nseg * 8), "alloc segment sums");
    HIPT(c->scan_tmp.ensure((size_t)scan_tmp_len(j.nseg) * 8), "alloc scan");
    // the quotient kernels read the table through these from now on; the
    // table counts as valid only once its build has completed
    t.U = U;
    t.ud = ud;
    t.mu = mu_lim;
    t.L = j.NL;
    t.nseg = j.nseg;
    t.musegs = j.musegs;
    return MDR_OK;
}

// Enqueues the build of a prepared table: kernel launches only (capturable).
mdr_status enqueue_table(mdr_ctx* c, const TableJob& j) {
    SieveArgs a;
    a.U = j.U;
    a.ud = j.ud;
    a.L = j.NL;
    a.mu_segs = j.musegs;
    a.stencil = c->stencil.as<uint8_t>();
    a.pattern = c->pattern.as<uint8_t>();
    a.pattern3 = c->pattern3.as<uint8_t>();
    a.wprimes = c->wprimes.as<uint32_t>();
    a.wrcp = c->wrcp.as<double>();
    a.wweights = c->wweights.as<uint8_t>();
    a.n_w = c->n_w;
    a.n_small = c->n_small;
    a.wlarge = c->wlarge.as<PrimeConsts>();
    a.n_split = c->n_split;
    a.n_coop = c->n_coop;
    a.sqprimes = c->sqprimes.as<uint32_t>();
    a.n_sq = c->n_sq;
    a.n_sq_small = c->n_sq_small;
    a.Md = j.t->Md.as<int16_t>();
    a.cl = j.t->cl.as<CLine>();
    a.mu6 = j.t->mu6.as<uint32_t>();
    a.segsum = j.t->segsum.as<i64>();
    a.seg0 = 0;
    // a dispatch holds fewer than 2^32 work-items: at most 2^31 threads per launch
    const i64 maxgrid = ((i64)1 << 31) / SIEVE_SHAPES[j.shape].threads;
    for (i64 s0 = 0; s0 < j.nseg; s0 += maxgrid) {
        a.seg_first = s0;
        const i64 nb = std::min(maxgrid, j.nseg - s0);
        HIPT(launch_sieve(j.shape, (unsigned)nb, c->st, a), "launch sieve");
    }
    HIPT(scan_exclusive_i64(j.t->segsum.as<i64>(), j.t->segsum.as<i64>(), j.nseg, c->scan_tmp.as<i64>(),
                            scan_tmp_len(j.nseg), nullptr, c->st),
         "scan segments");
    FixupArgs f;
    f.ud = j.ud;
    f.L = j.NL;
    f.nd = j.ud;
    f.seg_shift = SIEVE_SHAPES[j.shape].shift;
    f.Md = j.t->Md.as<int16_t>();
    f.cl = j.t->cl.as<CLine>();
    f.segoff = j.t->segsum.as<i64>();
    f.seg0 = 0;
    f.carry = nullptr;
    const i64 nfix = std::max<i64>(1, std::max(j.ud, j.NL));
    const i64 fix_blocks = std::min<i64>((nfix + 255) / 256, 1 << 20);   // grid-stride kernel
    hipLaunchKernelGGL(k_table_fixup, dim3((unsigned)fix_blocks), dim3(256), 0, c->st, f);
    HIPT(hipGetLastError(), "launch fixup");
    return MDR_OK;
}

double table_bytes(const TableJob& j) {
    return 2.0 * (double)(j.ud + 1) + (double)sizeof(CLine) * (double)j.NL +
           (double)j.musegs * (SIEVE_SHAPES[j.shape].seg / 12);
}

// Builds (or reuses) a table and waits for it.
mdr_status build_table(mdr_ctx* c, TableStore& t, i64 U, i64 ud, i64 mu_lim, mdr_stats* stats, bool reuse = true,
                       int shape = 0) {
    TableJob j;
    mdr_status s = prepare_table(c, t, U, ud, mu_lim, reuse, shape, j);
    if (s != MDR_OK) return s;
    if (!j.build) {
        if (stats) stats->table_reused = 1;
        return MDR_OK;
    }
    HIPT(hipEventRecord(c->ev[0], c->st), "event");
    s = enqueue_table(c, j);
    if (s != MDR_OK) return s;
    HIPT(hipEventRecord(c->ev[1], c->st), "event");
    HIPT(hipEventSynchronize(c->ev[1]), "table build");
    t.valid = true;
    if (stats) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
        stats->table_ms += ms;
        stats->table_bytes += table_bytes(j);
    }
    return MDR_OK;
}

__global__ void k_lookup(TableParam tp, const i64* qs, const int* xi, int n, u64* out128) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const i64 v = get_M(tp, qs[i]);
    out128[2 * xi[i]] = (u64)v;
    out128[2 * xi[i] + 1] = (u64)(v >> 63);
}

// carry[0] = M(q) from the resident table (add == nullptr), else carry[0] += add[0]:
// the prefix a window's checkpoints start from.
__global__ void k_window_carry(TableParam tp, i64 q, i64* carry, const i64* add) {
    if (blockIdx.x == 0 && threadIdx.x == 0) carry[0] = add ? carry[0] + add[0] : (i64)get_M(tp, q);
}

// ---- one call: table, lookups, partition, quotient phase -------------------------------
// The call is prepared on the host (sizes, allocations, uploads), then enqueued
// as kernel launches only, either straight onto the engine's stream or
// captured into a hipGraph that is instantiated once per shape of the launch
// sequence and updated in place (hipGraphExecUpdate) for later calls.
struct BatchLaunch {
    TableJob tj;
    TableParam tp{};
    size_t nx = 0, nxs = 0, nlook = 0;
    i64 slots = 0, P = 0, cap = 0;
    bool learned = false;
    PartArgs pa{};
    QuotArgs qa{};
};

// Partition and quotient phase of one pass: slot lengths (and predicted
// costs), their scans, the chunk cuts and the persistent quotient kernel.
mdr_status enqueue_quotient(mdr_ctx* c, const BatchLaunch& b, const PartArgs& pa, const QuotArgs& qa, bool timed) {
    unsigned long long* ctr = c->ctr.as<unsigned long long>();
    if (timed) HIPT(hipEventRecord(c->ev[2], c->st), "event");
    hipLaunchKernelGGL(k_slot_len, dim3((unsigned)((b.slots + 255) / 256)), dim3(256), 0, c->st, pa);
    HIPT(hipGetLastError(), "launch slot lengths");
    HIPT(scan_exclusive_i64(pa.len, pa.len, b.slots, c->scan_tmp.as<i64>(), scan_tmp_len(b.slots),
                            (i64*)(ctr + 4), c->st),
         "scan positions");
    if (b.learned)
        HIPT(scan_exclusive_i64(pa.cost, pa.cost, b.slots, c->scan_tmp.as<i64>(), scan_tmp_len(b.slots),
                                (i64*)(ctr + 5), c->st),
             "scan costs");
    hipLaunchKernelGGL(k_cuts, dim3((unsigned)((b.P + 1 + 255) / 256)), dim3(256), 0, c->st, pa);
    HIPT(hipGetLastError(), "launch cuts");
    if (timed) HIPT(hipEventRecord(c->ev[3], c->st), "event");
    const unsigned grid = (unsigned)(c->ncu * c->quot_blocks_per_cu);
    hipLaunchKernelGGL(k_quot, dim3(grid), dim3(256), 0, c->st, qa);
    HIPT(hipGetLastError(), "launch quotient kernel");
    return MDR_OK;
}

// Kernel launches of one call (no allocation, no memset, no copy, no host
// wait), so the same code runs eagerly and under stream capture.
mdr_status enqueue_batch(mdr_ctx* c, const BatchLaunch& b, bool timed) {
    unsigned long long* ctr = c->ctr.as<unsigned long long>();
    if (timed) HIPT(hipEventRecord(c->ev[0], c->st), "event");
    const i64 nclear = std::max<i64>(8, (i64)std::max(b.nx, b.nxs) * 2);
    hipLaunchKernelGGL(k_clear3, dim3((unsigned)((nclear + 255) / 256)), dim3(256), 0, c->st, c->out128.as<u64>(),
                       (i64)b.nx * 2, (u64*)ctr, (i64)8, c->comp_out.as<u64>(), (i64)b.nxs * 2);
    HIPT(hipGetLastError(), "launch clear");
    if (b.tj.build) {
        mdr_status s = enqueue_table(c, b.tj);
        if (s != MDR_OK) return s;
    }
    if (timed) HIPT(hipEventRecord(c->ev[1], c->st), "event");
    if (b.nlook) {
        const char* lk = static_cast<const char*>(c->lookup.p);
        hipLaunchKernelGGL(k_lookup, dim3((unsigned)((b.nlook + 255) / 256)), dim3(256), 0, c->st, b.tp,
                           (const i64*)lk, (const int*)(lk + b.nlook * 8), (int)b.nlook, c->out128.as<u64>());
        HIPT(hipGetLastError(), "launch lookup");
    }
    if (b.nxs) {
        mdr_status s = enqueue_quotient(c, b, b.pa, b.qa, timed);
        if (s != MDR_OK) return s;
    }
    if (timed) HIPT(hipEventRecord(c->ev[4], c->st), "event");
    return MDR_OK;
}

// ---- windowed tables --------------------------------------------------------------------
// The resident table covers 1..U0 (mu, the dense tier, every M the constants
// and lookups read, and the compressed tier up to U0); the compressed tier of
// U0 < q <= U streams through windows [L, R] of at most wvals values (whole
// sieve segments).  Each window is sieved into its own lines, its checkpoints
// continue from the carried M(L - 1), and a quotient pass sums the S1
// lookups whose quotient falls in it (the positions do not depend on the
// window, so the sum over the passes equals the one-table sum).
// Reference: compute_mertens_windowed (mertens_dr_windowed.hpp:49-407), which
// streams int16 M windows over 1..u built from a retained mu sieve.
struct WindowPlan {
    i64 U0 = 0, U = 0, wvals = 0;
    int shape = 0;
    i64 count() const { return wvals > 0 && U > U0 ? (U - U0 + wvals - 1) / wvals : 0; }
};

mdr_status prepare_windows(mdr_ctx* c, const WindowPlan& w) {
    if (w.U > SIEVE_MAX_U) return fail(c, MDR_ERR_RANGE, "windows above the sieve's limit (U >= 24576 * 2^32)");
    mdr_status st = ensure_primes(c, w.U);   // the windows sieve up to sqrt(U)
    if (st != MDR_OK) return st;
    const i64 seg = SIEVE_SHAPES[w.shape].seg;
    const i64 nl = (w.wvals + LINE_VALUES - 1) / LINE_VALUES;
    const i64 ns = w.wvals / seg + 1;
    HIPT(c->wcl.fit((size_t)nl * sizeof(CLine)), "alloc window lines");
    HIPT(c->wseg.ensure((size_t)ns * 8), "alloc window sums");
    HIPT(c->wscan.ensure((size_t)scan_tmp_len(ns) * 8), "alloc window scan");
    HIPT(c->wcarry.ensure(16), "alloc window carry");
    return MDR_OK;
}

TableParam window_param(mdr_ctx* c, const TableParam& res, i64 L, i64 R) {
    TableParam t = res;   // mu, the dense tier and mu_lim stay the resident ones
    t.U = R;
    t.ud = L - 1;
    t.L = (R - L + 1 + LINE_VALUES - 1) / LINE_VALUES;
    t.cl = c->wcl.as<CLine>();
    t.pass = 2;
    t.qlo = L;
    t.qhi = R;
    return t;
}

// Sieves window [L, R] into the window lines and advances the carry.
mdr_status enqueue_window_table(mdr_ctx* c, const WindowPlan& w, i64 L, i64 R, const TableParam& res) {
    const SieveShape& sh = SIEVE_SHAPES[w.shape];
    const i64 seg = sh.seg;
    const i64 s0 = (L - 1) / seg, s1 = (R + seg - 1) / seg;
    const i64 NL = (R - L + 1 + LINE_VALUES - 1) / LINE_VALUES;
    i64* carry = c->wcarry.as<i64>();
    if (L == w.U0 + 1) {
        hipLaunchKernelGGL(k_window_carry, dim3(1), dim3(64), 0, c->st, res, w.U0, carry, (const i64*)nullptr);
        HIPT(hipGetLastError(), "launch window carry");
    }
    SieveArgs a;
    a.U = R;
    a.ud = L - 1;   // every value of the window lies above it: lines only
    a.L = NL;
    a.mu_segs = 0;  // mu comes from the resident table
    a.stencil = c->stencil.as<uint8_t>();
    a.pattern = c->pattern.as<uint8_t>();
    a.pattern3 = c->pattern3.as<uint8_t>();
    a.wprimes = c->wprimes.as<uint32_t>();
    a.wrcp = c->wrcp.as<double>();
    a.wweights = c->wweights.as<uint8_t>();
    a.n_w = c->n_w;
    a.n_small = c->n_small;
    a.wlarge = c->wlarge.as<PrimeConsts>();
    a.n_split = c->n_split;
    a.n_coop = c->n_coop;
    a.sqprimes = c->sqprimes.as<uint32_t>();
    a.n_sq = c->n_sq;
    a.n_sq_small = c->n_sq_small;
    a.Md = c->tab.Md.as<int16_t>();
    a.cl = c->wcl.as<CLine>();
    a.mu6 = c->tab.mu6.as<uint32_t>();
    a.segsum = c->wseg.as<i64>();
    a.seg0 = s0;
    const i64 maxgrid = ((i64)1 << 31) / sh.threads;
    for (i64 b0 = s0; b0 < s1; b0 += maxgrid) {
        a.seg_first = b0;
        const i64 nb = std::min(maxgrid, s1 - b0);
        HIPT(launch_sieve(w.shape, (unsigned)nb, c->st, a), "launch window sieve");
    }
    HIPT(scan_exclusive_i64(c->wseg.as<i64>(), c->wseg.as<i64>(), s1 - s0, c->wscan.as<i64>(), scan_tmp_len(s1 - s0),
                            carry + 1, c->st),
         "scan window");
    FixupArgs f;
    f.ud = L - 1;
    f.L = NL;
    f.nd = 0;
    f.seg_shift = sh.shift;
    f.Md = c->tab.Md.as<int16_t>();
    f.cl = c->wcl.as<CLine>();
    f.segoff = c->wseg.as<i64>();
    f.seg0 = s0;
    f.carry = carry;
    const i64 fix_blocks = std::min<i64>((std::max<i64>(NL, 1) + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_table_fixup, dim3((unsigned)fix_blocks), dim3(256), 0, c->st, f);
    HIPT(hipGetLastError(), "launch window fixup");
    hipLaunchKernelGGL(k_window_carry, dim3(1), dim3(64), 0, c->st, res, (i64)0, carry, (const i64*)(carry + 1));
    HIPT(hipGetLastError(), "launch window carry");
    return MDR_OK;
}

// Captures the call into a graph and launches it, reusing the instantiated
// graph of the previous call with the same launch sequence.
mdr_status launch_graph(mdr_ctx* c, const BatchLaunch& b, bool& instantiated) {
    HIPT(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal), "begin capture");
    const mdr_status s = enqueue_batch(c, b, false);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->st, &g);
    if (s != MDR_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
    }
    if (ec != hipSuccess) return hip_fail(c, ec, "end capture");
    // one cached executable per launch sequence: with and without the table build,
    // with and without the learned cost scan, with and without quotient work
    const int key = (b.tj.build ? 1 : 0) | (b.learned ? 2 : 0) | (b.nxs ? 4 : 0) | (b.nlook ? 8 : 0);
    hipGraphExec_t& ex = c->graph_exec[key];
    instantiated = true;
    if (ex) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult r = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(ex, g, &bad, &r) == hipSuccess && r == hipGraphExecUpdateSuccess) {
            instantiated = false;
        } else {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(ex);
            ex = nullptr;
        }
    }
    hipError_t e = hipSuccess;
    if (instantiated) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        ex = nullptr;
        return hip_fail(c, e, "instantiate graph");
    }
    HIPT(hipEventRecord(c->ev[5], c->st), "event");
    HIPT(hipGraphLaunch(ex, c->st), "launch graph");
    HIPT(hipEventRecord(c->ev[6], c->st), "event");
    return MDR_OK;
}

mdr_status run_batch(mdr_ctx* c, std::vector<XPlan>& plans, const mdr_opts& o, long long* out,
                     mdr_stats* stats) {
    const size_t nx = plans.size();
    // table covering every X
    i64 U = 1, mu_lim = 1, ud = 0;
    bool any_sum = false, need_table = false;
    for (const XPlan& p : plans) {
        if (p.kind == 0) continue;
        need_table = true;
        U = std::max(U, std::max(p.u, p.mu_lim));
        if (p.kind == 1) U = std::max(U, p.X);
        if (p.kind == 2) {
            any_sum = true;
            ud = std::max(ud, p.ud);
            mu_lim = std::max(mu_lim, p.mu_lim);
        }
    }
    // one dense tier for the batch: the largest any X asked for; lookups
    // alone keep a dense table only up to 5e7 (where |M| fits the int16 tier)
    ud = any_sum ? std::min(U, ud) : U <= 50000000LL ? U : 1048576;
    if (ud < U) ud = (ud / LINE_VALUES) * LINE_VALUES;
    // sieve workgroup shape: explicit, else the tuned one of the largest X
    int shape = o.sieve_shape > 0 ? o.sieve_shape - 1 : 0;
    if (need_table && o.sieve_shape == 0 && o.params == MDR_PARAMS_TUNED) {
        i64 xmax = 1;
        for (const XPlan& p : plans)
            if (p.kind != 0) xmax = std::max(xmax, p.X);
        shape = tune_sieve_shape(xmax);
    }
    // windowed table: the resident part holds whole segments up to what mu,
    // the dense tier, the constants (A, B <= cx sqrt(X) < mu_limit) and the
    // lookups read; the compressed tier above it streams through windows
    WindowPlan win;
    i64 wreq = o.window;   // explicit, else a tuned plan's
    for (const XPlan& p : plans)
        if (o.window == 0 && p.kind == 2) wreq = std::max(wreq, p.window);
    if (need_table && any_sum && wreq > 0) {
        const i64 seg = SIEVE_SHAPES[shape].seg;
        i64 need = std::max(mu_lim, ud);
        for (const XPlan& p : plans)
            if (p.kind == 1) need = std::max(need, p.X);
        const i64 U0 = std::max<i64>(seg, (need + seg - 1) / seg * seg);
        if (U0 < U) {
            win.U0 = U0;
            win.U = U;
            win.wvals = std::max<i64>(seg, (wreq + seg - 1) / seg * seg);
            win.shape = shape;
            U = U0;
            if (ud > U) ud = U;
        }
    }
    const bool windowed = win.count() > 0;
    // tuned calls: a cached table that covers the request serves it, and every
    // summed X takes the cached cutoff (M(X) does not depend on u)
    if (need_table && !windowed && !o.no_table_reuse && c->tab.valid && o.params == MDR_PARAMS_TUNED &&
        o.tune_u == 0 && c->tab.U >= U && c->tab.mu >= mu_lim) {
        for (XPlan& p : plans) {
            if (p.kind != 2 || p.u >= c->tab.U) continue;
            if (p.X <= c->tab.U) {
                p.kind = 1;
                p.u = p.X;
                p.mu_lim = 1;
                continue;
            }
            p.u = c->tab.U;
            p.N = p.X / p.u;
            const i64 A2 = (i64)(p.cx * std::sqrt((double)p.X)) + 1000;
            p.mu_lim = std::max(A2, p.N + 1000);
        }
        U = c->tab.U;
        ud = c->tab.ud;
        mu_lim = c->tab.mu;
    }
    BatchLaunch b;
    b.nx = nx;
    if (need_table) {
        if (!windowed) c->wcl.release();   // a one-table call needs the memory
        mdr_status s = prepare_table(c, c->tab, U, ud, mu_lim, !o.no_table_reuse, shape, b.tj);
        if (s != MDR_OK) return s;
        if (stats && !b.tj.build) stats->table_reused = 1;
        if (windowed) {
            s = prepare_windows(c, win);
            if (s != MDR_OK) return s;
        }
    }
    b.tp = c->tab.tp();

    std::vector<XParam> xps;
    std::vector<int> xidx;
    std::vector<i64> lq;
    std::vector<int> lx;
    i64 slots = 0;
    for (size_t i = 0; i < nx; ++i) {
        const XPlan& p = plans[i];
        if (p.kind == 1) { lq.push_back(p.X); lx.push_back((int)i); }
        if (p.kind != 2) continue;
        XParam xp{};
        xp.X = p.X;
        xp.u = p.u;
        xp.N = p.N;
        xp.T = count_coprime6(p.N);
        xp.slot_base = slots;
        xp.cx = p.cx;
        xp.n6 = p.N / 6;
        xp.n3 = p.N / 3;
        xp.n2 = p.N / 2;
        slots += 2 * xp.T;
        xps.push_back(xp);
        xidx.push_back((int)i);
    }
    b.nxs = xps.size();
    b.nlook = lq.size();
    b.slots = slots;

    // ---- allocations and uploads: nothing after this block allocates ----
    const auto th0 = std::chrono::steady_clock::now();
    HIPT(c->out128.ensure(nx * 16 + 16), "alloc results");
    HIPT(c->ctr.ensure(64), "alloc counters");
    HIPT(c->comp_out.ensure(b.nxs * 16 + 16), "alloc accumulators");
    if (b.nlook) {
        HIPT(c->lookup.ensure(b.nlook * 12 + 16), "alloc lookups");
        HIPT(hipMemcpyAsync(c->lookup.p, lq.data(), b.nlook * 8, hipMemcpyHostToDevice, c->st), "upload");
        HIPT(hipMemcpyAsync((char*)c->lookup.p + b.nlook * 8, lx.data(), b.nlook * 4, hipMemcpyHostToDevice, c->st),
             "upload");
    }
    if (b.nxs) {
        i64 xmax_sum = 0;
        for (const XPlan& p : plans)
            if (p.kind == 2) xmax_sum = std::max(xmax_sum, p.X);
        b.learned = o.partition == MDR_PARTITION_LEARNED && c->mlp_ok && xmax_sum >= LEARNED_MIN_X;
        const i64 nwaves = (i64)c->ncu * c->quot_blocks_per_cu * 4;
        i64 tuned_chunks = 0;
        if (o.params == MDR_PARAMS_TUNED) {
            i64 xmax = 1;
            for (const XPlan& p : plans)
                if (p.kind == 2) xmax = std::max(xmax, p.X);
            tuned_chunks = tune_chunks(xmax);
        }
        b.P = o.chunk > 0 ? o.chunk
              : tuned_chunks > 0 ? tuned_chunks
                                 : nwaves * (o.sched == MDR_SCHED_DYNAMIC ? 64 : 4);
        HIPT(c->xparam.ensure(b.nxs * sizeof(XParam)), "alloc params");
        HIPT(c->slot_len.ensure((size_t)slots * 8), "alloc slot lengths");
        if (b.learned) HIPT(c->slot_cost.ensure((size_t)slots * 8), "alloc slot costs");
        HIPT(c->bounds.ensure((size_t)(b.P + 1) * 8), "alloc chunk bounds");
        HIPT(c->scan_tmp.ensure((size_t)std::max(scan_tmp_len(slots), scan_tmp_len(b.tj.nseg)) * 8), "alloc scan");
        b.cap = o.record_cycles ? (i64)1 << 22 : 0;
        if (b.cap) HIPT(c->stamps.ensure((size_t)b.cap * sizeof(StampRec)), "alloc cycle log");
        HIPT(hipMemcpyAsync(c->xparam.p, xps.data(), b.nxs * sizeof(XParam), hipMemcpyHostToDevice, c->st),
             "upload params");
        // ctr layout: [0] work queue, [1..3] position counts, [4] total positions,
        // [5] total predicted cycles, [6] cycle-log count
        unsigned long long* ctr = c->ctr.as<unsigned long long>();
        PartArgs& pa = b.pa;
        pa.xp = c->xparam.as<XParam>();
        pa.nx = (int)b.nxs;
        pa.total_slots = slots;
        pa.tp = b.tp;
        pa.mlp = b.learned ? c->mlp.as<float>() : nullptr;
        pa.len = c->slot_len.as<i64>();
        pa.cost = b.learned ? c->slot_cost.as<i64>() : nullptr;
        pa.type_pos = ctr + 1;
        pa.totals = reinterpret_cast<const i64*>(ctr + 4);
        pa.P = b.P;
        pa.bounds = c->bounds.as<i64>();
        // automatic chunk counts adapt to the call's work (explicit ones are kept)
        pa.min_chunk = o.chunk > 0 ? 0 : MIN_CHUNK_POSITIONS;
        pa.by_slot = o.partition == MDR_PARTITION_TERMS ? 1 : 0;
        QuotArgs& qa = b.qa;
        qa.xp = pa.xp;
        qa.nx = pa.nx;
        qa.tp = b.tp;
        qa.off = pa.len;
        qa.nslots = slots;
        qa.bounds = pa.bounds;
        qa.P = b.P;
        qa.next = ctr;
        qa.dynamic = o.sched == MDR_SCHED_DYNAMIC ? 1 : 0;
        qa.stamps = b.cap ? c->stamps.as<StampRec>() : nullptr;
        qa.nstamps = ctr + 6;
        qa.stamp_cap = b.cap;
        qa.stamp_stride = o.record_cycles > 1 ? o.record_cycles : 1;
        qa.out128 = c->comp_out.as<u64>();
        qa.wave_span = nullptr;
        if (o.wave_span) {
            HIPT(c->wspan.ensure((size_t)nwaves * 16), "alloc wave stamps");
            qa.wave_span = c->wspan.as<u64>();
        }
    }
    if (stats) stats->h2d_copy_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - th0).count();

    // ---- launch ----
    bool instantiated = false;
    const bool graph = o.use_graph && !windowed;   // windowed calls run eagerly
    double tab_ms = 0, part_ms = 0, quot_ms = 0, graph_ms = 0, pred_cycles = 0;
    auto phase_ms = [&](int e0, int e1) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev[e0], c->ev[e1]);
        return (double)ms;
    };
    if (graph) {
        mdr_status s = launch_graph(c, b, instantiated);
        if (s != MDR_OK) return s;
    } else {
        mdr_status s = enqueue_batch(c, b, true);
        if (s != MDR_OK) return s;
    }
    HIPT(hipStreamSynchronize(c->st), "batch");
    if (b.tj.build) c->tab.valid = true;
    if (b.learned && windowed) {   // the resident pass's predicted total
        long long pc = 0;
        HIPT(d2h(c, &pc, c->ctr.as<unsigned long long>() + 5, 8), "download");
        pred_cycles += (double)pc;
    }
    if (graph) {
        graph_ms = phase_ms(5, 6);
    } else {
        if (b.tj.build) tab_ms += phase_ms(0, 1);
        if (b.nxs) {
            part_ms += phase_ms(2, 3);
            quot_ms += phase_ms(3, 4);
        }
    }
    // windows of the compressed tier above the resident part: sieve the
    // window, then a quotient pass over the lookups that fall in it
    const TableParam rtp = b.tp;
    for (i64 L = win.U0 + 1; windowed && L <= win.U; L += win.wvals) {
        const i64 R = std::min(win.U, L + win.wvals - 1);
        HIPT(hipEventRecord(c->ev[0], c->st), "event");
        mdr_status s = enqueue_window_table(c, win, L, R, rtp);
        if (s != MDR_OK) return s;
        HIPT(hipEventRecord(c->ev[1], c->st), "event");
        PartArgs pa = b.pa;
        QuotArgs qa = b.qa;
        pa.tp = window_param(c, rtp, L, R);
        qa.tp = pa.tp;
        hipLaunchKernelGGL(k_clear3, dim3(1), dim3(64), 0, c->st, (u64*)c->ctr.as<unsigned long long>(), (i64)1,
                           (u64*)nullptr, (i64)0, (u64*)nullptr, (i64)0);   // the chunk queue
        HIPT(hipGetLastError(), "launch clear");
        s = enqueue_quotient(c, b, pa, qa, true);
        if (s != MDR_OK) return s;
        HIPT(hipEventRecord(c->ev[4], c->st), "event");
        HIPT(hipStreamSynchronize(c->st), "window");
        if (b.learned) {   // this pass's predicted total (the scan overwrites it per pass)
            long long pc = 0;
            HIPT(d2h(c, &pc, c->ctr.as<unsigned long long>() + 5, 8), "download");
            pred_cycles += (double)pc;
        }
        tab_ms += phase_ms(0, 1);
        part_ms += phase_ms(2, 3);
        quot_ms += phase_ms(3, 4);
        if (stats) {
            stats->windows += 1;
            stats->table_bytes += (double)sizeof(CLine) * (double)pa.tp.L;
        }
    }

    // ---- results ----
    const auto td0 = std::chrono::steady_clock::now();
    std::vector<u64> res(nx * 2);
    if (nx) HIPT(d2h(c, res.data(), c->out128.p, nx * 16), "download");
    unsigned long long ch[7] = {};
    if (b.nxs) {
        std::vector<u64> acc_h(b.nxs * 2);
        HIPT(d2h(c, acc_h.data(), c->comp_out.p, b.nxs * 16), "download");
        for (size_t i = 0; i < b.nxs; ++i) {
            res[2 * xidx[i]] = acc_h[2 * i];
            res[2 * xidx[i] + 1] = acc_h[2 * i + 1];
        }
        HIPT(d2h(c, ch, c->ctr.p, sizeof ch), "download");
    }
    for (size_t i = 0; i < nx; ++i) {
        if (plans[i].kind == 0) { out[i] = plans[i].value; continue; }
        const i64 lo = (i64)res[2 * i], hi = (i64)res[2 * i + 1];
        if (hi != (lo >> 63)) return fail(c, MDR_ERR_RANGE, "result does not fit in 64 bits");
        out[i] = lo;
    }
    if (b.nxs && o.wave_span && stats) {
        // every resident wave stamps once per launch (last launch of the call)
        const i64 nwaves = (i64)c->ncu * c->quot_blocks_per_cu * 4;
        std::vector<u64> sp((size_t)nwaves * 2);
        HIPT(d2h(c, sp.data(), c->wspan.p, sp.size() * 8), "download wave stamps");
        u64 t0 = ~0ull, t1 = 0;
        double busy = 0;
        for (i64 w = 0; w < nwaves; ++w) {
            t0 = std::min(t0, sp[2 * w]);
            t1 = std::max(t1, sp[2 * w + 1]);
            busy += (double)(sp[2 * w + 1] - sp[2 * w]);
        }
        const double khz = c->wall_khz > 0 ? (double)c->wall_khz : 100000.0;
        stats->wave_span_ms = (double)(t1 - t0) / khz;
        stats->wave_busy_ms = busy / (double)nwaves / khz;
        stats->wave_busy_frac = t1 > t0 ? busy / (double)nwaves / (double)(t1 - t0) : 0.0;
    }
    if (b.cap) {
        const i64 n = std::min<i64>((i64)ch[6], b.cap);
        c->last_stamps.resize((size_t)n);
        if (n) HIPT(d2h(c, c->last_stamps.data(), c->stamps.p, (size_t)n * sizeof(StampRec)),
                    "download cycle log");
    }
    if (stats) {
        stats->d2h_copy_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - td0).count();
        if (graph) {
            stats->graph_ms += graph_ms;
            stats->graph_used = 1;
            stats->graph_instantiated += instantiated ? 1 : 0;
            stats->kernel_time += graph_ms / 1000.0;
        } else {
            stats->table_ms += tab_ms;
            stats->partition_ms += part_ms;
            stats->quotient_ms += quot_ms;
            stats->kernel_time = (stats->table_ms + stats->partition_ms + stats->quotient_ms) / 1000.0;
        }
        if (b.tj.build) stats->table_bytes += table_bytes(b.tj);
        if (b.nxs) {
            stats->chunks += b.P;
            stats->slots += slots;
            stats->s1_terms += (long long)(ch[1] + ch[2]);
            stats->s1_comp_terms += (long long)ch[2];
            stats->s2_terms += (long long)ch[3];
            stats->predicted_cycles += windowed ? pred_cycles : b.learned ? (double)(long long)ch[5] : 0.0;
            stats->measured_cycles += (graph ? 0.0 : quot_ms) * (double)c->clock_khz *
                                      (double)((i64)c->ncu * c->quot_blocks_per_cu * 4);
            stats->quotient_bytes += 2.0 * (double)ch[1] + 20.0 * (double)ch[2] + 3.0 * (double)ch[3];
            stats->partition_used = b.learned ? MDR_PARTITION_LEARNED
                                    : o.partition == MDR_PARTITION_TERMS ? MDR_PARTITION_TERMS
                                                                         : MDR_PARTITION_STATIC;
        }
        stats->U = windowed ? win.U : c->tab.U;
        stats->u_dense = c->tab.ud;
        stats->mu_limit = c->tab.mu;
        stats->W = c->tab.L;
        if (!plans.empty()) {
            const XPlan& p = plans.back();
            stats->u = p.u;
            stats->N = p.N;
            stats->cx = p.cx;
        }
    }
    return MDR_OK;
}

}  // namespace

// ---- Dirichlet cross-check engine -----------------------------------------------------
// M(x) from sum_{d <= v} M(floor(v/d)) = 1, an algorithm independent of the
// Deleglise-Rivat split (the reference's optional dirichlet_engine.hpp,
// bench_dr.cpp:9-12, test_dr.cpp:8-11 and :71-88, is not shipped with it).  With
// t = isqrt(v) and the d grouped by q = floor(v/d) <= t,
//   M(v) = 1 - sum_{2 <= d <= v/(t+1)} M(floor(v/d)) - sum_{q <= t} (floor(v/q) - floor(v/(q+1))) M(q).
// M(q) for q <= L comes from a table (L about x^(2/3)); the large values
// v = floor(x/k), k <= K = x/(L+1), are computed from large k down: M(x/k)
// needs M(x/(k d)) only for k d <= K, so every k in (K/2^(j+1), K/2^j] can run
// in parallel once the levels above it are done (one workgroup per k).

__global__ __launch_bounds__(256) void k_dirichlet_level(TableParam tp, i64 x, i64 klo, i64 khi, i64* big) {
    __shared__ i64 red[4];
    const i64 k = khi - (i64)blockIdx.x;
    if (k <= klo) return;
    const i64 v = div_exact(x, k);
    i64 t = (i64)__builtin_sqrt((double)v);
    while ((u64)t * (u64)t > (u64)v) --t;
    while ((u64)(t + 1) * (u64)(t + 1) <= (u64)v) ++t;
    const i64 n1 = div_exact(v, t + 1);
    const double dv = (double)v;
    i64 s = 0;
    for (i64 d = 2 + (i64)threadIdx.x; d <= n1; d += blockDim.x) {
        const i64 q = (i64)qdiv_big((u64)v, dv, (u64)d, rcp_fast((double)d));
        s += q <= tp.U ? (i64)get_M(tp, q) : big[k * d];
    }
    for (i64 q = 1 + (i64)threadIdx.x; q <= t; q += blockDim.x) {
        const i64 a = (i64)qdiv_big((u64)v, dv, (u64)q, rcp_fast((double)q));
        const i64 b = (i64)qdiv_big((u64)v, dv, (u64)(q + 1), rcp_fast((double)(q + 1)));
        s += (a - b) * (i64)get_M(tp, q);
    }
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) big[k] = 1 - (red[0] + red[1] + red[2] + red[3]);
}

__global__ void k_dirichlet_small(TableParam tp, i64 x, i64* big) {
    if (blockIdx.x == 0 && threadIdx.x == 0) big[1] = get_M(tp, x);
}

mdr_status dirichlet_mertens(mdr_ctx* c, i64 X, long long* out, mdr_stats* stats) {
    if (X < 1) { *out = 0; return MDR_OK; }
    // the large values take about 4 x / sqrt(L) term evaluations (about 1.5e11
    // per second): a cross-check engine, limited to X <= 1e18
    if (X > 1000000000000000000LL) return fail(c, MDR_ERR_RANGE, "the Dirichlet engine takes X <= 1e18");
    // table to L ~ x^(2/3) (at least sqrt(x) + 1, at most the one-table cap)
    const i64 r = isqrt_host(X);
    i64 L = (i64)std::ceil(std::pow((double)X, 2.0 / 3.0));
    L = std::max<i64>({L, r + 1, 100000});
    L = std::min<i64>({L, tune_u_cap(), X});
    L = std::max<i64>(L, std::min<i64>(X, r + 1));
    const i64 ud = L <= 1048576 ? L : 1048576;   // a whole number of lines
    mdr_status s = build_table(c, c->tab, L, ud, 1, stats);
    if (s != MDR_OK) return s;
    const TableParam tp = c->tab.tp();
    const i64 K = X / (L + 1);
    HIPT(c->big.ensure((size_t)(K + 2) * 8), "alloc large values");
    i64* big = c->big.as<i64>();
    HIPT(hipEventRecord(c->ev[3], c->st), "event");
    if (K == 0) {
        hipLaunchKernelGGL(k_dirichlet_small, dim3(1), dim3(64), 0, c->st, tp, X, big);
        HIPT(hipGetLastError(), "launch dirichlet");
    }
    for (i64 khi = K; khi >= 1; khi /= 2) {
        const i64 klo = khi / 2;
        hipLaunchKernelGGL(k_dirichlet_level, dim3((unsigned)(khi - klo)), dim3(256), 0, c->st, tp, X, klo, khi, big);
        HIPT(hipGetLastError(), "launch dirichlet level");
    }
    HIPT(hipEventRecord(c->ev[4], c->st), "event");
    long long m = 0;
    HIPT(d2h(c, &m, big + 1, 8), "download");
    *out = m;
    if (stats) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev[3], c->ev[4]);
        stats->quotient_ms += ms;
        stats->kernel_time = (stats->table_ms + stats->quotient_ms) / 1000.0;
        stats->U = L;
        stats->u = L;
        stats->N = K;
    }
    return MDR_OK;
}

// ---- allocation-free entry point ----------------------------------------------------
// mdr_mertens_noalloc computes M(X) with no device allocation: every device
// buffer of the call -- the stencil and sieving primes (generated by kernels,
// no upload), the table, the partition arrays, the counters and the result --
// is carved out of one file-scope device array of NA_SCRATCH_BYTES.  Each GPU
// holds its own copy of the array (loaded with the code object, held until the
// process ends); its address on a device is read once, with that device
// current, by mdr_init (hipGetSymbolAddress) and kept per device, and only
// that address is handed to kernels.  Calls on one device share the array:
// they are serialised by a per-device lock inside this process, and calls
// from several processes each use their own copy.
namespace {

constexpr size_t NA_SCRATCH_BYTES = (size_t)256 << 20;
constexpr int NA_MAX_DEVICES = 64;
// sieving primes up to this root (one workgroup sieves them in LDS), so the
// table of a call stays below NA_MAX_ROOT^2 values
constexpr i64 NA_MAX_ROOT = 65535;

static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

void* g_na_ptr[NA_MAX_DEVICES] = {};
std::mutex g_na_mu[NA_MAX_DEVICES];
std::mutex g_na_init_mu;

// Reads this device's address of the scratch array (the device is current).
hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_na_init_mu);
    if (g_na_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_na_ptr[device] = ptr;
    return hipSuccess;
}

// The periodic stencil (weights of 2, 3, 5, 7, 11 and the square flags of 4,
// 9, 25) for n = x + 1, as ensure_primes builds it on the host.
__global__ void k_na_stencil(uint8_t* st, i64 n) {
    const i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const unsigned v = (unsigned)(x % STENCIL_PERIOD) + 1;
    unsigned a = 0;
    if (v % 2 == 0) a += 1;    // bit_width(p - 1) | 1 for p = 2, 3, 5, 7, 11
    if (v % 3 == 0) a += 3;
    if (v % 5 == 0) a += 3;
    if (v % 7 == 0) a += 3;
    if (v % 11 == 0) a += 5;
    if (v % 4 == 0 || v % 9 == 0 || v % 25 == 0) a |= 0x80u;
    st[x] = (uint8_t)a;
}

// The weights of 13, 17, 19 and 23 for n = x + 1, as ensure_primes builds them.
__global__ void k_na_pattern(uint8_t* pt, i64 n) {
    const i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) pt[x] = (uint8_t)pattern_byte((unsigned)(x % PATTERN_PERIOD) + 1);
}

// The weights of 29, 31 and 37 for n = x + 1, as ensure_primes builds them.
__global__ void k_na_pattern3(uint8_t* pt, i64 n) {
    const i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) pt[x] = (uint8_t)pattern3_byte((unsigned)(x % PATTERN3_PERIOD) + 1);
}

// The sieving primes up to root (< 65536) in the order ensure_primes uploads
// them: primes >= 13 with 1/p and their byte weights (and from 256 on, at
// index k - n_small, their large-prime constants, followed by PRIME_RECORD_PAD
// end records), primes >= 7 for the square flags.  One workgroup: an LDS
// sieve of root + 1 flags, then an ordered compaction, 1024 numbers per step.
__global__ __launch_bounds__(1024) void k_na_primes(int root, int n_small, uint32_t* wp, double* wr, uint8_t* ww,
                                                    PrimeConsts* wl, uint32_t* sq) {
    __shared__ uint8_t comp[NA_MAX_ROOT + 1];
    __shared__ int cnt_w[1024 / 64], cnt_s[1024 / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i <= root; i += 1024) comp[i] = i < 2 ? 1 : 0;
    __syncthreads();
    for (int p = 2; p * p <= root; ++p) {
        if (comp[p]) continue;   // read by every thread after the previous barrier
        for (int m = p * p + tid * p; m <= root; m += 1024 * p) comp[m] = 1;
        __syncthreads();
    }
    int base_w = 0, base_s = 0;
    for (int b = 0; b <= root; b += 1024) {
        const int n = b + tid;
        const bool pr = n <= root && !comp[n];
        const bool in_w = pr && n >= 13, in_s = pr && n >= 7;
        const u64 mw = __ballot(in_w), ms = __ballot(in_s);
        const u64 below = (lane ? ~0ull >> (64 - lane) : 0ull);
        if (lane == 0) {
            cnt_w[wave] = __popcll(mw);
            cnt_s[wave] = __popcll(ms);
        }
        __syncthreads();
        int ow = base_w, os = base_s, tw = 0, ts = 0;
        for (int w = 0; w < 1024 / 64; ++w) {
            if (w < wave) { ow += cnt_w[w]; os += cnt_s[w]; }
            tw += cnt_w[w];
            ts += cnt_s[w];
        }
        if (in_w) {
            const int k = ow + __popcll(mw & below);
            wp[k] = (uint32_t)n;
            wr[k] = 1.0 / (double)n;
            ww[k] = (uint8_t)((32u - (uint32_t)__clz((uint32_t)n - 1u)) | 1u);
            if (n >= 256) wl[k - n_small] = large_prime_consts((uint32_t)n);
        }
        if (in_s) sq[os + __popcll(ms & below)] = (uint32_t)n;
        base_w += tw;
        base_s += ts;
        __syncthreads();
    }
    for (int i = tid; i < PRIME_RECORD_PAD; i += 1024) wl[(base_w > n_small ? base_w - n_small : 0) + i] = PRIME_RECORD_END;
}

__global__ void k_na_xparam(XParam xp, XParam* dst) {
    if (blockIdx.x == 0 && threadIdx.x == 0) dst[0] = xp;
}

// M(q) of the table into a 128-bit result (an X read from the table).
__global__ void k_na_read(TableParam tp, i64 q, u64* out128) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const i64 v = get_M(tp, q);
        out128[0] = (u64)v;
        out128[1] = (u64)(v >> 63);
    }
}

// Byte offsets of one call's arrays in the scratch array.
struct NaLayout {
    size_t stencil = 0, pattern = 0, pattern3 = 0, wp = 0, wr = 0, ww = 0, wl = 0, sq = 0, Md = 0, cl = 0, mu6 = 0, segsum = 0,
           scan = 0;
    size_t xparam = 0, len = 0, bounds = 0, ctr = 0, out = 0, end = 0;
};

size_t na_take(size_t& at, size_t bytes) {
    const size_t off = at;
    at += (bytes + 255) & ~(size_t)255;
    return off;
}

// Counts of the sieving primes up to root, as the kernel lays them out.
void na_prime_counts(i64 root, int& n_w, int& n_small, int& n_split, int& n_coop, int& n_sq, int& n_sq_small) {
    uint64_t comp[(NA_MAX_ROOT + 64) / 64] = {};
    n_w = n_small = n_split = n_coop = n_sq = n_sq_small = 0;
    for (i64 i = 2; i <= root; ++i) {
        if (comp[i >> 6] >> (i & 63) & 1) continue;
        for (i64 m = i * i; m <= root; m += i) comp[m >> 6] |= 1ull << (m & 63);
        if (i >= 13) {
            ++n_w;
            if (i < 256) ++n_small;
            if (i < SQ_SPLIT) ++n_split;
            if (i < COOP_PRIME_LIMIT) ++n_coop;
        }
        if (i >= 7) { ++n_sq; if (i < 256) ++n_sq_small; }
    }
}

struct NaPlan {
    int kind = 0;       // 0 constant, 1 table lookup, 2 Deleglise-Rivat sum
    i64 value = 0;
    i64 U = 0, ud = 0, mu_lim = 0, u = 0, N = 0, T = 0, slots = 0, P = 0, nseg = 0, musegs = 0, L = 0;
    double cx = 0;
    int shape = 0;
    NaLayout lay;
};

// Lays out a plan's arrays; false when they do not fit the scratch array.
bool na_layout(NaPlan& q, i64 nwaves) {
    const SieveShape& sh = SIEVE_SHAPES[q.shape];
    q.nseg = (q.U + sh.seg - 1) / sh.seg;
    q.musegs = (q.mu_lim + sh.seg - 1) / sh.seg;
    q.L = q.U > q.ud ? (q.U - q.ud + LINE_VALUES - 1) / LINE_VALUES : 0;
    const i64 root = isqrt_host(q.U) + 1;
    if (root > NA_MAX_ROOT) return false;
    NaLayout& l = q.lay;
    size_t at = 0;
    l.stencil = na_take(at, (size_t)STENCIL_PERIOD + SEG_MAX);
    l.pattern = na_take(at, (size_t)PATTERN_REPEAT * PATTERN_PERIOD + SEG_MAX);
    l.pattern3 = na_take(at, (size_t)PATTERN_REPEAT * PATTERN3_PERIOD + SEG_MAX);
    l.wp = na_take(at, (size_t)root * 4 + 4);
    l.wr = na_take(at, (size_t)root * 8 + 8);
    l.ww = na_take(at, (size_t)root + 4);
    l.wl = na_take(at, ((size_t)root / 2 + PRIME_RECORD_PAD) * sizeof(PrimeConsts));   // pi(root) < root / 2
    l.sq = na_take(at, (size_t)root * 4 + 4);
    l.Md = na_take(at, (size_t)(q.ud + 1) * 2);
    l.cl = na_take(at, (size_t)q.L * sizeof(CLine));
    l.mu6 = na_take(at, (size_t)q.musegs * (sh.seg / 48) * 4);
    l.segsum = na_take(at, (size_t)q.nseg * 8);
    q.P = q.kind == 2 ? std::max<i64>(1, nwaves * 64) : 0;
    l.scan = na_take(at, (size_t)std::max(scan_tmp_len(q.nseg), scan_tmp_len(q.slots)) * 8);
    l.xparam = na_take(at, sizeof(XParam));
    l.len = na_take(at, (size_t)std::max<i64>(q.slots, 1) * 8);
    l.bounds = na_take(at, (size_t)(q.P + 1) * 8);
    l.ctr = na_take(at, 64);
    l.out = na_take(at, 16);
    l.end = at;
    return at <= NA_SCRATCH_BYTES;
}

// The plan of an allocation-free call: the tuned split, with the cutoff
// lowered (M(X) does not depend on it) until the call's arrays fit.
mdr_status na_plan(mdr_ctx* c, i64 X, NaPlan& q, i64 nwaves) {
    mdr_opts o;
    mdr_default_opts(&o);
    o.params = MDR_PARAMS_TUNED;
    const XPlan p = plan_x(X, o);
    q = NaPlan{};
    q.kind = p.kind;
    q.value = p.value;
    if (p.kind == 0) return MDR_OK;
    q.shape = tune_sieve_shape(X);
    if (p.kind == 1) {
        q.U = q.ud = q.mu_lim = std::max<i64>(p.X, 1);
        if (!na_layout(q, nwaves))
            return fail(c, MDR_ERR_RANGE, "X beyond what the allocation-free call's scratch array holds");
        return MDR_OK;
    }
    q.cx = p.cx;
    const i64 A2 = (i64)(p.cx * std::sqrt((double)X)) + 1000;
    for (i64 u = p.u;; u = u / 5 * 4) {
        u = std::max<i64>(u / LINE_VALUES * LINE_VALUES, LINE_VALUES);
        q.u = u;
        q.N = X / u;
        q.mu_lim = std::max(A2, q.N + 1000);
        q.U = std::max(u, q.mu_lim);
        q.ud = std::min<i64>(std::min(p.ud, u), q.U);
        if (q.ud < q.U) q.ud = q.ud / LINE_VALUES * LINE_VALUES;
        q.T = count_coprime6(q.N);
        q.slots = 2 * q.T;
        if (X <= q.U) {   // the lowered cutoff reaches X: a table read
            q.kind = 1;
            q.U = q.ud = q.mu_lim = X;
            q.slots = 0;
        }
        if (na_layout(q, nwaves)) return MDR_OK;
        if (u <= A2) break;   // the split's own limit: lowering u further only adds slots
    }
    return fail(c, MDR_ERR_RANGE,
                "X beyond what the allocation-free call's scratch array holds (" +
                    std::to_string(NA_SCRATCH_BYTES >> 20) + " MiB; about X <= 8e15)");
}

}  // namespace

// ---- C ABI --------------------------------------------------------------------------
extern "C" {

void mdr_default_opts(mdr_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->strategy = MDR_STRATEGY_AUTO;
    o->partition = MDR_PARTITION_STATIC;
    o->params = MDR_PARAMS_REFERENCE;
    o->sched = MDR_SCHED_DYNAMIC;
}

mdr_status mdr_init(mdr_ctx** out, int device) {
    if (!out) return MDR_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MDR_ERR_NODEVICE;
    if (device < 0 || device >= n) return MDR_ERR_NODEVICE;
    mdr_ctx* c = new mdr_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    for (int i = 0; i < 8 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->clock_khz, hipDeviceAttributeClockRate, device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device);
    int occ = 0;
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_quot, 256, 0);
    if (e != hipSuccess) {
        mdr_shutdown(c);
        return MDR_ERR_HIP;
    }
    c->quot_blocks_per_cu = std::max(1, occ);
    // the allocation-free entry point: this device's address of its scratch
    // array, and a pinned word pair for its result
    if (noalloc_setup(device) != hipSuccess || hipHostMalloc(&c->na_host, 64, hipHostMallocDefault) != hipSuccess) {
        c->na_host = nullptr;
        mdr_shutdown(c);
        return MDR_ERR_HIP;
    }
    size_t mem_free = 0, mem_total = 0;
    if (hipMemGetInfo(&mem_free, &mem_total) == hipSuccess && mem_total > 0) tune_set_device_memory(mem_total);
    // the trained cost model: MDR_COST_MODEL, else cost_model.bin beside the library
    std::string path;
    if (const char* e = std::getenv("MDR_COST_MODEL")) path = e;
    else {
        Dl_info info;
        if (dladdr((void*)&mdr_init, &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/cost_model.bin";
        }
    }
    if (!path.empty()) (void)mdr_cost_model_load(c, path.c_str());   // absent or invalid: static partition
    c->err.clear();
    *out = c;
    return MDR_OK;
}

void mdr_shutdown(mdr_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->st) (void)hipStreamSynchronize(c->st);
    DevBuf* bufs[] = {&c->tab.Md, &c->tab.cl, &c->tab.mu6, &c->tab.segsum, &c->scan_tmp, &c->stencil, &c->pattern, &c->pattern3,
                      &c->wprimes, &c->wrcp, &c->wweights, &c->wlarge, &c->sqprimes, &c->xparam, &c->slot_len, &c->slot_cost, &c->bounds,
                      &c->out128, &c->ctr, &c->stamps, &c->comp_out, &c->lookup, &c->mlp,
                      &c->wcl, &c->wseg, &c->wscan, &c->wcarry, &c->big, &c->wspan};
    for (auto& g : c->graph_exec)
        if (g) (void)hipGraphExecDestroy(g);
    for (DevBuf* b : bufs) b->release();
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->na_host) (void)hipHostFree(c->na_host);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

const char* mdr_last_error(mdr_ctx* c) { return c ? c->err.c_str() : "null context"; }

long long mdr_quot_waves(mdr_ctx* c) { return c ? (long long)c->ncu * c->quot_blocks_per_cu * 4 : 0; }

mdr_status mdr_plan(mdr_ctx* c, long long X, const mdr_opts* opts, int* streaming, long long* u,
                    long long* u_dense, long long* mu_limit, long long* N, double* cx) {
    // Pure host logic: ctx may be NULL (then errors carry no message).
    mdr_opts o;
    mdr_default_opts(&o);
    if (opts) o = *opts;
    mdr_ctx scratch;
    mdr_ctx* cc = c ? c : &scratch;
    std::unique_lock<std::mutex> lk(cc->mu);
    mdr_status s = validate_opts(cc, o);
    if (s != MDR_OK) return s;
    const XPlan p = plan_x(X, o);
    if (streaming) *streaming = p.streaming;
    if (u) *u = p.u;
    if (u_dense) *u_dense = p.ud;
    if (mu_limit) *mu_limit = p.mu_lim;
    if (N) *N = p.N;
    if (cx) *cx = p.cx;
    return MDR_OK;
}

double mdr_param_cx(long long X, const mdr_opts* opts) {
    mdr_opts o;
    mdr_default_opts(&o);
    if (opts) o = *opts;
    return ref_cx(X, o);
}

long long mdr_param_dense_cutoff(long long X, const mdr_opts* opts) {
    mdr_opts o;
    mdr_default_opts(&o);
    if (opts) o = *opts;
    return ref_dense_u(X, o);
}

long long mdr_param_streaming_cutoff(long long X, const mdr_opts* opts) {
    mdr_opts o;
    mdr_default_opts(&o);
    if (opts) o = *opts;
    return ref_stream_u(X, o, ref_cx(X, o));
}

mdr_status mdr_mertens_batch(mdr_ctx* c, const long long* X, size_t n, const mdr_opts* opts, long long* out,
                             mdr_stats* stats) {
    if (!c || (!X && n) || (!out && n)) return MDR_ERR_INVALID;
    mdr_opts o;
    mdr_default_opts(&o);
    if (opts) o = *opts;
    std::lock_guard<std::mutex> lk(c->mu);
    c->err.clear();
    mdr_status s = validate_opts(c, o);
    if (s != MDR_OK) return s;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const auto t0 = std::chrono::steady_clock::now();
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MDR_ERR_HIP, "hipSetDevice failed");
    std::vector<XPlan> plans(n);
    for (size_t i = 0; i < n; ++i) plans[i] = plan_x(X[i], o);
    share_batch_cutoff(plans, o);
    s = run_batch(c, plans, o, out, stats);
    if (stats) stats->total_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return s;
}

mdr_status mdr_mertens(mdr_ctx* c, long long X, const mdr_opts* opts, long long* out, mdr_stats* stats) {
    return mdr_mertens_batch(c, &X, 1, opts, out, stats);
}

mdr_status mdr_mertens_noalloc(mdr_ctx* c, long long X, long long* out, mdr_stats* stats) {
    if (!c || !out) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    c->err.clear();
    if (stats) *stats = mdr_stats{};
    const auto t0 = std::chrono::steady_clock::now();
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MDR_ERR_HIP, "hipSetDevice failed");
    if (c->device >= NA_MAX_DEVICES || !g_na_ptr[c->device] || !c->na_host)
        return fail(c, MDR_ERR_HIP, "allocation-free scratch array not set up on this device");
    std::lock_guard<std::mutex> dlk(g_na_mu[c->device]);
    const i64 nwaves = (i64)c->ncu * c->quot_blocks_per_cu * 4;
    NaPlan q;
    mdr_status s = na_plan(c, X, q, nwaves);
    if (s != MDR_OK) return s;
    if (q.kind == 0) {
        *out = q.value;
        return MDR_OK;
    }
    unsigned char* sc = static_cast<unsigned char*>(g_na_ptr[c->device]);
    const NaLayout& l = q.lay;
    const SieveShape& sh = SIEVE_SHAPES[q.shape];
    // stencil and sieving primes, generated on the device
    const i64 nst = (i64)STENCIL_PERIOD + SEG_MAX;
    const i64 root = isqrt_host(q.U) + 1;
    int n_w = 0, n_small = 0, n_split = 0, n_coop = 0, n_sq = 0, n_sq_small = 0;
    na_prime_counts(root, n_w, n_small, n_split, n_coop, n_sq, n_sq_small);
    HIPT(hipEventRecord(c->ev[0], c->st), "event");
    hipLaunchKernelGGL(k_na_stencil, dim3((unsigned)((nst + 255) / 256)), dim3(256), 0, c->st, sc + l.stencil, nst);
    const i64 npt = (i64)PATTERN_REPEAT * PATTERN_PERIOD + SEG_MAX;
    hipLaunchKernelGGL(k_na_pattern, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, c->st, sc + l.pattern, npt);
    const i64 npt3 = (i64)PATTERN_REPEAT * PATTERN3_PERIOD + SEG_MAX;
    hipLaunchKernelGGL(k_na_pattern3, dim3((unsigned)((npt3 + 255) / 256)), dim3(256), 0, c->st, sc + l.pattern3,
                       npt3);
    hipLaunchKernelGGL(k_na_primes, dim3(1), dim3(1024), 0, c->st, (int)root, n_small, (uint32_t*)(sc + l.wp),
                       (double*)(sc + l.wr), sc + l.ww, (PrimeConsts*)(sc + l.wl), (uint32_t*)(sc + l.sq));
    HIPT(hipGetLastError(), "launch prime setup");
    // the table
    SieveArgs a;
    a.U = q.U;
    a.ud = q.ud;
    a.L = q.L;
    a.seg_first = 0;
    a.mu_segs = q.musegs;
    a.stencil = sc + l.stencil;
    a.pattern = sc + l.pattern;
    a.pattern3 = sc + l.pattern3;
    a.wprimes = (const uint32_t*)(sc + l.wp);
    a.wrcp = (const double*)(sc + l.wr);
    a.wweights = sc + l.ww;
    a.n_w = n_w;
    a.n_small = n_small;
    a.wlarge = (const PrimeConsts*)(sc + l.wl);
    a.n_split = n_split;
    a.n_coop = n_coop;
    a.sqprimes = (const uint32_t*)(sc + l.sq);
    a.n_sq = n_sq;
    a.n_sq_small = n_sq_small;
    a.Md = (int16_t*)(sc + l.Md);
    a.cl = (CLine*)(sc + l.cl);
    a.mu6 = (uint32_t*)(sc + l.mu6);
    a.segsum = (i64*)(sc + l.segsum);
    a.seg0 = 0;
    HIPT(launch_sieve(q.shape, (unsigned)q.nseg, c->st, a), "launch sieve");
    i64* scan = (i64*)(sc + l.scan);
    const i64 scan_len = (i64)((l.xparam - l.scan) / 8);
    HIPT(scan_exclusive_i64(a.segsum, a.segsum, q.nseg, scan, scan_len, nullptr, c->st), "scan segments");
    FixupArgs f;
    f.ud = q.ud;
    f.L = q.L;
    f.nd = q.ud;
    f.seg_shift = sh.shift;
    f.Md = a.Md;
    f.cl = a.cl;
    f.segoff = a.segsum;
    f.seg0 = 0;
    f.carry = nullptr;
    const i64 nfix = std::max<i64>(1, std::max(q.ud, q.L));
    hipLaunchKernelGGL(k_table_fixup, dim3((unsigned)std::min<i64>((nfix + 255) / 256, 1 << 20)), dim3(256), 0, c->st,
                       f);
    HIPT(hipGetLastError(), "launch fixup");
    HIPT(hipEventRecord(c->ev[1], c->st), "event");
    TableParam tp;
    tp.U = q.U;
    tp.ud = q.ud;
    tp.L = q.L;
    tp.mu_lim = q.mu_lim;
    tp.Md = a.Md;
    tp.cl = a.cl;
    tp.mu6 = a.mu6;
    tp.pass = 0;
    tp.qlo = q.ud + 1;
    tp.qhi = q.U;
    u64* out128 = (u64*)(sc + l.out);
    u64* ctr = (u64*)(sc + l.ctr);
    hipLaunchKernelGGL(k_clear3, dim3(1), dim3(64), 0, c->st, ctr, (i64)8, out128, (i64)2, (u64*)nullptr, (i64)0);
    HIPT(hipGetLastError(), "launch clear");
    if (q.kind == 1) {
        hipLaunchKernelGGL(k_na_read, dim3(1), dim3(64), 0, c->st, tp, (i64)X, out128);
        HIPT(hipGetLastError(), "launch read");
    } else {
        XParam xp{};
        xp.X = X;
        xp.u = q.u;
        xp.N = q.N;
        xp.T = q.T;
        xp.slot_base = 0;
        xp.cx = q.cx;
        xp.n6 = q.N / 6;
        xp.n3 = q.N / 3;
        xp.n2 = q.N / 2;
        XParam* dxp = (XParam*)(sc + l.xparam);
        hipLaunchKernelGGL(k_na_xparam, dim3(1), dim3(64), 0, c->st, xp, dxp);
        PartArgs pa{};
        pa.xp = dxp;
        pa.nx = 1;
        pa.total_slots = q.slots;
        pa.tp = tp;
        pa.mlp = nullptr;
        pa.len = (i64*)(sc + l.len);
        pa.cost = nullptr;
        pa.type_pos = (unsigned long long*)(ctr + 1);
        pa.totals = (const i64*)(ctr + 4);
        pa.P = q.P;
        pa.bounds = (i64*)(sc + l.bounds);
        pa.min_chunk = MIN_CHUNK_POSITIONS;
        pa.by_slot = 0;
        hipLaunchKernelGGL(k_slot_len, dim3((unsigned)((q.slots + 255) / 256)), dim3(256), 0, c->st, pa);
        HIPT(hipGetLastError(), "launch slot lengths");
        HIPT(scan_exclusive_i64(pa.len, pa.len, q.slots, scan, scan_len, (i64*)(ctr + 4), c->st), "scan positions");
        hipLaunchKernelGGL(k_cuts, dim3((unsigned)((q.P + 1 + 255) / 256)), dim3(256), 0, c->st, pa);
        HIPT(hipGetLastError(), "launch cuts");
        QuotArgs qa{};
        qa.xp = dxp;
        qa.nx = 1;
        qa.tp = tp;
        qa.off = pa.len;
        qa.nslots = q.slots;
        qa.bounds = pa.bounds;
        qa.P = q.P;
        qa.next = (unsigned long long*)ctr;
        qa.out128 = out128;
        qa.dynamic = 1;
        qa.stamps = nullptr;
        qa.nstamps = (unsigned long long*)(ctr + 6);
        qa.stamp_cap = 0;
        qa.stamp_stride = 1;
        qa.wave_span = nullptr;
        HIPT(hipEventRecord(c->ev[2], c->st), "event");
        hipLaunchKernelGGL(k_quot, dim3((unsigned)(c->ncu * c->quot_blocks_per_cu)), dim3(256), 0, c->st, qa);
        HIPT(hipGetLastError(), "launch quotient kernel");
        HIPT(hipEventRecord(c->ev[3], c->st), "event");
    }
    HIPT(hipMemcpyAsync(c->na_host, out128, 16, hipMemcpyDeviceToHost, c->st), "download");
    HIPT(hipStreamSynchronize(c->st), "allocation-free call");
    const u64* r = static_cast<const u64*>(c->na_host);
    const i64 lo = (i64)r[0], hi = (i64)r[1];
    if (hi != (lo >> 63)) return fail(c, MDR_ERR_RANGE, "result does not fit in 64 bits");
    *out = lo;
    if (stats) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
        stats->table_ms = ms;
        if (q.kind == 2) {
            (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
            stats->quotient_ms = ms;
        }
        stats->U = q.U;
        stats->u_dense = q.ud;
        stats->mu_limit = q.mu_lim;
        stats->W = q.L;
        stats->u = q.u;
        stats->N = q.N;
        stats->slots = q.slots;
        stats->chunks = q.P;
        stats->cx = q.cx;
        stats->partition_used = MDR_PARTITION_STATIC;
        stats->kernel_time = (stats->table_ms + stats->quotient_ms) / 1000.0;
        stats->total_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return MDR_OK;
}

mdr_status mdr_mertens_dirichlet(mdr_ctx* c, long long X, long long* out, mdr_stats* stats) {
    if (!c || !out) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    c->err.clear();
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MDR_ERR_HIP, "hipSetDevice failed");
    if (stats) *stats = mdr_stats{};
    const auto t0 = std::chrono::steady_clock::now();
    mdr_status s = dirichlet_mertens(c, X, out, stats);
    if (stats) stats->total_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return s;
}

mdr_status mdr_table_build(mdr_ctx* c, long long U, long long u_dense, long long mu_limit) {
    if (!c || U < 1 || u_dense < 0 || mu_limit < 0) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    c->err.clear();
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MDR_ERR_HIP, "hipSetDevice failed");
    return build_table(c, c->tab, U, u_dense, mu_limit, nullptr);
}

}  // extern "C"

// component kernels ------------------------------------------------------------------
namespace {

__global__ void k_read_M(TableParam tp, i64 first, i64 count, int32_t* out) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = get_M(tp, first + i);
}
__global__ void k_read_mu(TableParam tp, i64 first, i64 count, int8_t* out) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const i64 n = first + i;
    int8_t m = 0;
    if (n >= 1 && n <= tp.mu_lim) {
        // mu(n) = mu(odd part) * (-1) when 2 || n, 0 when 4 | n; likewise for 3.
        i64 v = n;
        int sgn = 1;
        bool zero = false;
        if (v % 4 == 0) zero = true;
        else if (v % 2 == 0) { v /= 2; sgn = -sgn; }
        if (v % 9 == 0) zero = true;
        else if (v % 3 == 0) { v /= 3; sgn = -sgn; }
        if (!zero) m = (int8_t)(sgn * mu6_mu(tp.mu6, v / 3));
    }
    out[i] = m;
}

__global__ void k_component(TableParam tp, Seg sg, u64 y, int geval, int comb2, i64 u, double cx, u64* out128) {
    const int lane = threadIdx.x & 63;
    const i64 wave = ((i64)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const i64 nw = ((i64)gridDim.x * blockDim.x) >> 6;
    u64 lo = 0;
    i64 hi = 0;
    if (!geval) {
        const i64 L = sg.count;
        const i64 a = (i64)((u128)wave * (u64)L / (u64)nw), b = (i64)((u128)(wave + 1) * (u64)L / (u64)nw);
        if (a < b) {
            const i64 v = run_seg(tp, sg, y, (u64)a, (u64)b, lane);
            add128(lo, hi, (u64)v, v >> 63);
        }
    } else {
        XParam xp{};
        xp.u = u;
        Frame f{};
        f.y = (i64)y;
        f.coef = 1;
        f.comb2 = comb2;
        f.A = split_point((i64)y, cx, u);
        f.ky = (i64)y / (f.A + 1);
        if (comb2) {
            f.B = split_point((i64)y >> 1, cx, u);
            f.ky2 = ((i64)y >> 1) / (f.B + 1);
        }
        i64 L = 0;
        const int ns = frame_nseg(f);
        for (int i = 0; i < ns; ++i) L += frame_seg(f, xp, tp, i).count;
        const i64 a = (i64)((u128)wave * (u64)L / (u64)nw), b = (i64)((u128)(wave + 1) * (u64)L / (u64)nw);
        i64 s0 = 0;
        for (int i = 0; i < ns; ++i) {
            const Seg s = frame_seg(f, xp, tp, i);
            const i64 s1 = s0 + s.count;
            const i64 l = a > s0 ? a : s0, h = b < s1 ? b : s1;
            if (l < h) {
                const i64 v = s.sign * run_seg(tp, s, y, (u64)(l - s0), (u64)(h - s0), lane);
                add128(lo, hi, (u64)v, v >> 63);
            }
            s0 = s1;
        }
        if (wave == 0 && lane == 0) {
            const i128 cst = frame_const(f, tp);
            add128(lo, hi, (u64)cst, (i64)(cst >> 64));
        }
    }
    wave_reduce128(lo, hi);
    if (lane == 0) atomic_add128(out128, lo, hi);
}

// out[i] = M(q[i]) (gathers for table objects)
__global__ void k_gather_M(TableParam tp, const i64* q, i64 n, int32_t* out) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = get_M(tp, q[i]);
}

// The compressed tier in the reference's layout (mertens_dr.hpp:670-687,
// :741-752) for a dense limit ref_ud at or above this table's own (ud): one
// thread per 64-value word w of chunk c0 + w / 4, covering n = ref_ud + 1 +
// 64 w' ...; the word's bits are this table's plane bits at offset
// n - ud - 1 (one or two of its 64-value groups, expanded from their kept
// bits), its checkpoint M(n - 1).  Words past lim read as zero with a zero
// checkpoint.
__global__ void k_export_chunks(TableParam tp, i64 lim, i64 ref_ud, i64 c0, i64 nwords, u64* chunks, int32_t* ck) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nwords) return;
    const i64 w = 4 * c0 + t;                  // word of the reference tier
    const i64 n0 = ref_ud + 1 + 64 * w;
    u64 nz = 0, ng = 0;
    int32_t m0 = 0;
    if (n0 <= lim) {
        m0 = get_M(tp, n0 - 1);
        const u64 d = (u64)(n0 - tp.ud - 1);   // >= 0: ref_ud >= ud
        const u64 gw = d >> 6;
        const uint32_t b = (uint32_t)(d & 63);
        u64 z0, g0;
        raw_planes(tp.cl, gw, z0, g0);
        nz = z0 >> b;
        ng = g0 >> b;
        // the next group holds values only when it starts at or below U
        if (b && (i64)(64 * (gw + 1)) < tp.U - tp.ud) {
            u64 z1, g1;
            raw_planes(tp.cl, gw + 1, z1, g1);
            nz |= z1 << (64 - b);
            ng |= g1 << (64 - b);
        }
        const i64 cnt = lim - n0 + 1;
        if (cnt < 64) {
            const u64 m = (1ull << cnt) - 1;
            nz &= m;
            ng &= m;
        }
    }
    // chunk layout: nonzero bytes of words 0..3, then sign bytes of words 0..3
    u64* ch = chunks + 8 * (t >> 2);
    ch[t & 3] = nz;
    ch[4 + (t & 3)] = ng;
    ck[t] = m0;
}

mdr_status run_component(mdr_ctx* c, const Seg& sg, i64 y, int geval, int comb2, i64 u, double cx, long long* lo,
                         long long* hi) {
    if (!c->tab.valid) return fail(c, MDR_ERR_NOTABLE, "no table built");
    HIPT(c->comp_out.ensure(64), "alloc");
    HIPT(hipMemsetAsync(c->comp_out.p, 0, 16, c->st), "clear");
    hipLaunchKernelGGL(k_component, dim3(64), dim3(256), 0, c->st, c->tab.tp(), sg, (u64)y, geval, comb2, u, cx,
                       c->comp_out.as<u64>());
    HIPT(hipGetLastError(), "launch component");
    u64 r[2];
    HIPT(hipMemcpyAsync(r, c->comp_out.p, 16, hipMemcpyDeviceToHost, c->st), "download");
    HIPT(hipStreamSynchronize(c->st), "component");
    *lo = (long long)r[0];
    *hi = (long long)r[1];
    return MDR_OK;
}

}  // namespace

extern "C" {

// One dispatch holds fewer than 2^32 work-items: reads and model batches are
// limited to 2^31 entries per call.
constexpr long long MAX_CALL_ITEMS = 1LL << 31;

}  // extern "C"

namespace {

// Reads of one table into host memory, through a device staging buffer of
// at most 2^27 entries per dispatch (the caller holds the context's lock).
constexpr long long READ_STAGE = 1LL << 27;

mdr_status read_M(mdr_ctx* c, const TableParam& tp, i64 first, i64 count, int32_t* out) {
    if (count == 0) return MDR_OK;
    DevBuf tmp;
    HIPT(tmp.ensure((size_t)std::min(count, READ_STAGE) * 4), "alloc");
    hipError_t e = hipSuccess;
    for (i64 k = 0; k < count && e == hipSuccess; k += READ_STAGE) {
        const i64 m = std::min(READ_STAGE, count - k);
        hipLaunchKernelGGL(k_read_M, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->st, tp, first + k, m,
                           tmp.as<int32_t>());
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(out + k, tmp.p, (size_t)m * 4, hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    }
    tmp.release();
    if (e != hipSuccess) return hip_fail(c, e, "read M");
    return MDR_OK;
}

mdr_status read_mu(mdr_ctx* c, const TableParam& tp, i64 first, i64 count, int8_t* out) {
    if (count == 0) return MDR_OK;
    DevBuf tmp;
    HIPT(tmp.ensure((size_t)std::min(count, READ_STAGE)), "alloc");
    hipError_t e = hipSuccess;
    for (i64 k = 0; k < count && e == hipSuccess; k += READ_STAGE) {
        const i64 m = std::min(READ_STAGE, count - k);
        hipLaunchKernelGGL(k_read_mu, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->st, tp, first + k, m,
                           tmp.as<int8_t>());
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(out + k, tmp.p, (size_t)m, hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    }
    tmp.release();
    if (e != hipSuccess) return hip_fail(c, e, "read mu");
    return MDR_OK;
}

}  // namespace

// A table object: its own device buffers, built and read on its context's GPU.
struct mdr_table {
    mdr_ctx* c = nullptr;
    TableStore s;
    i64 U = 0, mu = 0;   // as requested (the build may hold one more value)
    TableParam tp() const {
        TableParam t = s.tp();
        t.mu_lim = mu;   // mu reads beyond the requested limit give 0
        return t;
    }
};

extern "C" {

mdr_status mdr_table_read_M(mdr_ctx* c, long long first, long long count, int32_t* out) {
    if (!c || count < 0 || count > MAX_CALL_ITEMS || (count && !out) || first < 0) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->tab.valid) return fail(c, MDR_ERR_NOTABLE, "no table built");
    if (count && (first > c->tab.U || count - 1 > c->tab.U - first))
        return fail(c, MDR_ERR_INVALID, "read beyond the table");
    return read_M(c, c->tab.tp(), first, count, out);
}

mdr_status mdr_table_read_mu(mdr_ctx* c, long long first, long long count, int8_t* out) {
    if (!c || count < 0 || count > MAX_CALL_ITEMS || (count && !out) || first < 0 || first > LLONG_MAX - count)
        return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->tab.valid) return fail(c, MDR_ERR_NOTABLE, "no table built");
    return read_mu(c, c->tab.tp(), first, count, out);
}

mdr_status mdr_table_create(mdr_ctx* c, long long U, long long u_dense, long long mu_limit, mdr_table** out) {
    if (!c || !out || U < 0 || u_dense < 0 || mu_limit < 0 || mu_limit > U) return MDR_ERR_INVALID;
    *out = nullptr;
    std::lock_guard<std::mutex> lk(c->mu);
    c->err.clear();
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MDR_ERR_HIP, "hipSetDevice failed");
    const i64 Ub = std::max<i64>(U, 1);
    i64 ud = std::min<i64>(u_dense, Ub);
    if (ud < Ub) ud = (ud / LINE_VALUES) * LINE_VALUES;
    mdr_table* t = new mdr_table();
    t->c = c;
    t->U = U;
    t->mu = mu_limit;
    const mdr_status s = build_table(c, t->s, Ub, ud, std::max<i64>(mu_limit, 1), nullptr, false, 0);
    if (s != MDR_OK) {
        t->s.release();
        delete t;
        return s;
    }
    *out = t;
    return MDR_OK;
}

void mdr_table_free(mdr_table* t) {
    if (!t) return;
    {
        std::lock_guard<std::mutex> lk(t->c->mu);
        (void)hipSetDevice(t->c->device);
        (void)hipStreamSynchronize(t->c->st);
        t->s.release();
    }
    delete t;
}

mdr_status mdr_table_info(const mdr_table* t, long long* U, long long* u_dense, long long* mu_limit) {
    if (!t) return MDR_ERR_INVALID;
    if (U) *U = t->U;
    if (u_dense) *u_dense = std::min(t->s.ud, t->U);
    if (mu_limit) *mu_limit = t->mu;
    return MDR_OK;
}

mdr_status mdr_table_copy_M(mdr_table* t, long long first, long long count, int32_t* out) {
    if (!t || count < 0 || (count && !out) || first < 0) return MDR_ERR_INVALID;
    mdr_ctx* c = t->c;
    std::lock_guard<std::mutex> lk(c->mu);
    c->err.clear();
    if (count && (first > t->U || count - 1 > t->U - first)) return fail(c, MDR_ERR_INVALID, "read beyond the table");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MDR_ERR_HIP, "hipSetDevice failed");
    return read_M(c, t->tp(), first, count, out);
}

mdr_status mdr_table_copy_mu(mdr_table* t, long long first, long long count, int8_t* out) {
    if (!t || count < 0 || (count && !out) || first < 0 || first > LLONG_MAX - count) return MDR_ERR_INVALID;
    mdr_ctx* c = t->c;
    std::lock_guard<std::mutex> lk(c->mu);
    c->err.clear();
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MDR_ERR_HIP, "hipSetDevice failed");
    return read_mu(c, t->tp(), first, count, out);
}

mdr_status mdr_table_gather_M(mdr_table* t, const long long* q, size_t n, int32_t* out) {
    if (!t || (n && (!q || !out))) return MDR_ERR_INVALID;
    mdr_ctx* c = t->c;
    std::lock_guard<std::mutex> lk(c->mu);
    c->err.clear();
    for (size_t i = 0; i < n; ++i)
        if (q[i] < 0 || q[i] > t->U) return fail(c, MDR_ERR_INVALID, "gather index outside the table");
    if (n == 0) return MDR_OK;
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MDR_ERR_HIP, "hipSetDevice failed");
    const i64 stage = std::min<i64>((i64)n, READ_STAGE);
    DevBuf dq, dout;
    hipError_t e = dq.ensure((size_t)stage * 8);
    if (e == hipSuccess) e = dout.ensure((size_t)stage * 4);
    for (i64 k = 0; k < (i64)n && e == hipSuccess; k += stage) {
        const i64 m = std::min<i64>(stage, (i64)n - k);
        e = hipMemcpyAsync(dq.p, q + k, (size_t)m * 8, hipMemcpyHostToDevice, c->st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_gather_M, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->st, t->tp(),
                               dq.as<i64>(), m, dout.as<int32_t>());
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(out + k, dout.p, (size_t)m * 4, hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    }
    dq.release();
    dout.release();
    if (e != hipSuccess) return hip_fail(c, e, "gather M");
    return MDR_OK;
}

mdr_status mdr_table_copy_chunks(mdr_table* t, long long u_dense_ref, long long first_chunk, long long count,
                                 void* chunks, int32_t* checkpoints) {
    if (!t || first_chunk < 0 || count < 0 || (count && (!chunks || !checkpoints))) return MDR_ERR_INVALID;
    mdr_ctx* c = t->c;
    std::lock_guard<std::mutex> lk(c->mu);
    c->err.clear();
    if (u_dense_ref < t->s.ud || u_dense_ref > t->U)
        return fail(c, MDR_ERR_INVALID, "u_dense_ref outside [the table's dense tier, U]");
    if (first_chunk > (LLONG_MAX - u_dense_ref - 1) / 256 - count)   // values n of the chunks stay below 2^63
        return fail(c, MDR_ERR_INVALID, "chunk range beyond 64-bit values");
    if (count == 0) return MDR_OK;
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MDR_ERR_HIP, "hipSetDevice failed");
    const i64 stage = std::min<i64>(count, READ_STAGE / 4);   // chunks per dispatch
    DevBuf dch, dck;
    hipError_t e = dch.ensure((size_t)stage * 64);
    if (e == hipSuccess) e = dck.ensure((size_t)stage * 16);
    for (i64 k = 0; k < count && e == hipSuccess; k += stage) {
        const i64 m = std::min<i64>(stage, count - k);
        hipLaunchKernelGGL(k_export_chunks, dim3((unsigned)((4 * m + 255) / 256)), dim3(256), 0, c->st, t->tp(), t->U,
                           (i64)u_dense_ref, (i64)(first_chunk + k), 4 * m, dch.as<u64>(), dck.as<int32_t>());
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync((char*)chunks + 64 * k, dch.p, (size_t)m * 64, hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(checkpoints + 4 * k, dck.p, (size_t)m * 16, hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    }
    dch.release();
    dck.release();
    if (e != hipSuccess) return hip_fail(c, e, "copy chunks");
    return MDR_OK;
}

mdr_status mdr_s2_interval(mdr_ctx* c, int piece, long long first, long long last, long long y, long long* lo,
                           long long* hi) {
    if (!c || !lo || !hi || piece < 1 || piece > 12 || y < 0) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->tab.valid) return fail(c, MDR_ERR_NOTABLE, "no table built");
    if (first < 1) first = 1;
    if (last > c->tab.mu) return fail(c, MDR_ERR_INVALID, "interval beyond the mu table");
    Seg sg{};
    sg.type = SEG_S2;
    sg.piece = piece;
    sg.sign = 1;
    if (first <= last) {
        sg.first = count_coprime6(first - 1);
        sg.count = count_coprime6(last) - sg.first;
    }
    return run_component(c, sg, y, 0, 0, 0, 0, lo, hi);
}

mdr_status mdr_s1_sum(mdr_ctx* c, long long y, long long first, long long last, int stride, long long* lo,
                      long long* hi) {
    if (!c || !lo || !hi || (stride != 1 && stride != 2) || first < 1 || y < 0) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->tab.valid) return fail(c, MDR_ERR_NOTABLE, "no table built");
    if (first <= last && y / first > c->tab.U) return fail(c, MDR_ERR_INVALID, "quotient beyond the table");
    // split at the tier boundary like the engine does
    const i64 split = y / (c->tab.ud + 1);
    u64 rlo = 0;
    i64 rhi = 0;
    for (int comp = 1; comp >= 0; --comp) {
        Seg sg{};
        sg.type = comp ? SEG_S1_COMP : SEG_S1_DENSE;
        sg.sign = 1;
        sg.stride = stride;
        i64 a = first, b = last;
        if (comp) {
            if (b > split) b = split;
        } else if (a <= split) {
            a = split + 1;
            if (stride == 2 && ((a - first) & 1)) ++a;
        }
        if (a > b) continue;
        if (comp && c->tab.L == 0) return fail(c, MDR_ERR_INVALID, "quotient beyond the dense table");
        sg.first = a;
        sg.count = (b - a) / stride + 1;
        long long l, h;
        mdr_status s = run_component(c, sg, y, 0, 0, 0, 0, &l, &h);
        if (s != MDR_OK) return s;
        add128(rlo, rhi, (u64)l, (i64)h);
    }
    *lo = (long long)rlo;
    *hi = (long long)rhi;
    return MDR_OK;
}

mdr_status mdr_g_eval(mdr_ctx* c, long long y, int comb2, long long u, double cx, long long* lo, long long* hi) {
    if (!c || !lo || !hi || y < 4 || u < 2 || !(cx > 0)) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->tab.valid) return fail(c, MDR_ERR_NOTABLE, "no table built");
    if (u > c->tab.U) return fail(c, MDR_ERR_INVALID, "u beyond the table");
    if ((i64)(cx * std::sqrt((double)y)) > c->tab.mu) return fail(c, MDR_ERR_INVALID, "split point beyond the mu table");
    return run_component(c, Seg{}, y, 1, comb2 ? 1 : 0, u, cx, lo, hi);
}

}  // extern "C"

// ---- learned cost model --------------------------------------------------------------
namespace {
const char MODEL_MAGIC[8] = {'M', 'D', 'R', 'M', 'L', 'P', '0', '1'};

mdr_status set_model(mdr_ctx* c, const float* w, int n) {
    if (n == 0) {
        c->mlp_ok = false;
        c->mlp_host.clear();
        return MDR_OK;
    }
    if (n != MLP_NPARAM || !w) return fail(c, MDR_ERR_INVALID, "cost model needs MDR_MLP_NPARAM weights");
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(w[i])) return fail(c, MDR_ERR_INVALID, "cost model weights must be finite");
    HIPT(c->mlp.ensure(sizeof(float) * n), "alloc cost model");
    HIPT(h2d(c, c->mlp.p, w, sizeof(float) * n), "upload cost model");
    c->mlp_host.assign(w, w + n);
    c->mlp_ok = true;
    return MDR_OK;
}

struct DevTmp {
    void* p = nullptr;
    ~DevTmp() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t b) { return hipMalloc(&p, b < 16 ? 16 : b); }
    template <typename T> T* as() { return static_cast<T*>(p); }
};

struct PinnedTmp {
    void* p = nullptr;
    ~PinnedTmp() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t alloc(size_t b) { return hipHostMalloc(&p, b < 16 ? 16 : b, hipHostMallocDefault); }
};

struct EventTmp {
    hipEvent_t e = nullptr;
    ~EventTmp() {
        if (e) (void)hipEventDestroy(e);
    }
    hipError_t create() { return hipEventCreateWithFlags(&e, hipEventDisableTiming); }
};
}  // namespace

extern "C" {

mdr_status mdr_cost_model_set(mdr_ctx* c, const float* w, int n) {
    if (!c) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    return set_model(c, w, n);
}

mdr_status mdr_cost_model_get(mdr_ctx* c, float* w, int n) {
    if (!c || !w) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->mlp_ok) return fail(c, MDR_ERR_INVALID, "no cost model");
    if (n < MLP_NPARAM) return fail(c, MDR_ERR_INVALID, "buffer too small");
    std::memcpy(w, c->mlp_host.data(), sizeof(float) * MLP_NPARAM);
    return MDR_OK;
}

mdr_status mdr_cost_model_load(mdr_ctx* c, const char* path) {
    if (!c || !path) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(c, MDR_ERR_INVALID, std::string("cannot open ") + path);
    char magic[8];
    std::vector<float> w(MLP_NPARAM);
    const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, MODEL_MAGIC, 8) == 0 &&
                    std::fread(w.data(), sizeof(float), MLP_NPARAM, f) == (size_t)MLP_NPARAM;
    std::fclose(f);
    if (!ok) return fail(c, MDR_ERR_INVALID, std::string("not a cost model file: ") + path);
    return set_model(c, w.data(), MLP_NPARAM);
}

mdr_status mdr_cost_model_save(mdr_ctx* c, const char* path) {
    if (!c || !path) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->mlp_ok) return fail(c, MDR_ERR_INVALID, "no cost model");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(c, MDR_ERR_INVALID, std::string("cannot write ") + path);
    const bool ok = std::fwrite(MODEL_MAGIC, 1, 8, f) == 8 &&
                    std::fwrite(c->mlp_host.data(), sizeof(float), MLP_NPARAM, f) == (size_t)MLP_NPARAM;
    std::fclose(f);
    return ok ? MDR_OK : fail(c, MDR_ERR_INVALID, "short write");
}

mdr_status mdr_cycle_log(mdr_ctx* c, mdr_stamp* out, size_t cap, size_t* n) {
    if (!c || !n || (cap && !out)) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    const size_t k = std::min(cap, c->last_stamps.size());
    if (k) std::memcpy(out, c->last_stamps.data(), k * sizeof(mdr_stamp));
    *n = c->last_stamps.size();
    return MDR_OK;
}

mdr_status mdr_mlp_features(const mdr_stamp* s, size_t n, float* feats, float* targets) {
    if ((n && (!s || !feats || !targets))) return MDR_ERR_INVALID;
    for (size_t i = 0; i < n; ++i) {
        if (s[i].y < 1 || s[i].start < 1 || s[i].end < s[i].start || s[i].count == 0) return MDR_ERR_INVALID;
        piece_features(feats + i * MLP_IN, s[i].y, s[i].start, s[i].end, (int)s[i].kind);
        targets[i] = std::log2((float)(s[i].cycles > 0 ? s[i].cycles : 1) / (float)s[i].count);
    }
    return MDR_OK;
}

mdr_status mdr_mlp_predict(mdr_ctx* c, const float* w, const float* feats, size_t n, float* out) {
    if (!c || !w || (n && (!feats || !out)) || n > (size_t)MAX_CALL_ITEMS) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    if (n == 0) return MDR_OK;
    DevTmp dw, dx, dy;
    HIPT(dw.alloc(sizeof(float) * MLP_NPARAM), "alloc");
    HIPT(dx.alloc(sizeof(float) * MLP_IN * n), "alloc");
    HIPT(dy.alloc(sizeof(float) * n), "alloc");
    HIPT(h2d(c, dw.p, w, sizeof(float) * MLP_NPARAM), "upload");
    HIPT(h2d(c, dx.p, feats, sizeof(float) * MLP_IN * n), "upload");
    HIPT(mlp_predict(dw.as<float>(), dx.as<float>(), (i64)n, dy.as<float>(), c->st), "predict");
    HIPT(hipMemcpyAsync(out, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->st), "download");
    HIPT(hipStreamSynchronize(c->st), "predict");
    return MDR_OK;
}

mdr_status mdr_mlp_grad(mdr_ctx* c, const float* w, const float* feats, const float* targets, size_t n, float* grad,
                        float* loss) {
    if (!c || !w || !feats || !targets || !grad || n == 0 || n > (size_t)MAX_CALL_ITEMS) return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    DevTmp dw, dx, dt, buf, err, dg, dl;
    HIPT(dw.alloc(sizeof(float) * MLP_NPARAM), "alloc");
    HIPT(dx.alloc(sizeof(float) * MLP_IN * n), "alloc");
    HIPT(dt.alloc(sizeof(float) * n), "alloc");
    HIPT(buf.alloc(sizeof(float) * mlp_row_floats() * n), "alloc");
    HIPT(err.alloc(sizeof(float) * n), "alloc");
    HIPT(dg.alloc(sizeof(float) * MLP_NPARAM), "alloc");
    HIPT(dl.alloc(sizeof(float)), "alloc");
    HIPT(h2d(c, dw.p, w, sizeof(float) * MLP_NPARAM), "upload");
    HIPT(h2d(c, dx.p, feats, sizeof(float) * MLP_IN * n), "upload");
    HIPT(h2d(c, dt.p, targets, sizeof(float) * n), "upload");
    HIPT(mlp_grad(dw.as<float>(), dx.as<float>(), dt.as<float>(), nullptr, (i64)n, buf.as<float>(), err.as<float>(),
                  dg.as<float>(), dl.as<float>(), c->st),
         "gradient");
    HIPT(hipMemcpyAsync(grad, dg.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c->st), "download");
    float l = 0;
    HIPT(hipMemcpyAsync(&l, dl.p, sizeof(float), hipMemcpyDeviceToHost, c->st), "download");
    HIPT(hipStreamSynchronize(c->st), "gradient");
    if (loss) *loss = l;
    return MDR_OK;
}

mdr_status mdr_mlp_train(mdr_ctx* c, const float* feats, const float* targets, size_t n, int epochs, int batch,
                         float lr, unsigned int seed, float* final_loss) {
    if (!c || !feats || !targets || n == 0 || n > (size_t)MAX_CALL_ITEMS || epochs < 1 || batch < 1 || !(lr > 0))
        return MDR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(c->mu);
    std::vector<float> w0 = c->mlp_host;
    std::mt19937 rng(seed);
    if (!c->mlp_ok) {
        w0.assign(MLP_NPARAM, 0.f);
        auto fill = [&](int off, int cnt, int fan_in) {
            std::uniform_real_distribution<float> d(-std::sqrt(6.f / fan_in), std::sqrt(6.f / fan_in));
            for (int i = 0; i < cnt; ++i) w0[off + i] = d(rng);
        };
        fill(MLP_W1, MLP_H * MLP_IN, MLP_IN);
        fill(MLP_W2, MLP_H * MLP_H, MLP_H);
        fill(MLP_W3, MLP_H, MLP_H);
        double mean = 0;
        for (size_t i = 0; i < n; ++i) mean += targets[i];
        w0[MLP_B3] = (float)(mean / (double)n);
    }
    const size_t bsz = std::min<size_t>((size_t)batch, n);
    DevTmp dw, dx, dt, buf, err, dg, dm, dv, dl, didx[2];
    PinnedTmp hidx[2];
    EventTmp used[2];
    HIPT(dw.alloc(sizeof(float) * MLP_NPARAM), "alloc");
    HIPT(dx.alloc(sizeof(float) * MLP_IN * n), "alloc");
    HIPT(dt.alloc(sizeof(float) * n), "alloc");
    HIPT(buf.alloc(sizeof(float) * mlp_row_floats() * std::max(bsz, n)), "alloc");
    HIPT(err.alloc(sizeof(float) * n), "alloc");
    HIPT(dg.alloc(sizeof(float) * MLP_NPARAM), "alloc");
    HIPT(dm.alloc(sizeof(float) * MLP_NPARAM), "alloc");
    HIPT(dv.alloc(sizeof(float) * MLP_NPARAM), "alloc");
    for (int k = 0; k < 2; ++k) {
        HIPT(didx[k].alloc(sizeof(int) * n), "alloc");
        HIPT(hidx[k].alloc(sizeof(int) * n), "alloc");
        HIPT(used[k].create(), "event");
    }
    HIPT(dl.alloc(sizeof(float)), "alloc");
    HIPT(h2d(c, dw.p, w0.data(), sizeof(float) * MLP_NPARAM), "upload");
    HIPT(h2d(c, dx.p, feats, sizeof(float) * MLP_IN * n), "upload");
    HIPT(h2d(c, dt.p, targets, sizeof(float) * n), "upload");
    HIPT(clear_sync(c, dm.p, sizeof(float) * MLP_NPARAM), "clear");
    HIPT(clear_sync(c, dv.p, sizeof(float) * MLP_NPARAM), "clear");
    std::vector<int> idx(n);
    for (size_t i = 0; i < n; ++i) idx[i] = (int)i;
    // Two pinned/device index buffers: the host shuffles epoch e + 1 while the
    // GPU runs epoch e, and waits only before refilling the buffer of epoch e - 1.
    int step = 0;
    for (int e = 0; e < epochs; ++e) {
        const int k = e & 1;
        std::shuffle(idx.begin(), idx.end(), rng);
        if (e >= 2) HIPT(hipEventSynchronize(used[k].e), "epoch");
        std::memcpy(hidx[k].p, idx.data(), sizeof(int) * n);
        HIPT(hipMemcpyAsync(didx[k].p, hidx[k].p, sizeof(int) * n, hipMemcpyHostToDevice, c->st), "upload");
        for (size_t b0 = 0; b0 + bsz <= n; b0 += bsz) {
            HIPT(mlp_train_step(dw.as<float>(), dx.as<float>(), dt.as<float>(), didx[k].as<int>() + b0, (i64)bsz,
                                buf.as<float>(), dm.as<float>(), dv.as<float>(), lr, ++step, c->st),
                 "train step");
        }
        HIPT(hipEventRecord(used[k].e, c->st), "epoch");
    }
    // loss over all samples with the final weights
    HIPT(mlp_grad(dw.as<float>(), dx.as<float>(), dt.as<float>(), nullptr, (i64)n, buf.as<float>(), err.as<float>(),
                  dg.as<float>(), dl.as<float>(), c->st),
         "loss");
    std::vector<float> w(MLP_NPARAM);
    float l = 0;
    HIPT(hipMemcpyAsync(w.data(), dw.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c->st), "download");
    HIPT(hipMemcpyAsync(&l, dl.p, sizeof(float), hipMemcpyDeviceToHost, c->st), "download");
    HIPT(hipStreamSynchronize(c->st), "train");
    if (final_loss) *final_loss = l;
    return set_model(c, w.data(), MLP_NPARAM);
}

}  // extern "C"
