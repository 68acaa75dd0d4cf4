/*
 * C ABI of the MI355X Deleglise-Rivat Mertens engine (libmdr_hip.so).
 *
 * Plain C types only: 64-bit integers, doubles, pointers and sizes. Nothing
 * throws across this boundary; every call returns an mdr_status and writes its
 * results through caller-owned pointers.  A context owns one GPU's device
 * buffers (tables are cached across calls) and serialises the calls made on
 * it; use one context per host thread.
 *
 * Reference interfaces each entry point replaces:
 *   mdr_mertens        mertens_dr::DelégliseRivatEngine::compute_mertens
 *                      (mertens_dr.hpp:968) and compute_compressed (:1442);
 *                      mertens_cuda::CudaDelégliseRivatEngine::compute_mertens
 *                      with its TimingStats out-pointer (mertens_cuda.cuh:410-416, :441)
 *   mdr_mertens_batch  the same, for many X sharing one table (no reference
 *                      counterpart beyond repeated compute_mertens calls)
 *   mdr_table_build/read_*  the context's cached table, read as SieveTable /
 *                      CompressedMertensTable get_M and get_mu (mertens_dr.hpp:404-425,
 *                      :707-716)
 *   mdr_table_create, _free, _info, _copy_M, _copy_mu, _gather_M, _copy_chunks
 *                      table objects: the
 *                      constructors SieveTable(limit, threads, max_mu, ...) and
 *                      CompressedMertensTable(total_limit, dense_limit, threads,
 *                      max_mu) (mertens_dr.hpp:386-402, :689-702) with their reads
 *                      get_M / get_mu / data / dense_data (:404-425, :707-716) and
 *                      the compressed tier's BitChunk + checkpoint arrays
 *                      (:670-687, :741-752)
 *   mdr_s2_interval    run_s2_fast<Piece> (mertens_dr.hpp:484-617)
 *   mdr_s1_sum         sum_dense_quotients<Stride> (mertens_dr.hpp:88-128) and the
 *                      compressed-tier loops (:1499-1504)
 *   mdr_g_eval         eval_single_S / eval_comb2 (mertens_dr.hpp:1081-1354)
 *   mdr_param_*        choose_cx / choose_sieve_limit / choose_compressed_limit
 *                      (mertens_dr.hpp:914-962, :1401-1419)
 *   mdr_plan           choose_cx / choose_sieve_limit / choose_compressed_limit and
 *                      the strategy switch (mertens_dr.hpp:914-962, :1027-1041,
 *                      :1401-1461)
 */
#ifndef MERTENS_HIP_H
#define MERTENS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdr_ctx mdr_ctx;
typedef struct mdr_table mdr_table;

typedef enum {
    MDR_OK = 0,
    MDR_ERR_INVALID = 1,   /* bad argument or option out of its documented range */
    MDR_ERR_HIP = 2,       /* a HIP runtime call failed (see mdr_last_error) */
    MDR_ERR_NOMEM = 3,     /* device allocation failed */
    MDR_ERR_NODEVICE = 4,  /* no usable GPU */
    MDR_ERR_RANGE = 5,     /* result or work size outside the supported range */
    MDR_ERR_NOTABLE = 6    /* component call made before a table was built */
} mdr_status;

/* strategy: MERTENS_USE_CHMT (mertens_dr.hpp:1027-1029, :1421-1435) */
enum {
    MDR_STRATEGY_AUTO = 0,       /* variable unset: streaming from 1e11 unless fx/u_cap_dense set */
    MDR_STRATEGY_DENSE = 1,      /* variable set, not 1/y/Y/t/T */
    MDR_STRATEGY_STREAMING = 2,  /* variable set to 1/y/Y/t/T: streaming from 1e11 */
    MDR_STRATEGY_FORCE_STREAMING = 3  /* compute_compressed: streaming for X > 1e5 */
};
/* chunk partition of the quotient phase: static = equal quotient positions per
 * chunk; learned = equal predicted cycles (the cost model; applied when the
 * call's largest summed X is at least 1e14, the low end of the model's
 * training range, and static otherwise); terms = equal
 * numbers of outer terms (G evaluations) per chunk, the reference GPU engine's
 * unit of work (one term per warp, mertens_cuda.cuh:307-403) */
enum { MDR_PARTITION_STATIC = 0, MDR_PARTITION_LEARNED = 1, MDR_PARTITION_TERMS = 2 };
enum { MDR_PARAMS_REFERENCE = 0, MDR_PARAMS_TUNED = 1 };
enum { MDR_SCHED_DYNAMIC = 0, MDR_SCHED_STATIC = 1 };

typedef struct {
    int strategy;             /* MDR_STRATEGY_* */
    double cx;                /* MERTENS_CX, 0 = default; accepted 0.1 < cx < 5.0 */
    long long u_cap;          /* MERTENS_CHMT_U_CAP, 0 = default 8e9; accepted 1e8..1e10 */
    double fx;                /* MERTENS_FX, 0 = default 0.32; accepted 0.1..3.0 */
    long long u_cap_dense;    /* MERTENS_U_CAP, 0 = default 1.5e9; accepted 1e8..4e9 */
    int fx_set;               /* MERTENS_FX present in the environment (disables auto streaming) */
    int u_cap_dense_set;      /* MERTENS_U_CAP present in the environment (same) */
    int partition;            /* MDR_PARTITION_* */
    int params;               /* MDR_PARAMS_*: reference parameters or the autotuned GPU split */
    int sched;                /* MDR_SCHED_* */
    int host_threads;         /* accepted for interface parity; the GPU path ignores it */
    long long chunk;          /* quotient-phase chunk count P, 0 = automatic */
    int record_cycles;        /* s > 0: log cycles of every cost piece of 1 in s G evaluations */
    int no_table_reuse;       /* 1: rebuild the table even when the cached one fits */
    /* explicit GPU split for MDR_PARAMS_TUNED (0 = from the tuning cache):
     * sieve cutoff (<= 1e12; a table whose compressed tier U - u_dense
     * exceeds 256 * (2^32 - 1) values fails with MDR_ERR_RANGE), dense-tier
     * limit and S1/S2 split factor */
    long long tune_u;
    long long tune_ud;
    double tune_cx;
    int use_graph;            /* 1: capture the call's launches into a hipGraph and replay it
                                 (instantiated once per launch sequence, updated in place) */
    double tune_batch;        /* batch cutoff factor for MDR_PARAMS_TUNED batches
                                 (0 = from the tuning cache; accepted 0.05..20) */
    int sieve_shape;          /* sieve segment length / workgroup: 0 = automatic (tuned with
                                 MDR_PARAMS_TUNED), 1 = 49152/1024, 2 = 24576/256,
                                 3 = 98304/1024, 4 = 49152/256, 5 = 49152/512,
                                 6 = 24576/512; the table is the same */
    long long window;         /* windowed table (compute_mertens_windowed,
                                 mertens_dr_windowed.hpp:49-407): 0 = the whole table is
                                 resident; n > 0 = keep only mu, the dense tier and the
                                 values up to max(cx sqrt(X) + 1000, X/u + 1000) resident and
                                 stream the compressed tier above them through windows of
                                 n values (rounded up to whole sieve segments); the result
                                 is the same, the table's memory is the resident part plus
                                 one window, and the cutoff may exceed the one-table limit
                                 (up to tune_u <= 1e12); such calls run eagerly (no graph) */
    int wave_span;            /* 1: every wave of the quotient kernel records its first and
                                 last clock (constant-rate wall clock); stats then report the
                                 waves' busy fraction and the idle tail (0 = off) */
} mdr_opts;

typedef struct {
    /* mirrors CudaDelégliseRivatEngine::TimingStats (seconds), all host-timed */
    double host_sieve_time;   /* always 0: the sieve runs on the GPU */
    double h2d_copy_time;     /* parameter upload */
    double kernel_time;       /* table build + partition + quotient phase (GPU events) */
    double d2h_copy_time;     /* result download */
    double total_time;
    /* GPU phase times in milliseconds (hipEvents on the engine's stream) */
    double table_ms, partition_ms, quotient_ms;
    int table_reused;
    long long U, u_dense, mu_limit, W;   /* table actually used; W = compressed 64-byte lines */
    long long u, N, slots, chunks;       /* last X of the call */
    double cx;
    long long s1_terms, s2_terms;        /* quotient positions (sum over the call) */
    double table_bytes;                  /* bytes of the table arrays written */
    long long s1_comp_terms;             /* S1 positions answered by the compressed tier */
    int partition_used;                  /* MDR_PARTITION_* actually applied (learned falls
                                            back to static without a cost model and when
                                            every summed X is below 1e14) */
    int graph_used;                      /* the call ran as one hipGraph launch */
    int graph_instantiated;              /* graphs instantiated (not updated in place) */
    double graph_ms;                     /* GPU time of the graph launches; the per-phase
                                            times stay 0 for graph calls */
    long long windows;                   /* compressed-tier windows streamed (windowed calls) */
    /* quotient phase against its models: the cost model's predicted cycles (learned
     * partition; summed over the G evaluations), the measured wave-cycles
     * (quotient_ms x shader clock x resident waves) and the algorithmic bytes
     * (2 B per dense S1 lookup, 20 B per compressed one, 3 B per S2 position) */
    double predicted_cycles, measured_cycles, quotient_bytes;
    /* opts.wave_span = 1 (last quotient launch of the call): span = last wave
     * end - first wave start, busy = mean over waves of (end - start); their
     * ratio is the fraction of the span the average wave worked, so
     * 1 / wave_busy_frac bounds what any partition of the same work can gain
     * in the quotient kernel */
    double wave_span_ms, wave_busy_ms, wave_busy_frac;
} mdr_stats;

/* One cycle-log record: a run of S1 (n) or S2 (j) positions of one G(y)
 * evaluation processed by one wave, and the shader-clock cycles it took. */
typedef struct {
    long long y, start, end;   /* start/end: first and last n or j of the run */
    unsigned int kind;         /* 0..3 S1 {dense,compressed}x{stride 1,2}; 4..15 S2 pieces 1..12 */
    unsigned int cycles;
    unsigned long long count;  /* positions */
} mdr_stamp;

#define MDR_MLP_NFEAT 5
#define MDR_MLP_NPARAM 1281

void mdr_default_opts(mdr_opts* opts);
mdr_status mdr_init(mdr_ctx** ctx, int device);
void mdr_shutdown(mdr_ctx* ctx);
const char* mdr_last_error(mdr_ctx* ctx);

/* M(X) for one X; *out receives the exact signed 64-bit value. */
mdr_status mdr_mertens(mdr_ctx* ctx, long long X, const mdr_opts* opts, long long* out,
                       mdr_stats* stats);
/* M(X[i]) for n values sharing one table sized to the largest cutoff. */
mdr_status mdr_mertens_batch(mdr_ctx* ctx, const long long* X, size_t n, const mdr_opts* opts,
                             long long* out, mdr_stats* stats);
/* M(X) with no device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory during the call (code running
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every device buffer of the call lives in one 256 MiB array
 * that each GPU holds for the life of the process (set up by mdr_init), so
 * the call performs no hipMalloc, hipFree or other allocation; inputs and
 * outputs are the caller's.  The plan is the tuned split with the cutoff
 * lowered until the call fits the array (M(X) does not depend on it); X
 * beyond what fits (about X > 8e15) fails with MDR_ERR_RANGE.  Static
 * partition, dynamic chunk queue, table rebuilt every call.  All calls on
 * one GPU share its array: calls through contexts of one process are
 * serialised per device, and calls that overlap on one GPU by other means
 * would race.  stats (optional): table_ms, quotient_ms, U, u_dense,
 * mu_limit, W, u, N, slots, chunks, cx. */
mdr_status mdr_mertens_noalloc(mdr_ctx* ctx, long long X, long long* out, mdr_stats* stats);

/* Cross-check engine: dirichlet::DirichletEngine::compute_mertens (the
 * reference's optional dirichlet_engine.hpp, included by bench_dr.cpp:9-12 and
 * test_dr.cpp:8-11 and cross-validated at test_dr.cpp:71-88; not shipped with
 * it).  Exact M(X) from sum_{d <= v} M(floor(v/d)) = 1 over the quotients
 * v = floor(X/k), with a table to about X^(2/3) (at most the one-table cap):
 * an algorithm independent of the Deleglise-Rivat split, about
 * 4 X / sqrt(table) term evaluations.  X > 1e18 fails with MDR_ERR_RANGE.
 * stats: table_ms, quotient_ms (the large values), U = the table limit,
 * N = the large values computed. */
mdr_status mdr_mertens_dirichlet(mdr_ctx* ctx, long long X, long long* out, mdr_stats* stats);

/* Waves the persistent quotient kernel keeps resident on the context's GPU
 * (the natural chunk count of a static round-robin schedule). */
long long mdr_quot_waves(mdr_ctx* ctx);

/* Parameters the engine would use for X (reference rules or tuned).  Host-only;
 * ctx may be NULL. */
mdr_status mdr_plan(mdr_ctx* ctx, long long X, const mdr_opts* opts, int* streaming, long long* u,
                    long long* u_dense, long long* mu_limit, long long* N, double* cx);

/* The reference's parameter rules as the engine applies them (host-only, no
 * context).  opts may be NULL (the reference's defaults); out-of-range
 * overrides are ignored as the reference ignores out-of-range variables.
 *   mdr_param_cx                choose_cx (mertens_dr.hpp:914-923): opts->cx
 *                               when 0.1 < cx < 5, else 1.35 / 1.15 / 0.95 / 0.70
 *                               by decade of X
 *   mdr_param_dense_cutoff      choose_sieve_limit (:931-962): the dense cutoff,
 *                               with opts->fx (MERTENS_FX) and opts->u_cap_dense
 *                               (MERTENS_U_CAP)
 *   mdr_param_streaming_cutoff  choose_compressed_limit (:1401-1419): the
 *                               streaming cutoff, with opts->u_cap
 *                               (MERTENS_CHMT_U_CAP) and the cx above */
double mdr_param_cx(long long X, const mdr_opts* opts);
long long mdr_param_dense_cutoff(long long X, const mdr_opts* opts);
long long mdr_param_streaming_cutoff(long long X, const mdr_opts* opts);

/* Component entry points (parity tests).  mdr_table_build sieves n = 1..U
 * with the dense tier up to u_dense (a multiple of 256 when below U; U - u_dense
 * at most 256 * (2^32 - 1), else MDR_ERR_RANGE) and packed mu up to mu_limit; the reads
 * copy M(first..first+count-1) and mu(...) to host. */
mdr_status mdr_table_build(mdr_ctx* ctx, long long U, long long u_dense, long long mu_limit);
mdr_status mdr_table_read_M(mdr_ctx* ctx, long long first, long long count, int32_t* out);
mdr_status mdr_table_read_mu(mdr_ctx* ctx, long long first, long long count, int8_t* out);
/* Table objects: device-resident tables built by the engine's sieve on the
 * context's GPU, any number per context besides its cached table, each
 * living until mdr_table_free (which must come before mdr_shutdown of its
 * context).  A table holds M(n) for 0 <= n <= U and mu(n) for n <= mu_limit
 * (mu reads beyond it give 0, as SieveTable::get_mu does); u_dense is the
 * table's int16 tier (rounded down to a multiple of 256 below U; the values
 * read are the same for any u_dense).  Calls on a table are serialised with
 * the other calls on its context.
 *   create       U >= 0; 0 <= mu_limit <= U (0 = none); *out receives the table
 *   info         the sizes actually built
 *   copy_M, _mu  M(first .. first+count-1) as int32, mu(...) as int8, into host
 *                memory; first >= 0, first + count - 1 <= U for M
 *   gather_M     out[i] = M(q[i]) for host arrays q, out (0 <= q[i] <= U)
 *   copy_chunks  the compressed tier of CompressedMertensTable with dense
 *                limit u_dense_ref (u_dense_ref <= U): chunks
 *                first_chunk .. first_chunk+count-1 of 256 values from
 *                u_dense_ref + 1 on, each 64 bytes {nonzero[32], sign[32]}
 *                (bit i of the chunk: mu != 0, mu < 0 of value i), and four
 *                int32 checkpoints per chunk (M just before each 64-value
 *                word); values beyond U read as mu = 0 */
mdr_status mdr_table_create(mdr_ctx* ctx, long long U, long long u_dense, long long mu_limit, mdr_table** out);
void mdr_table_free(mdr_table* table);
mdr_status mdr_table_info(const mdr_table* table, long long* U, long long* u_dense, long long* mu_limit);
mdr_status mdr_table_copy_M(mdr_table* table, long long first, long long count, int32_t* out);
mdr_status mdr_table_copy_mu(mdr_table* table, long long first, long long count, int8_t* out);
mdr_status mdr_table_gather_M(mdr_table* table, const long long* q, size_t n, int32_t* out);
mdr_status mdr_table_copy_chunks(mdr_table* table, long long u_dense_ref, long long first_chunk, long long count,
                                 void* chunks, int32_t* checkpoints);

/* sum over j in [first, last], gcd(j,6)=1 of mu(j)*piece(floor(y/j)); piece 1..8 are
 * Piece1..8, 9..12 SinglePiece1..4.  128-bit result as (lo, hi). */
mdr_status mdr_s2_interval(mdr_ctx* ctx, int piece, long long first, long long last, long long y,
                           long long* lo, long long* hi);
/* sum over n = first, first+stride, ... <= last of M(floor(y/n)); stride 1 or 2. */
mdr_status mdr_s1_sum(mdr_ctx* ctx, long long y, long long first, long long last, int stride,
                      long long* lo, long long* hi);
/* G(y) (comb2 = 0) or G(y) - G(floor(y/2)) (comb2 = 1) with cutoff u and split cx. */
mdr_status mdr_g_eval(mdr_ctx* ctx, long long y, int comb2, long long u, double cx, long long* lo,
                      long long* hi);

/* Learned cost model: a 5 -> 32 -> 32 -> 1 FP32 MLP (ReLU) predicting
 * log2(cycles per position) of a quotient run; no reference counterpart (the
 * reference schedules whole terms dynamically, mertens_dr.hpp:1616).
 * Weights are MDR_MLP_NPARAM floats: W1[32][5], b1[32], W2[32][32], b2[32],
 * W3[32], b3.  set with n = 0 removes the model (static partition). */
mdr_status mdr_cost_model_set(mdr_ctx* ctx, const float* w, int n);
mdr_status mdr_cost_model_get(mdr_ctx* ctx, float* w, int n);
mdr_status mdr_cost_model_load(mdr_ctx* ctx, const char* path);
mdr_status mdr_cost_model_save(mdr_ctx* ctx, const char* path);
/* Cycle log of the last call made with opts.record_cycles = stride > 0. */
mdr_status mdr_cycle_log(mdr_ctx* ctx, mdr_stamp* out, size_t cap, size_t* n);
/* Features (n x 5) and targets log2(cycles/count) of cycle-log records (host). */
mdr_status mdr_mlp_features(const mdr_stamp* s, size_t n, float* feats, float* targets);
/* GPU kernels of the network: inference, one forward+backward pass (mean
 * squared error over n samples; grad has MDR_MLP_NPARAM entries) and Adam
 * training from the context's model (He-uniform from seed when it has none);
 * training installs the result as the context's cost model. */
mdr_status mdr_mlp_predict(mdr_ctx* ctx, const float* w, const float* feats, size_t n, float* out);
mdr_status mdr_mlp_grad(mdr_ctx* ctx, const float* w, const float* feats, const float* targets, size_t n,
                        float* grad, float* loss);
mdr_status mdr_mlp_train(mdr_ctx* ctx, const float* feats, const float* targets, size_t n, int epochs,
                         int batch, float lr, unsigned int seed, float* final_loss);

#ifdef __cplusplus
}
#endif

#endif
