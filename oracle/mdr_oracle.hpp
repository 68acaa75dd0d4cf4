// CPU oracle for the Deleglise-Rivat Mertens path.
//
// TEST INFRASTRUCTURE ONLY. This is a plain C++20/OpenMP restatement of the
// reference CPU engine (mertens_dr.hpp) used to check the HIP engine. Only the
// tests, the smoke entry point and the CPU-baseline leg of the benchmark may
// load it, and only to check results. The product never links or calls it.
//
// Every function cites the reference lines it restates. The arithmetic is
// written from the mathematics: floor divisions are exact integer divisions
// (the reference's fast_div is exact as well, mertens_dr.hpp:76-86), S2
// summands use their closed floor forms rather than the reference's lookup
// tables, and sums are accumulated in 128-bit integers.
//
// Pins (each checked in tests/test_oracle.py):
//   isqrt                      -> test_sieve.cpp:44-51 cases
//   mobius_bytelog_*           -> Eratosthenes (test_sieve.cpp:16-29 semantics)
//   DenseTable/CompressedTable -> prefix sums of the Eratosthenes mu
//   finalize_byte              -> closed form at test_sieve.cpp:62-67
//   piece_*                    -> floor-form identities checked against the
//                                 reference's LUT forms (mertens_dr.hpp:462-482)
//   s2_interval                -> brute force (test_sieve.cpp:98-116)
//   g_single / g_comb2         -> brute-force G(y) and G(y)-G(y/2)
//   mertens                    -> known values (test_dr.cpp:26-47,
//                                 test_cuda.cu:59, README.md:188), the
//                                 recurrence (test_sieve.cpp:30-41) and the
//                                 compiled reference under oracle/_ref.
#ifndef MDR_ORACLE_HPP
#define MDR_ORACLE_HPP

#include <cstdint>
#include <vector>
#include <string>

namespace mdr_oracle {

using i64 = long long;
using u64 = unsigned long long;
using i128 = __int128;

// ---- primitives -----------------------------------------------------------
i64 isqrt(i64 n);                         // mertens_dr.hpp:65-74
inline i64 qdiv(i64 y, i64 d) { return y / d; }  // mertens_dr.hpp:80-86 (exact)
unsigned bit_width_u64(u64 v);
unsigned bytelog_weight(i64 p);           // mertens_dr.hpp:195 (bit_width(p-1)|1)
// Byte-log finalisation of one accumulator for value n (n >= 1).
// mertens_dr.hpp:143-171; n = 1 and n = 2 need the fix-ups of :294-295, :836-839.
int8_t finalize_byte(uint8_t acc, i64 n);

// ---- sieves ---------------------------------------------------------------
// Textbook Eratosthenes mu(0..limit), mu(0) = 0 (test_sieve.cpp:16-29).
std::vector<int8_t> mobius_eratosthenes(i64 limit);
// Base primes up to r and their byte-log weights.
struct BasePrimes { std::vector<i64> p; std::vector<uint8_t> w; };
BasePrimes base_primes(i64 r);
// mu(n) for n in [low, low+len) by the full-number byte-log sieve with primes
// up to isqrt(sieve_limit) (mertens_dr.hpp:814-839); low >= 1.
void mobius_bytelog_block(i64 low, i64 len, const BasePrimes& bp, int8_t* out);
// Odd-only table mu_odd[k] = mu(2k-1), k = 1..half_u (mertens_dr.hpp:173-296).
void mobius_bytelog_odd(int8_t* mu_odd, i64 half_u, i64 u, int threads);
int8_t mu_from_odd(i64 n, const int8_t* mu_odd);   // mertens_dr.hpp:130-141

// ---- tables ---------------------------------------------------------------
// SieveTable (mertens_dr.hpp:376-457): mu up to mu_limit, int16 M up to m_limit.
struct DenseTable {
    i64 u = 0, mu_limit = 0, m_limit = 0;
    std::vector<int8_t> mu;
    std::vector<int16_t> M;
    DenseTable(i64 limit, int threads, i64 max_mu = 0, i64 max_M = -1);
    int32_t get_M(i64 t) const { return M[(size_t)t]; }
    int8_t get_mu(i64 t) const { return t <= mu_limit ? mu[(size_t)t] : 0; }
};

// CompressedMertensTable (mertens_dr.hpp:675-902): int16 dense tier up to
// u_dense, 256-value chunks of two bit planes with four int32 checkpoints
// above it, mu up to mu_limit.
struct CompressedTable {
    i64 u_dense = 0, u_total = 0, mu_limit = 0, compressed_size = 0, num_chunks = 0;
    std::vector<int16_t> dense_M;
    std::vector<int8_t> mu;
    std::vector<u64> nz, neg;          // one 64-bit word per 64 values
    std::vector<int32_t> checkpoints;  // M immediately before each word
    CompressedTable(i64 total_limit, i64 dense_limit, int threads, i64 max_mu = 0);
    int32_t get_M(i64 q) const;
    int8_t get_mu(i64 j) const { return j <= mu_limit ? mu[(size_t)j] : 0; }
};

// ---- S2 summands (mertens_dr.hpp:459-482), closed floor forms ---------------
// piece ids: 1..8 = Piece1..Piece8, 9..12 = SinglePiece1..SinglePiece4.
i64 piece_closed(int piece, i64 q);
i64 piece_lut(int piece, i64 q);      // LUT forms, restated for the identity test
// sum_{j in [first,last], gcd(j,6)=1} mu(j) * piece(floor(y/j)) (mertens_dr.hpp:484-617)
i128 s2_interval(int piece, i64 first, i64 last, i64 y, const int8_t* mu);
i128 s2_combined(i64 y, i64 A, i64 B, const int8_t* mu);   // mertens_dr.hpp:622-642
i128 s2_single(i64 y, i64 A, const int8_t* mu);            // mertens_dr.hpp:647-660

// ---- parameters (mertens_dr.hpp:914-962, 1390-1435, 1451-1461) --------------
// Environment strings exactly as getenv would return them (nullptr = unset).
struct Controls {
    const char* use_chmt = nullptr;     // MERTENS_USE_CHMT
    const char* cx = nullptr;           // MERTENS_CX
    const char* chmt_u_cap = nullptr;   // MERTENS_CHMT_U_CAP
    const char* fx = nullptr;           // MERTENS_FX
    const char* u_cap = nullptr;        // MERTENS_U_CAP
};
double choose_cx(i64 X, const Controls& c);
i64 choose_sieve_limit(i64 X, const Controls& c);
i64 choose_compressed_limit(i64 X, const Controls& c);
bool use_streaming(i64 X, const Controls& c);   // :1027-1029 with :1421-1435

struct Plan {
    bool streaming = false;
    i64 u = 0, u_dense = 0, mu_limit = 0, N = 0;
    double cx = 0;
};
Plan plan_for(i64 X, const Controls& c);

// ---- term evaluation --------------------------------------------------------
// Abstract M lookup so one evaluator serves both tables.
struct MLookup {
    const int16_t* dense = nullptr; i64 dense_limit = 0;
    const CompressedTable* comp = nullptr;
    int32_t operator()(i64 q) const { return q <= dense_limit ? dense[q] : comp->get_M(q); }
};
// G(y) = 1 - S1 + kappa*M(A) - S2 (mertens_dr.hpp:1081-1151, :1490-1525)
i128 g_single(i64 y, double cx, i64 u, const MLookup& M, const int8_t* mu);
// G(y) - G(floor(y/2)) (mertens_dr.hpp:1158-1354, :1527-1614)
i128 g_comb2(i64 y, double cx, i64 u, const MLookup& M, const int8_t* mu);
// Brute force G(y) = sum_{n <= y/u} M(y/n) using a full prefix table.
i128 g_brute(i64 y, i64 u, const std::vector<int32_t>& prefix);

struct Term { i64 k; int8_t mu_k; uint8_t range; };
// Outer terms k <= N, gcd(k,6)=1, mu(k) != 0 (mertens_dr.hpp:1061-1071)
std::vector<Term> outer_terms(i64 N, const int8_t* mu);
i128 term_value(const Term& t, i64 X, double cx, i64 u, const MLookup& M, const int8_t* mu);

// ---- entry points -------------------------------------------------------------
i64 mertens(i64 X, int threads, const Controls& c);        // compute_mertens :968
i64 mertens_compressed(i64 X, int threads, const Controls& c);  // :1442
i64 mertens_small(i64 X);                                   // linear odd sieve :982-1010
i64 mertens_recurrence(i64 X, i64 base_limit);              // test_sieve.cpp:30-41

}  // namespace mdr_oracle

#endif
