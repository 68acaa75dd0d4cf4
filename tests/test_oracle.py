"""Pins the CPU oracle (CPU only).

Each check compares the oracle against something independent of its own
formulas: the reference tests' golden values and cases, the compiled
reference's outputs (tests/golden/ref_values.json), textbook routines
(Eratosthenes, the quotient recurrence, brute-force sums) or numbers derived
independently from the reference's parameter formulas.
"""
import math
import random

import numpy as np
import pytest

import oracle


def test_isqrt_reference_cases():
    # test_sieve.cpp:44-51
    large = 3037000499
    cases = [(-1, 0), (0, 0), (1, 1), (2, 1), (3, 1), (4, 2), (15, 3), (16, 4), (17, 4),
             (9007199254740992, 94906265), (10**16, 10**8), (large * large - 1, large - 1),
             (large * large, large), (large * large + 1, large), (9223372036854775807, large)]
    for n, r in cases:
        assert oracle.isqrt(n) == r


def segmented_mobius(lo, hi):
    """mu(n) for lo <= n <= hi by trial division with the primes up to sqrt(hi)."""
    r = math.isqrt(hi)
    base = oracle.mobius(r)
    primes = [p for p in range(2, r + 1) if base[p] != 0 and all(p % q for q in range(2, math.isqrt(p) + 1))]
    rem = np.arange(lo, hi + 1, dtype=np.int64)
    mu = np.ones(hi - lo + 1, dtype=np.int64)
    for p in primes:
        start = (-lo) % p
        idx = np.arange(start, hi - lo + 1, p)
        mu[idx] *= -1
        rem[idx] //= p
        sq = (-lo) % (p * p)
        mu[np.arange(sq, hi - lo + 1, p * p)] = 0
    mu[rem > 1] *= -1
    return mu


def test_eratosthenes_small_known():
    mu = oracle.mobius(30)
    assert list(mu[1:13]) == [1, -1, -1, 0, -1, 1, -1, 0, 0, 1, -1, 0]
    assert mu[30] == -1


@pytest.mark.parametrize("low,length", [(1, 200000), (2**31 - 40000, 100000), (2**33 - 5000, 60000),
                                        (10**10, 50000), (2 * 10**10 - 30000, 60000), (10**11 + 1, 40000)])
def test_bytelog_block_matches_eratosthenes(low, length):
    # byte-log full-number sieve (mertens_dr.hpp:814-839) against trial division
    got = oracle.mobius_bytelog_block(low, length, low + length - 1)
    assert np.array_equal(got.astype(np.int64), segmented_mobius(low, low + length - 1))


@pytest.mark.parametrize("u", [1, 2, 3, 7, 31, 256, 257, 262145, 1000033])
def test_bytelog_odd_matches_eratosthenes(u):
    half = (u + 1) // 2
    odd = oracle.mobius_bytelog_odd(half, u, 3)
    mu = oracle.mobius(u)
    assert np.array_equal(odd[1:half + 1], mu[1:2 * half:2])


@pytest.fixture(scope="module")
def ref_mu():
    return oracle.mobius(1 << 20)


@pytest.fixture(scope="module")
def ref_prefix(ref_mu):
    return np.cumsum(ref_mu.astype(np.int64))


@pytest.mark.parametrize("limit", [0, 1, 2, 3, 4, 7, 31, 256, 257, 262145, 1000033])
def test_dense_table(limit, ref_mu, ref_prefix):
    # SieveTable against the prefix of the Eratosthenes mu (test_sieve.cpp:165-172)
    mu, M = oracle.dense_table(limit, 3)
    assert np.array_equal(mu, ref_mu[:limit + 1])
    assert np.array_equal(M.astype(np.int64), ref_prefix[:limit + 1])


@pytest.mark.parametrize("total", [0, 1, 63, 64, 65, 255, 256, 257, 13860, 17325, 131072, 131073, 262147, 1000033])
@pytest.mark.parametrize("dense", [0, 1, 64, 256, 1000, 131071])
def test_compressed_table(total, dense, ref_mu, ref_prefix):
    # CompressedMertensTable (test_sieve.cpp:79-96, :178-182)
    if dense > total:
        dense = total
    M, mu, lim = oracle.compressed_table(total, dense, 3, total + 17)
    assert lim == total
    assert np.array_equal(M.astype(np.int64), ref_prefix[:total + 1])
    assert np.array_equal(mu, ref_mu[:total + 1])


def test_pieces_closed_equal_reference_lut():
    # closed floor forms = the reference's LUT forms (mertens_dr.hpp:462-482)
    rng = random.Random(7)
    qs = list(range(5000)) + [rng.randrange(2**62) for _ in range(2000)] + [2**63 - 2 - i for i in range(24)]  # (q+1)>>1 overflows only at INT64_MAX
    for p in range(1, 13):
        for q in qs:
            assert oracle.piece_closed(p, q) == oracle.piece_lut(p, q), (p, q)


def py_piece(p, q):
    """Inclusion-exclusion over the multiples j, 2j, 3j, 6j (with y/2 halves)."""
    def s(q, m2, m3, m6):
        return q - (q // 2 if m2 else 0) - (q // 3 if m3 else 0) + (q // 6 if m6 else 0)
    single = {9: (1, 1, 1), 10: (1, 1, 0), 11: (1, 0, 0), 12: (0, 0, 0)}
    if p in single:
        return s(q, *single[p])
    # comb2 piece: which multiples stay in S2(y) and in S2(y/2)
    y_part = {1: (1, 1, 1), 2: (1, 1, 1), 3: (1, 1, 0), 4: (1, 1, 0), 5: (1, 0, 0), 6: (1, 0, 0), 7: (0, 0, 0), 8: (0, 0, 0)}
    h_part = {1: (1, 1, 1), 2: (1, 1, 0), 3: (1, 1, 0), 4: (1, 0, 0), 5: (1, 0, 0), 6: (0, 0, 0), 7: (0, 0, 0), 8: None}
    v = s(q, *y_part[p])
    if h_part[p] is not None:
        v -= s(q // 2, *h_part[p])
    return v


def test_pieces_from_inclusion_exclusion():
    rng = random.Random(3)
    for p in range(1, 13):
        for q in list(range(300)) + [rng.randrange(10**18) for _ in range(300)]:
            assert oracle.piece_closed(p, q) == py_piece(p, q)


@pytest.mark.parametrize("piece", range(1, 13))
def test_s2_interval_brute_force(piece, ref_mu):
    # test_sieve.cpp:98-116 grid against integer division
    for y in [37, 65535, 999999937, 9007199254740991, 9007199254740992, 9007199254740993, 10**16, 10**17]:
        for first in [1, 2, 5, 6, 7, 11, 12, 13, 97, 99991]:
            for count in [0, 1, 2, 5, 6, 7, 23, 24, 25, 57]:
                last = first + count - 1
                exp = sum(int(ref_mu[j]) * py_piece(piece, y // j)
                          for j in range(first, last + 1) if j % 2 and j % 3)
                assert oracle.s2_interval(piece, first, last, y, ref_mu) == exp


def test_g_terms_independent_of_split():
    # G(y) = sum_{n <= y/u} M(y/n) for every cx; comb2 = G(y) - G(y/2).
    # The brute force reads M up to y itself, so y stays within the prefix.
    lim = 20_000_000
    mu = oracle.mobius(lim)
    pre = np.cumsum(mu.astype(np.int64))
    M16 = pre.astype(np.int16)
    pre32 = pre.astype(np.int32)
    rng = random.Random(11)
    checked = 0
    for _ in range(200):
        u = rng.choice([5000, 40000, 300000])
        y = rng.randrange(u + 1, lim + 1)
        cx = rng.choice([0.3, 0.7, 0.95, 1.35, 2.0])
        if int(cx * math.sqrt(y)) + 1 >= u:
            continue
        g = oracle.g_brute(y, u, pre32)
        assert oracle.g_single(y, cx, u, M16, mu) == g
        assert oracle.g_comb2(y, cx, u, M16, mu) == g - oracle.g_brute(y // 2, u, pre32)
        checked += 1
    assert checked > 100


def test_mertens_known_values(known):
    for X, v in known.items():
        if X <= 10**13:
            assert oracle.mertens(X, 8) == v, X


def test_mertens_dense_strategy(known):
    for X in [10**9, 10**10, 10**11, 10**12]:
        assert oracle.mertens(X, 8, use_chmt="0") == known[X]


def test_mertens_small_paths():
    for X in list(range(0, 400)) + [99999, 100000]:
        assert oracle.mertens_small(X) == oracle.mertens(X, 2) == oracle.mertens_recurrence(X, 64)


def test_mertens_recurrence_boundaries():
    # test_sieve.cpp:183-187 list
    for X in [100000, 100001, 262143, 262144, 262145, 999983, 1000000, 49999999, 50000000,
              50000001, 50000003, 987654321, 10000000000]:
        assert oracle.mertens(X, 8) == oracle.mertens_recurrence(X, 1 << 22), X


def test_compressed_entry(ref_values):
    # test_sieve.cpp:205-214
    assert oracle.mertens_compressed(99975680, 3) == oracle.mertens_recurrence(99975680, 1 << 22)
    assert oracle.mertens_compressed(0, 0) == 0
    assert oracle.mertens_compressed(10**8, 3) == 1928
    assert oracle.mertens_compressed(10**10, 3) == -33722


def test_oracle_matches_compiled_reference(ref_values):
    # values computed by oracle/_ref/bench_dr (the reference, compiled here)
    rng = random.Random(5)
    xs = [x for x in ref_values if x <= 3 * 10**12]
    for X in rng.sample(xs, min(60, len(xs))):
        assert oracle.mertens(X, 8) == ref_values[X], X


def test_env_controls_do_not_change_results(known):
    X = 10**12
    for kw in [dict(cx="0.5"), dict(cx="2.5"), dict(cx="7"), dict(use_chmt="1"), dict(use_chmt="y", chmt_u_cap="100000000"),
               dict(fx="0.2"), dict(u_cap="100000000"), dict(use_chmt="no")]:
        assert oracle.mertens(X, 8, **kw) == known[X], kw


def test_plan_matches_independent_derivation():
    # parameters derived independently from the reference formulas
    p = oracle.plan(10**12, use_chmt="0")
    assert (p["streaming"], p["u"], p["N"], p["mu_limit"], p["cx"]) == (False, 14381928, 69531, 951000, 0.95)
    p = oracle.plan(10**15)
    assert (p["streaming"], p["u"], p["u_dense"], p["N"], p["mu_limit"]) == (True, 5379446016, 128000000, 185892, 42691748)
    p = oracle.plan(10**17)
    assert (p["u"], p["u_dense"], p["N"], p["mu_limit"], p["cx"]) == (8000000000, 426908672, 12500000, 426908484, 1.35)
    # strategy switch (mertens_dr.hpp:1027-1029, :1421-1435)
    assert oracle.plan(10**11)["streaming"] and not oracle.plan(10**11 - 1)["streaming"]
    assert not oracle.plan(10**12, fx="0.3")["streaming"]
    assert not oracle.plan(10**12, u_cap="2000000000")["streaming"]
    assert oracle.plan(10**12, fx="0.3", use_chmt="T")["streaming"]
    assert not oracle.plan(10**12, use_chmt="")["streaming"]
