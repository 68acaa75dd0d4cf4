"""GPU quotient-phase parity: S2 piece intervals, S1 table sums and whole G
evaluations against the oracle (and brute force), on the same inputs.

Reference: run_s2_fast<Piece> (mertens_dr.hpp:484-617, pinned as in
test_sieve.cpp:98-116), sum_dense_quotients (:88-128, test_sieve.cpp:118-138)
and eval_single_S / eval_comb2 (:1081-1354).
"""
import math
import random

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

LIM = 1 << 21


@pytest.fixture(scope="module")
def mu_ref():
    return oracle.mobius(LIM)


@pytest.fixture(scope="module")
def pre_ref(mu_ref):
    return np.cumsum(mu_ref.astype(np.int64))


@pytest.fixture(scope="module")
def table(engine):
    # two tiers: dense up to 4096, compressed above
    engine.table_build(LIM, 4096, LIM)
    return engine


Y_GRID = [37, 65535, 999999937, 9007199254740991, 9007199254740992, 9007199254740993,
          10**16, 10**17, 10**18, 9 * 10**18, 2**63 - 1]


@pytest.mark.parametrize("piece", range(1, 13))
def test_s2_interval_grid(table, mu_ref, piece):
    print("s2 piece", piece, flush=True)
    for y in Y_GRID:
        for first in [1, 2, 5, 6, 7, 11, 12, 13, 97, 99991]:
            for count in [0, 1, 2, 5, 6, 7, 23, 24, 25, 57]:
                last = first + count - 1
                assert table.s2_interval(piece, first, last, y) == oracle.s2_interval(piece, first, last, y, mu_ref), \
                    (piece, y, first, last)


@pytest.mark.parametrize("piece", range(1, 13))
def test_s2_interval_long_ranges(table, mu_ref, piece):
    rng = random.Random(piece)
    for _ in range(6):
        y = rng.choice([rng.randrange(10**6, 10**12), rng.randrange(10**12, 2**53), rng.randrange(2**53, 2**63)])
        first = rng.randrange(1, 1000)
        last = rng.randrange(first, LIM)
        assert table.s2_interval(piece, first, last, y) == oracle.s2_interval(piece, first, last, y, mu_ref)


def test_s1_sums_both_tiers(table, pre_ref):
    rng = random.Random(9)
    for _ in range(60):
        stride = rng.choice([1, 2])
        y = rng.choice([rng.randrange(LIM, 10**9), rng.randrange(10**9, 2**53), rng.randrange(2**53, 2**62)])
        first = max(1, y // LIM + 1 + rng.randrange(0, 50))
        last = first + rng.randrange(0, 300000)
        n = np.arange(first, last + 1, stride, dtype=object)
        q = [y // int(v) for v in n]
        exp = int(sum(int(pre_ref[v]) for v in q))
        assert table.s1_sum(y, first, last, stride) == exp, (y, first, last, stride)


def test_g_eval_matches_oracle(table, mu_ref, pre_ref):
    M16 = pre_ref.astype(np.int16)
    rng = random.Random(21)
    n = 0
    for _ in range(300):
        u = rng.choice([20000, 300000, LIM])
        y = rng.randrange(u + 1, min(u * u // 16, 10**13))
        cx = rng.choice([0.3, 0.7, 0.95, 1.35, 2.2])
        if int(cx * math.sqrt(y)) + 1 >= u:
            continue
        for comb in (0, 1):
            want = (oracle.g_comb2 if comb else oracle.g_single)(y, cx, u, M16, mu_ref)
            assert table.g_eval(y, comb, u, cx) == want, (y, comb, u, cx)
        n += 1
    assert n > 150


def test_dense_only_table_g_eval(engine, mu_ref, pre_ref):
    engine.table_build(LIM, LIM, LIM)
    M16 = pre_ref.astype(np.int16)
    for y, u, cx in [(10**9, 200000, 0.95), (10**11, LIM, 1.35), (987654321987, LIM, 1.15)]:
        for comb in (0, 1):
            assert engine.g_eval(y, comb, u, cx) == (oracle.g_comb2 if comb else oracle.g_single)(y, cx, u, M16, mu_ref)


def lcm_upto(k):
    v = 1
    for i in range(2, k + 1):
        v = v * i // math.gcd(v, i)
    return v


# y with very many divisors (remainder 0 at most n), and y -+ 1 beside them
# (remainders n - 1 and 1): the points where an estimated quotient is off by one
DIV_Y = []
for base in [lcm_upto(20), lcm_upto(30), lcm_upto(40), lcm_upto(42), lcm_upto(42) * 42, 720720 * 2**40]:
    DIV_Y += [base - 1, base, base + 1]


@pytest.mark.parametrize("piece", [1, 4, 7, 9, 12])
def test_s2_interval_exact_division_edges(table, mu_ref, piece):
    for y in DIV_Y:
        assert y < 2**63
        for first, last in [(1, 4000), (1, LIM), (LIM // 2, LIM)]:
            assert table.s2_interval(piece, first, last, y) == oracle.s2_interval(piece, first, last, y, mu_ref), \
                (piece, y, first, last)


def test_s1_sums_exact_division_edges(table, pre_ref):
    # S1 reads M(y/n) <= LIM from this table, so y/LIM <= n <= sqrt(y) needs y <= LIM^2
    ran = 0
    for y in DIV_Y:
        first = y // LIM + 1
        last = min(first + 200000, math.isqrt(y))
        if last < first:
            continue
        for stride in (1, 2):
            n = np.arange(first, last + 1, stride, dtype=object)
            exp = int(sum(int(pre_ref[y // int(v)]) for v in n))
            assert table.s1_sum(y, first, last, stride) == exp, (y, first, last, stride)
            ran += 1
    assert ran >= 12
