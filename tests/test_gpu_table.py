"""GPU table build parity: mu and M from the HIP sieve against Eratosthenes.

Reference behaviour: SieveTable / CompressedMertensTable get_M and get_mu
(mertens_dr.hpp:404-425, :707-716), pinned as in test_sieve.cpp:79-96 and
:165-182 -- every entry against the prefix of an independent mu.
"""
import math

import numpy as np
import pytest

import mdr_hip
import oracle
from test_oracle import segmented_mobius

pytestmark = pytest.mark.gpu

SEG = 49152
LIM = 1 << 21


@pytest.fixture(scope="module")
def mu_ref():
    return oracle.mobius(LIM)


@pytest.fixture(scope="module")
def pre_ref(mu_ref):
    return np.cumsum(mu_ref.astype(np.int64))


def cases():
    out = []
    for U in [1, 2, 3, 4, 63, 64, 65, 255, 256, 257, 13859, 13860, 13861, 17325, SEG - 1, SEG, SEG + 1,
              2 * SEG + 5, 277199, 277200, 277201, 1000033, LIM]:
        for ud in sorted({U, 0, 256, 4096, SEG, SEG + 256, (U // 512) * 256}):
            if ud > U or (ud < U and ud % 256):
                continue
            for ml in sorted({U, max(1, U // 3)}):
                out.append((U, ud, ml))
    return out


@pytest.mark.parametrize("U,ud,ml", cases())
def test_table_matches_eratosthenes(engine, mu_ref, pre_ref, U, ud, ml):
    print(f"table U={U} ud={ud} mu_limit={ml}", flush=True)
    engine.table_build(U, ud, ml)
    M = engine.read_M(0, U + 1).astype(np.int64)
    assert np.array_equal(M, pre_ref[:U + 1])
    mu = engine.read_mu(0, ml + 1)
    assert np.array_equal(mu, mu_ref[:ml + 1])


def test_large_table_windows(engine, known, ref_values):
    # A 2e10 table (two tiers, ~6 GB): mu windows against trial division,
    # M differences against mu, and M at the known / reference-computed points.
    U, ud = 20_000_000_000, 128_000_000
    print("large table", U, flush=True)
    engine.table_build(U, ud, 60_000_000)
    for X in [10**9, 10**10]:
        assert int(engine.read_M(X, 1)[0]) == known[X]
    pts = [x for x in ref_values if 10**8 <= x <= U][::25]
    for X in pts:
        assert int(engine.read_M(X, 1)[0]) == ref_values[X], X
    for lo in [127_999_000, 10**9 - 3000, 2**31 - 2000, 10**10, U - 40000]:
        n = 40000
        M = engine.read_M(lo - 1, n + 1).astype(np.int64)
        mu = segmented_mobius(lo, lo + n - 1)
        assert np.array_equal(np.diff(M), mu), lo


def test_table_mu_high_window(engine):
    # packed mu above 1e9 (mu is kept for j coprime to 6 and rebuilt for 2, 3)
    U = 1_200_000_000
    engine.table_build(U, U, U)
    lo, n = 1_100_000_000, 30000
    assert np.array_equal(engine.read_mu(lo, n).astype(np.int64), segmented_mobius(lo, lo + n - 1))


def test_read_beyond_table_is_an_error(engine):
    import mdr_hip
    engine.table_build(1000, 1000, 1000)
    with pytest.raises(mdr_hip.MdrError):
        engine.read_M(900, 200)


def test_compressed_tier_limit_is_an_error(engine):
    # the compressed tier holds at most 256 * (2^32 - 1) values (32-bit line
    # indices); larger requests fail before anything is allocated, and an
    # explicit cutoff above 1e12 is rejected as well
    ud = 1048576
    with pytest.raises(mdr_hip.MdrError, match="RANGE|compressed-tier"):
        engine.table_build(ud + 256 * 2**32, ud, 10**6)
    with pytest.raises(mdr_hip.MdrError, match="INVALID|tune_u"):
        engine.mertens(10**18, mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, tune_u=10**12 + 256,
                                                 tune_ud=ud, no_table_reuse=1))


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5, 6])
def test_sieve_shapes_build_the_same_table(engine, mu_ref, pre_ref, known, shape):
    # every sieve workgroup shape (segment length / threads) builds the same
    # table: a call leaves its table in the context, whose M and mu windows
    # are then checked against Eratosthenes, and M(X) against the known values
    for X in (10**12, 10**16):
        assert engine.mertens(X, params=mdr_hip.PARAMS_TUNED, sieve_shape=shape, no_table_reuse=1) == known[X]
    st = engine.last_stats
    assert st.U > LIM
    M = engine.read_M(0, LIM + 1).astype(np.int64)
    assert np.array_equal(M, pre_ref[:LIM + 1])
    ml = min(st.mu_limit, LIM)
    assert np.array_equal(engine.read_mu(0, ml + 1), mu_ref[:ml + 1])
    hi = st.U - 40000   # the top of the table: M differences against trial division
    Mh = engine.read_M(hi - 1, 40001).astype(np.int64)
    assert np.array_equal(np.diff(Mh), segmented_mobius(hi, hi + 39999))


def test_sieve_shape_out_of_range(engine):
    with pytest.raises(mdr_hip.MdrError):
        engine.mertens(10**12, sieve_shape=7)


def test_segment_prefixes_across_a_large_table(engine, ref_values):
    # the offset each segment gets from the scan of the segment sums: M at
    # points spread over a 6e10 table against the compiled reference's values,
    # and M differences against trial division across segment edges
    U, ud = 60_000_000_000, 1048576
    engine.table_build(U, ud, 10**6)
    pts = [x for x in ref_values if x <= U][::9]
    assert len(pts) > 20
    for X in pts:
        assert int(engine.read_M(X, 1)[0]) == ref_values[X], X
    for k in [1, 2, 777, 123_457, U // SEG - 1]:
        lo = k * SEG - 5000
        M = engine.read_M(lo - 1, 10001).astype(np.int64)
        assert np.array_equal(np.diff(M), segmented_mobius(lo, lo + 9999)), k
