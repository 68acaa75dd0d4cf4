"""Table objects on the GPU (mdr_table_create and its reads), bit-exact.

Reference behaviour: SieveTable / CompressedMertensTable construction and
get_M / get_mu / data (mertens_dr.hpp:376-457, :675-752), and the compressed
tier's record layout -- 256-value chunks {nonzero[32], sign[32]} with four
int32 checkpoints, M just before each 64-value word (:670-687, :741-752).
Every M and mu is checked against Eratosthenes' mu and its prefix sums; the
chunk records against the bit definition of the format.  Several tables live
at once beside the context's cached table.  The C++ drop-in program repeats
test_sieve.cpp:79-96 and :165-204 through the drop-in classes.
"""
import os
import subprocess

import numpy as np
import pytest

import mdr_hip
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LIM = 1 << 21


@pytest.fixture(scope="module")
def mu_ref():
    return oracle.mobius(LIM)


@pytest.fixture(scope="module")
def pre_ref(mu_ref):
    return np.cumsum(mu_ref.astype(np.int64))


def expected_chunks(mu, pre, U, ref_ud, c0, count):
    """Chunk records and checkpoints of the compressed tier above ref_ud."""
    ch = np.zeros((count, 64), dtype=np.uint8)
    ck = np.zeros(count * 4, dtype=np.int32)
    for c in range(count):
        for k in range(4):
            n0 = ref_ud + 1 + 256 * (c0 + c) + 64 * k
            if n0 <= U:
                ck[4 * c + k] = pre[n0 - 1]
            for i in range(64):
                n = n0 + i
                if n > U or mu[n] == 0:
                    continue
                bit = 64 * k + i
                ch[c, bit // 8] |= 1 << (bit % 8)
                if mu[n] < 0:
                    ch[c, 32 + bit // 8] |= 1 << (bit % 8)
    return ch, ck


@pytest.mark.parametrize("U,ud,ml", [(0, 0, 0), (1, 0, 1), (1, 1, 1), (63, 7, 63), (257, 0, 100), (13861, 1000, 13861),
                                     (49153, 4095, 20000), (262147, 131071, 262147), (LIM, 1024, LIM)])
def test_table_object_matches_eratosthenes(engine, mu_ref, pre_ref, U, ud, ml):
    with mdr_hip.Table(engine, U, ud, ml) as t:
        got_U, got_ud, got_mu = t.info()
        assert (got_U, got_mu) == (U, ml) and got_ud <= max(ud, 0)
        M = t.M(0, U + 1)
        assert np.array_equal(M, pre_ref[:U + 1]) and M[0] == 0
        mu = t.mu(0, U + 2)
        want = np.zeros(U + 2, dtype=np.int8)
        want[:ml + 1] = mu_ref[:ml + 1]
        assert np.array_equal(mu, want)
        rng = np.random.default_rng(U)
        qs = rng.integers(0, U + 1, size=min(5000, U + 1))
        assert np.array_equal(t.gather_M(qs), pre_ref[qs])
        if ud < U:
            nch = (U - ud + 255) // 256
            for c0, cnt in {(0, min(nch, 40)), (max(0, nch - 3), min(nch, 3))}:
                ch, ck = t.chunks(ud, c0, cnt)
                ech, eck = expected_chunks(mu_ref, pre_ref, U, ud, c0, cnt)
                assert np.array_equal(ch, ech), (c0, cnt)
                assert np.array_equal(ck, eck), (c0, cnt)


def test_table_objects_coexist_with_the_cached_table(engine, pre_ref, known):
    a = mdr_hip.Table(engine, 1000033, 256 * 100, 1000033)
    b = mdr_hip.Table(engine, 262147, 262147, 5000)
    try:
        assert engine.mertens(10**14, no_table_reuse=1) == known[10**14]   # rebuilds the context's table
        assert np.array_equal(a.M(999000, 1034), pre_ref[999000:1000034])
        assert np.array_equal(b.M(0, 262148), pre_ref[:262148])
    finally:
        a.close()
        b.close()
    assert engine.mertens(10**12) == known[10**12]


def test_table_object_errors(engine):
    with pytest.raises(mdr_hip.MdrError):
        mdr_hip.Table(engine, 100, 0, 101)          # mu limit above U
    with mdr_hip.Table(engine, 1000, 256, 1000) as t:
        with pytest.raises(mdr_hip.MdrError):
            t.M(990, 20)                              # beyond U
        with pytest.raises(mdr_hip.MdrError):
            t.gather_M([5, 1001])
        with pytest.raises(mdr_hip.MdrError):
            t.chunks(100, 0, 1)                       # below the table's own dense tier
        # ranges whose end does not fit in 64 bits are rejected, not wrapped
        with pytest.raises(mdr_hip.MdrError):
            t.M(2**63 - 10, 20)
        with pytest.raises(mdr_hip.MdrError):
            t.mu(2**63 - 5, 10)
        with pytest.raises(mdr_hip.MdrError):
            t.chunks(256, 2**62, 1)


def test_dropin_tables_cpp_program():
    exe = os.path.join(ROOT, "tests", "cpp", "test_dropin_tables")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
