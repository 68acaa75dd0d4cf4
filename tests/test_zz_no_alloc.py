"""The allocation-free entry point (mdr_mertens_noalloc) against mdr_mertens.

Both go through the C ABI.  Two calls in a row with the second X smaller
(the scratch array keeps the first call's contents, so the second call must
clear and rebuild what it reads), then the three plan kinds (constant, table
read, Deleglise-Rivat sum), the known values, and the size limit: an X whose
arrays do not fit the scratch array fails with MDR_ERR_RANGE.  Bit-exact
everywhere.
"""
import random

import pytest

import mdr_hip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = mdr_hip.Engine(0)
    yield e
    e.close()


def test_two_calls_second_smaller(eng, known):
    for X in (10**15, 123456789012):
        print("noalloc", X, flush=True)
        got = eng.mertens_noalloc(X)
        st = eng.last_stats
        want = eng.mertens(X, params=mdr_hip.PARAMS_TUNED, no_table_reuse=1)
        assert got == want, X
        if X in known:
            assert got == known[X]
        assert st.U > 0 and st.table_ms > 0


def test_plan_kinds_and_known_values(eng, known):
    for X in [-3, 0, 1, 2, 3, 4, 1000, 49999999, 50000000, 50000001, 10**8, 10**10, 10**12, 10**13, 10**14]:
        got = eng.mertens_noalloc(X)
        assert got == eng.mertens(X), X
        if X in known:
            assert got == known[X], X


def test_random_points_match_the_allocating_call(eng):
    rng = random.Random(7)
    xs = [rng.randrange(10**9, 8 * 10**15) for _ in range(6)]
    for X in sorted(xs, reverse=True):
        assert eng.mertens_noalloc(X) == eng.mertens(X, params=mdr_hip.PARAMS_TUNED), X


def test_beyond_the_scratch_array_is_an_error(eng, known):
    with pytest.raises(mdr_hip.MdrError, match="RANGE|scratch"):
        eng.mertens_noalloc(10**17)
    # the context still works afterwards, both ways
    assert eng.mertens_noalloc(10**11) == known[10**11]
    assert eng.mertens(10**11) == known[10**11]
