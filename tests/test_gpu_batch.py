"""Batched and table-sharing plans on the GPU (tuned parameters).

A tuned batch builds one table whose cutoff is chosen for the whole call
(every X above it is summed with u = that cutoff, every X below it is read
from the table), and a tuned call whose cached table already covers it reuses
that table with the cached cutoff.  M(X) does not depend on u, so every case
is checked bit-exactly against the compiled reference's values; the extreme
plans this creates (u just below X, so N = X / u is 1, 2, 3 ...) are checked
directly as well.
"""
import random

import pytest

import mdr_hip
import oracle

pytestmark = pytest.mark.gpu


def tuned(**kw):
    return mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, **kw)


def test_tuned_batch_of_all_reference_values(engine, ref_values):
    xs = sorted(ref_values)
    out = engine.batch(xs, tuned(no_table_reuse=1))
    assert out == [ref_values[x] for x in xs]
    st = engine.last_stats
    # the shared cutoff is far above the single-X cutoff of the largest X
    single = engine.plan(max(xs), tuned())
    assert st.U > single["u"]


def test_tuned_batch_sample_of_the_4096_workload(engine, ref_values, xlists):
    xs = xlists["batch_4096_1e9_1e14_seed3"]
    sub = random.Random(5).sample(xs, 600)
    for graph in (0, 1):
        out = engine.batch(sub, tuned(no_table_reuse=1, use_graph=graph,
                                      partition=mdr_hip.PARTITION_LEARNED))
        want = [ref_values[x] if x in ref_values else None for x in sub]
        assert all(w is None or o == w for o, w in zip(out, want))
        # X without a reference value: the same X alone with reference parameters
        rest = [i for i, w in enumerate(want) if w is None][:12]
        for i in rest:
            assert out[i] == engine.mertens(sub[i], no_table_reuse=1), sub[i]


@pytest.mark.parametrize("X", [10**9 + 7, 3 * 10**10 + 1, 999999999989, 10**12])
def test_cutoff_just_below_x(engine, X):
    want = oracle.mertens(X, 8) if X <= 10**10 else None
    base = engine.mertens(X, no_table_reuse=1)
    if want is not None:
        assert base == want
    for k in (1, 2, 3, 4, 5, 6, 7, 11, 13, 36, 97):
        u = X // k
        if u > 2 * 10**11 or u * u < 9 * X:
            continue
        r = engine.mertens(X, tuned(tune_u=u, no_table_reuse=1))
        assert r == base, (X, k, engine.last_stats.N)


def test_covering_table_reuse(engine, known, ref_values):
    # build a large table, then evaluate smaller X against it
    assert engine.mertens(10**16, tuned()) == known[10**16]
    big_u = engine.last_stats.U
    xs = sorted(x for x in ref_values if 10**9 <= x <= 10**15)
    for x in random.Random(7).sample(xs, 40):
        assert engine.mertens(x, tuned()) == ref_values[x], x
        st = engine.last_stats
        assert st.table_reused == 1 and st.U == big_u
    # reference parameters never take the covering path
    x = xs[0]
    assert engine.mertens(x) == ref_values[x]
    assert engine.last_stats.U != big_u


def test_lookup_only_batches(engine, ref_values):
    # a batch whose X all fall below a cached cutoff is answered by lookups
    assert engine.mertens(10**15, tuned()) is not None
    U = engine.last_stats.U
    xs = sorted(x for x in ref_values if 10**6 <= x <= U)[:300]
    out = engine.batch(xs, tuned())
    assert out == [ref_values[x] for x in xs]
    assert engine.last_stats.s1_terms == 0 and engine.last_stats.s2_terms == 0


def test_large_mixed_batch_learned_equals_static(engine, xlists, ref_values):
    # 64 X log-uniform in [1e16, 1e18] in one call (a shared cutoff of 1e12,
    # above the one-table cap, so the compressed tier is windowed): the
    # learned partition and the static (equal-positions) partition give the
    # same 64 values, and the values below 1.3e16 match the compiled reference
    xs = xlists["large_64_1e16_1e18_seed5"]
    print("large batch, learned partition", flush=True)
    learned = engine.batch(xs, tuned(no_table_reuse=1, partition=mdr_hip.PARTITION_LEARNED))
    assert engine.last_stats.partition_used == mdr_hip.PARTITION_LEARNED
    assert engine.last_stats.windows >= 2 and engine.last_stats.U > 5 * 10**11   # above the one-table cap
    print("large batch, static partition", flush=True)
    static = engine.batch(xs, tuned(partition=mdr_hip.PARTITION_STATIC, sched=mdr_hip.SCHED_STATIC))
    assert engine.last_stats.partition_used == mdr_hip.PARTITION_STATIC
    assert learned == static
    checked = [i for i, x in enumerate(xs) if x in ref_values]
    assert len(checked) >= 4
    assert all(learned[i] == ref_values[xs[i]] for i in checked)


def test_batch_with_an_x_below_the_shared_cutoff_stays_one_table(engine, ref_values):
    # the shared cutoff of these X is above the one-table cap, but an X at or
    # below it would become a lookup outside the resident part of a windowed
    # table: the batch keeps one table at the capped cutoff
    small = max(x for x in ref_values if x <= 2 * 10**11)
    xs = [small, 10**18, 9 * 10**17 + 7]
    out = engine.batch(xs, tuned(no_table_reuse=1))
    st = engine.last_stats
    assert st.windows == 0 and st.U <= 5 * 10**11   # the one-table cap
    assert out[0] == ref_values[small]
    assert out[1] == -46758740
    assert out[2] == engine.mertens(xs[2], tuned(no_table_reuse=1))


def test_batch_of_mixed_shapes(engine, known):
    # duplicates, unsorted order, constants (X <= 3 and below 1), the small
    # sieve-and-read path and summed X in one call: every entry equals its own
    # single call, under the reference's rules and the tuned split, eager and
    # as a graph launch; an empty batch is a no-op
    xs = [10**15, -3, 0, 1, 2, 3, 10**15, 12345, 10**12, 10**15 + 7, 5 * 10**7, 10**12, 4, 10**13]
    single = [engine.mertens(x) for x in xs]
    assert single[0] == known[10**15] and single[8] == known[10**12] and single[1:4] == [0, 0, 1]
    for kw in ({}, {"params": mdr_hip.PARAMS_TUNED}, {"params": mdr_hip.PARAMS_TUNED, "use_graph": 1},
               {"partition": mdr_hip.PARTITION_LEARNED, "no_table_reuse": 1}):
        print("mixed batch", kw, flush=True)
        assert engine.batch(xs, **kw) == single, kw
    assert engine.batch([]) == []
