"""The hipGraph launch path (mdr_opts.use_graph = 1).

A call's launches (clears, table build, lookups, partition, quotient kernel)
are captured into one graph, instantiated once per launch sequence and
updated in place for later calls.  Every replay is checked bit-exactly
against the compiled reference's values, with eager calls of other shapes
between the replays, as the table build and the counters are written by
kernels inside the graph.
"""
import random

import pytest

import mdr_hip
import oracle

pytestmark = pytest.mark.gpu


def sample(ref_values, n, seed, lo=10**6, hi=10**16):
    xs = sorted(x for x in ref_values if lo <= x <= hi)
    return random.Random(seed).sample(xs, n)


def test_graph_batches_with_eager_work_between(engine, ref_values, known):
    a = sample(ref_values, 48, 1) + [1, 2, 3, 100, 10**5]      # sums, lookups and constants
    b = sample(ref_values, 48, 2) + [4, 5, 6, 1000, 2 * 10**5]
    ga = mdr_hip.make_opts(use_graph=1, params=mdr_hip.PARAMS_TUNED)
    eager = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED)
    inst = []
    for rep in range(4):
        for xs in (a, b, a):
            print("graph batch", rep, xs[:2], flush=True)
            out = engine.batch(xs, ga)
            assert out == [ref_values[x] if x in ref_values else oracle.mertens(x, 1) for x in xs]
            assert engine.last_stats.graph_used == 1
            inst.append(engine.last_stats.graph_instantiated)
        # eager work of another shape between the replays
        assert engine.mertens(10**13, eager) == known[10**13]
        assert engine.last_stats.graph_used == 0
    # graphs are instantiated once per launch sequence, then updated in place
    assert sum(inst) <= 2, inst


def test_graph_rebuilds_table_inside_the_graph(engine, known):
    X = 10**15
    g = mdr_hip.make_opts(use_graph=1, no_table_reuse=1, params=mdr_hip.PARAMS_TUNED)
    for rep in range(5):
        assert engine.mertens(X, g) == known[X], rep
        st = engine.last_stats
        assert st.graph_used == 1 and st.table_reused == 0 and st.graph_ms > 0
        assert engine.mertens(10**12) == known[10**12]


def test_graph_modes_match_eager(engine, ref_values):
    xs = sample(ref_values, 64, 3, lo=10**9)
    want = [ref_values[x] for x in xs]
    for kw in [dict(partition=mdr_hip.PARTITION_LEARNED), dict(sched=mdr_hip.SCHED_STATIC),
               dict(partition=mdr_hip.PARTITION_LEARNED, sched=mdr_hip.SCHED_STATIC, chunk=1000),
               dict(strategy=mdr_hip.STRATEGY_FORCE_STREAMING), dict(record_cycles=4)]:
        for rep in range(3):
            print("graph mode", kw, rep, flush=True)
            assert engine.batch(xs, mdr_hip.make_opts(use_graph=1, **kw)) == want, (kw, rep)
            assert engine.batch(xs[:7], mdr_hip.make_opts(**kw)) == want[:7]


def test_graph_cycle_log_matches_eager(engine):
    X = 10**14
    o = dict(record_cycles=1, partition=mdr_hip.PARTITION_STATIC, sched=mdr_hip.SCHED_STATIC)
    engine.mertens(X, mdr_hip.make_opts(**o))
    n_eager = len(engine.cycle_log())
    engine.mertens(X, mdr_hip.make_opts(use_graph=1, **o))
    n_graph = len(engine.cycle_log())
    assert n_eager == n_graph > 0


def test_use_graph_out_of_range_rejected(engine):
    with pytest.raises(mdr_hip.MdrError):
        engine.mertens(10**12, mdr_hip.make_opts(use_graph=2))
