"""The benchmark configurations, each run as specified and checked bit-exactly.

C1  M(1e12) with the dense strategy (MERTENS_USE_CHMT=0: dense int16 table,
    the reference's cutoff rule and cx = 0.95), README.md:183.
C2  M(1e15) with the streaming compressed table under the reference's
    parameter rules (u = 5,379,446,016, dense prefix 1.28e8, cx = 1.35) and
    the static partition, README.md:186.
C3  the 4096 seeded log-uniform X in [1e9, 1e14] in ONE hipGraph launch with
    the tuned shared table: the compiled reference's values where it has them
    (every X below 1e11 and the 100 largest), and every X against the eager
    static-partition batch (another launch mode and partition of the same work).
C4  M(1e17) with the learned (MLP) partition and the tuned split, under the
    dynamic chunk queue and under the static round-robin schedule, README.md:188;
    the learned partition must actually be applied, and the static partition
    gives the same integer.
"""
import pytest

import mdr_hip

pytestmark = pytest.mark.gpu


def test_c1_dense_1e12(engine, known):
    X = 10**12
    o = mdr_hip.make_opts(strategy=mdr_hip.STRATEGY_DENSE, no_table_reuse=1)
    assert engine.mertens(X, o) == known[X] == 62366
    st = engine.last_stats
    plan = engine.plan(X, o)
    assert not plan["streaming"] and plan["cx"] == 0.95
    assert st.U == plan["u"] and st.u_dense == st.U      # one dense int16 tier
    assert st.s1_comp_terms == 0


def test_c2_streaming_1e15_reference_parameters(engine, known):
    X = 10**15
    o = mdr_hip.make_opts(strategy=mdr_hip.STRATEGY_STREAMING, partition=mdr_hip.PARTITION_STATIC,
                          no_table_reuse=1)
    assert engine.mertens(X, o) == known[X] == -3216373
    st = engine.last_stats
    assert (st.U, st.cx) == (5379446016, 1.35)
    assert st.partition_used == mdr_hip.PARTITION_STATIC
    assert st.s1_comp_terms > 0


def test_c3_full_batch_one_graph_launch(engine, xlists, ref_values):
    xs = xlists["batch_4096_1e9_1e14_seed3"]
    assert len(xs) == 4096
    g = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, partition=mdr_hip.PARTITION_LEARNED, use_graph=1,
                          no_table_reuse=1)
    out = engine.batch(xs, g)
    st = engine.last_stats
    assert st.graph_used == 1
    checked = [i for i, x in enumerate(xs) if x in ref_values]
    assert len(checked) >= 1771
    bad = [xs[i] for i in checked if out[i] != ref_values[xs[i]]]
    assert not bad, bad[:5]
    eager = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, partition=mdr_hip.PARTITION_STATIC, no_table_reuse=1)
    assert engine.batch(xs, eager) == out
    # a second replay of the same graph gives the same values
    assert engine.batch(xs, g) == out
    assert engine.last_stats.graph_used == 1


@pytest.mark.parametrize("sched", [mdr_hip.SCHED_DYNAMIC, mdr_hip.SCHED_STATIC])
def test_c4_learned_partition_1e17(engine, known, sched):
    X = 10**17
    kw = dict(params=mdr_hip.PARAMS_TUNED, sched=sched, no_table_reuse=1, wave_span=1)
    if sched == mdr_hip.SCHED_STATIC:
        kw["chunk"] = engine.waves()   # one chunk per resident wave: no queue to absorb imbalance
    learned = engine.mertens(X, mdr_hip.make_opts(partition=mdr_hip.PARTITION_LEARNED, **kw))
    st = engine.last_stats
    assert learned == known[X] == -21830254
    assert st.partition_used == mdr_hip.PARTITION_LEARNED
    assert st.predicted_cycles > 0
    assert 0.0 < st.wave_busy_frac <= 1.0
    static = engine.mertens(X, mdr_hip.make_opts(partition=mdr_hip.PARTITION_STATIC, **kw))
    assert static == learned
    assert engine.last_stats.partition_used == mdr_hip.PARTITION_STATIC
    print(f"sched {sched}: learned wave busy {st.wave_busy_frac:.4f}, "
          f"static {engine.last_stats.wave_busy_frac:.4f}", flush=True)
