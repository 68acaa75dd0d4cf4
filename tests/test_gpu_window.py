"""Windowed tables on the GPU: the compressed tier streamed through windows.

Reference behaviour: compute_mertens_windowed (mertens_dr_windowed.hpp:49-407)
keeps mu and M up to mu_limit resident and streams the M table above it in
windows; its result is the one-table result (the header states bit-identical
S1/S2 math).  Here every windowed M(X) is checked against the known values,
the compiled reference's values and the same engine's one-table result, over
window sizes that cut the compressed tier into one to a few dozen windows,
both partitions and schedules, every sieve shape and a cutoff beyond the
one-table compressed-tier limit.
"""
import random

import pytest

import mdr_hip

pytestmark = pytest.mark.gpu


def opts(**kw):
    kw.setdefault("no_table_reuse", 1)
    return mdr_hip.make_opts(**kw)


@pytest.mark.parametrize("X", [10**12, 10**13, 10**15, 10**16])
def test_windowed_known_values(engine, known, X):
    one = engine.mertens(X, opts(params=mdr_hip.PARAMS_TUNED))
    assert one == known[X]
    U = engine.last_stats.U
    for nwin in (1, 3, 17):
        w = max(1, U // nwin)
        print("windowed", X, "window", w, flush=True)
        assert engine.mertens(X, opts(params=mdr_hip.PARAMS_TUNED, window=w)) == known[X]
        st = engine.last_stats
        assert st.windows >= 1
        assert st.U == U


def test_windowed_reference_parameters(engine, ref_values):
    # the reference's own parameter rules (dense and streaming strategies)
    rng = random.Random(17)
    xs = rng.sample(sorted(x for x in ref_values if x >= 10**9), 12)
    for X in xs:
        for strategy in (mdr_hip.STRATEGY_AUTO, mdr_hip.STRATEGY_STREAMING, mdr_hip.STRATEGY_DENSE):
            print("windowed ref", X, strategy, flush=True)
            assert engine.mertens(X, opts(strategy=strategy, window=10**7)) == ref_values[X], (X, strategy)


def test_windowed_partitions_schedules_and_shapes(engine, known):
    X = 10**15
    for part in (mdr_hip.PARTITION_STATIC, mdr_hip.PARTITION_LEARNED, mdr_hip.PARTITION_TERMS):
        for sched in (mdr_hip.SCHED_DYNAMIC, mdr_hip.SCHED_STATIC):
            assert engine.mertens(X, opts(params=mdr_hip.PARAMS_TUNED, partition=part, sched=sched,
                                          window=3 * 10**8)) == known[X]
    for shape in (1, 2, 3, 4, 5, 6):
        print("windowed shape", shape, flush=True)
        assert engine.mertens(X, opts(params=mdr_hip.PARAMS_TUNED, sieve_shape=shape, window=2 * 10**8)) == known[X]


def test_windowed_random_points_match_one_table(engine):
    # the same engine with and without windows, at cutoffs and splits drawn at
    # random (window edges fall at arbitrary quotients)
    rng = random.Random(23)
    for _ in range(10):
        X = int(10 ** rng.uniform(11, 15.5))
        cx = rng.choice([0.7, 1.35, 2.0, 2.8])
        u = int(X ** (2 / 3) * rng.uniform(0.3, 2.0))
        o = dict(params=mdr_hip.PARAMS_TUNED, tune_u=u, tune_cx=cx, tune_ud=rng.choice([0, 65536, 1048576]))
        one = engine.mertens(X, opts(**o))
        w = rng.randrange(49152, max(49153, u // 2))
        print("windowed random", X, u, cx, w, flush=True)
        assert engine.mertens(X, opts(window=w, **o)) == one, (X, u, cx, w)


def test_windowed_batch(engine, ref_values):
    xs = sorted(x for x in ref_values if x >= 10**10)[::7]
    out = engine.batch(xs, opts(params=mdr_hip.PARAMS_TUNED, window=10**9))
    assert out == [ref_values[x] for x in xs]


def test_windowed_cutoff_beyond_the_one_table_limit(engine):
    # u = 1e12 is 250 GB of compressed lines as one table; windows of 2e11
    # hold 50 GB at a time.  M(1e18) = -46758740 is the external value (OEIS
    # A084237) the one-table path also reproduces
    X = 10**18
    o = dict(params=mdr_hip.PARAMS_TUNED, tune_u=10**12, tune_cx=2.8)
    assert engine.mertens(X, opts(window=2 * 10**11, **o)) == -46758740
    st = engine.last_stats
    assert st.windows == 5 and st.U >= 10**12


def test_windows_sieve_with_primes_up_to_the_full_cutoff():
    # a fresh context whose first table is windowed: the windows above the
    # resident part need sieving primes up to sqrt(U), not sqrt(U0)
    with mdr_hip.Engine(0) as e:
        X = 10**16
        u = 4 * 10**10
        o = opts(params=mdr_hip.PARAMS_TUNED, tune_u=u, tune_cx=2.0, window=10**10)
        assert e.mertens(X, o) == -3195437
        assert e.last_stats.windows == 4


def test_windowed_call_runs_eagerly_and_logs_cycles(engine, known):
    # a graph request is run eagerly when the table is windowed; the cycle log
    # collects every pass
    X = 10**15
    o = opts(params=mdr_hip.PARAMS_TUNED, window=5 * 10**8, use_graph=1, record_cycles=8)
    assert engine.mertens(X, o) == known[X]
    st = engine.last_stats
    assert st.windows >= 2 and st.graph_used == 0
    kinds = {s.kind for s in engine.cycle_log()}
    assert {2, 3} & kinds and any(k >= 4 for k in kinds)   # compressed S1 and S2 runs


def test_negative_window_is_rejected(engine):
    with pytest.raises(mdr_hip.MdrError):
        engine.mertens(10**12, opts(window=-1))


def test_windowed_batch_with_small_and_constant_x(engine):
    # a tuned batch whose shared cutoff exceeds the one-table cap is windowed;
    # its constant (X <= 3) and read-from-table (X <= 5e7) members live in the
    # resident part
    import oracle
    xs = [3, 10**7 + 19, 10**18, 9 * 10**17 + 7]
    out = engine.batch(xs, mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, no_table_reuse=1))
    assert engine.last_stats.windows >= 1
    assert out[0] == -1
    assert out[1] == oracle.mertens(xs[1], 8)
    assert out[2] == -46758740
    assert out[3] == engine.mertens(xs[3], mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, no_table_reuse=1))
