"""Cost model (the learned partition's MLP).

CPU: the numpy oracle's backpropagation is pinned by central finite
differences, and the engine's host feature code equals the oracle's.
GPU: the hand-written forward and backward kernels match the oracle,
training lowers the loss, and the learned partition gives bit-identical
M(X) to the static one (also with a nonsense model, and falls back to
static without one).
"""
import ctypes

import numpy as np
import pytest

import mdr_hip
import mlp_oracle as mo


def rand_weights(seed):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.4, mo.NPARAM).astype(np.float32)
    return w


def rand_data(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, (n, mo.IN)).astype(np.float32)
    T = rng.normal(4, 2, n).astype(np.float32)
    return X, T


def test_oracle_gradient_matches_finite_differences():
    w = rand_weights(1).astype(np.float64)
    X, T = rand_data(2, 64)
    _, g = mo.loss_and_grad(w, X, T)
    rng = np.random.default_rng(3)
    for p in rng.choice(mo.NPARAM, 80, replace=False):
        e = np.zeros(mo.NPARAM)
        e[p] = 1e-6
        num = (mo.loss_and_grad(w + e, X, T)[0] - mo.loss_and_grad(w - e, X, T)[0]) / 2e-6
        assert abs(num - g[p]) <= 1e-5 + 1e-4 * abs(num), p


def test_engine_features_match_oracle():
    rng = np.random.default_rng(4)
    n = 500
    st = (mdr_hip.Stamp * n)()
    ys, ss, es, ks = [], [], [], []
    for i in range(n):
        y = int(rng.integers(10**6, 10**18))
        s = int(rng.integers(1, 10**6))
        e = s + int(rng.integers(0, 10**6))
        k = int(rng.integers(0, 16))
        st[i].y, st[i].start, st[i].end, st[i].kind, st[i].cycles, st[i].count = y, s, e, k, 1000 + i, 10 + i
        ys.append(y), ss.append(s), es.append(e), ks.append(k)
    f, t = mdr_hip.mlp_features(st)
    assert np.allclose(f, mo.features(ys, ss, es, ks), rtol=1e-6, atol=1e-7)
    assert np.allclose(t, np.log2((1000 + np.arange(n)) / (10 + np.arange(n))), rtol=1e-5)


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(engine):
    w = rand_weights(5)
    X, _ = rand_data(6, 3000)
    got = engine.mlp_predict(w, X)
    assert np.allclose(got, mo.forward(w, X), rtol=1e-4, atol=1e-4)


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle(engine):
    # fp32 kernels against the float64 oracle: rtol 1e-3, atol 1e-5; batch
    # sizes that fill every workgroup (2048) and that leave the last one
    # partly empty (3001, 5: the forward kernel takes 8 samples per workgroup)
    for seed, n in ((0, 2048), (1, 3001), (2, 5)):
        w = rand_weights(10 + seed)
        X, T = rand_data(20 + seed, n)
        g, loss = engine.mlp_grad(w, X, T)
        l_ref, g_ref = mo.loss_and_grad(w, X, T)
        assert abs(loss - l_ref) <= 1e-3 * abs(l_ref) + 1e-5
        assert np.allclose(g, g_ref, rtol=1e-3, atol=1e-5)


@pytest.mark.gpu
def test_gpu_gradient_is_deterministic(engine):
    # one workgroup per parameter with a fixed summation order: the same batch
    # gives the same bits, also for a batch that is not a multiple of 256
    w = rand_weights(30)
    for n in (4096, 3001):
        X, T = rand_data(31, n)
        g1, l1 = engine.mlp_grad(w, X, T)
        g2, l2 = engine.mlp_grad(w, X, T)
        assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32)) and l1 == l2


@pytest.mark.gpu
def test_gpu_training_fits_a_known_function(engine):
    rng = np.random.default_rng(7)
    X = rng.uniform(0, 1, (20000, mo.IN)).astype(np.float32)
    T = (3 + 4 * X[:, 0] - 2 * X[:, 3] + 3 * np.abs(X[:, 1] - 0.5)).astype(np.float32)
    saved = None
    try:
        saved = engine.cost_model_get()
    except mdr_hip.MdrError:
        pass
    engine.cost_model_set(None)
    base = float(np.mean((T - T.mean()) ** 2))
    loss = engine.mlp_train(X, T, epochs=40, batch=512, lr=3e-3, seed=3)
    assert loss < 0.05 * base
    w = engine.cost_model_get()
    pred = mo.forward(w, X)
    assert abs(float(np.mean((pred - T) ** 2)) - loss) < 1e-3 + 0.05 * loss
    engine.cost_model_set(saved)


@pytest.mark.gpu
def test_learned_partition_is_bit_exact(engine, known):
    saved = None
    try:
        saved = engine.cost_model_get()
    except mdr_hip.MdrError:
        pass
    X = 10**15
    for w in [rand_weights(30), np.zeros(mo.NPARAM, np.float32), saved]:
        engine.cost_model_set(w)
        for sched in (mdr_hip.SCHED_STATIC, mdr_hip.SCHED_DYNAMIC):
            r = engine.mertens(X, partition=mdr_hip.PARTITION_LEARNED, sched=sched, params=mdr_hip.PARAMS_TUNED)
            assert r == known[X]
            assert engine.last_stats.partition_used == (mdr_hip.PARTITION_LEARNED if w is not None else 0)
    engine.cost_model_set(saved)


def test_invalid_model_rejected_without_device():
    lib = mdr_hip.load()
    assert lib.mdr_cost_model_set(None, None, 0) == 1


@pytest.mark.gpu
def test_stats_report_cycles_and_algorithmic_bytes(engine, known):
    # the learned partition's predicted total against the measured
    # wave-cycles of the quotient phase, and the algorithmic bytes
    X = 10**16
    assert engine.mertens(X, partition=mdr_hip.PARTITION_LEARNED, params=mdr_hip.PARAMS_TUNED,
                          no_table_reuse=1) == known[X]
    st = engine.last_stats
    dense = st.s1_terms - st.s1_comp_terms
    assert st.quotient_bytes == 2 * dense + 20 * st.s1_comp_terms + 3 * st.s2_terms
    assert st.measured_cycles > 0
    if st.partition_used == mdr_hip.PARTITION_LEARNED:
        print("predicted / measured", st.predicted_cycles / st.measured_cycles)
        assert 0.25 < st.predicted_cycles / st.measured_cycles < 4.0
    assert engine.mertens(X, params=mdr_hip.PARAMS_TUNED, no_table_reuse=1) == known[X]
    assert engine.last_stats.predicted_cycles == 0


@pytest.mark.gpu
def test_learned_partition_against_the_static_partition(engine, known):
    # What the trained network gains over the engine's own static partition
    # (equal quotient positions per chunk) under the same schedule, at
    # X = 1e17.  With one chunk per resident wave (static round robin) the
    # chunks' costs decide the finish time, and equal predicted cycles must
    # balance the waves better than equal positions: a higher wave busy
    # fraction and a shorter quotient phase.  Under the dynamic chunk queue the
    # queue itself absorbs imbalance; the learned cuts must not lose there.
    X = 10**17
    rr = engine.waves()

    def run(partition, sched, chunk=0):
        o = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, partition=partition, sched=sched, chunk=chunk,
                              no_table_reuse=1, wave_span=1)
        engine.mertens(X, o)
        best = None
        for _ in range(2):
            assert engine.mertens(X, o) == known[X]
            st = engine.last_stats
            if best is None or st.quotient_ms < best[0]:
                best = (st.quotient_ms, st.wave_busy_frac, st.partition_used)
        return best

    q_s, b_s, _ = run(mdr_hip.PARTITION_STATIC, mdr_hip.SCHED_STATIC, rr)
    q_l, b_l, used = run(mdr_hip.PARTITION_LEARNED, mdr_hip.SCHED_STATIC, rr)
    assert used == mdr_hip.PARTITION_LEARNED
    print(f"round robin: static {q_s:.1f} ms (busy {b_s:.3f}), learned {q_l:.1f} ms (busy {b_l:.3f}), "
          f"gain {q_s / q_l:.3f}", flush=True)
    assert b_l > b_s
    assert q_s > 1.05 * q_l
    dq_s, db_s, _ = run(mdr_hip.PARTITION_STATIC, mdr_hip.SCHED_DYNAMIC)
    dq_l, db_l, _ = run(mdr_hip.PARTITION_LEARNED, mdr_hip.SCHED_DYNAMIC)
    print(f"queue: static {dq_s:.1f} ms (busy {db_s:.3f}), learned {dq_l:.1f} ms (busy {db_l:.3f})", flush=True)
    assert dq_l < 1.03 * dq_s
