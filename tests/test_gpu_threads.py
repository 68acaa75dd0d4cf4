"""Host threads on the C ABI: one context per thread runs concurrently, and
calls on one shared context are serialised by the context (the boundary's
threading rule).  Results are checked against the compiled reference's values.
"""
import random
import threading

import pytest

import mdr_hip

pytestmark = pytest.mark.gpu


def pick(ref_values, n, seed, lo=10**11, hi=2 * 10**15):
    xs = sorted(x for x in ref_values if lo <= x <= hi)
    return random.Random(seed).sample(xs, n)


def run_threads(target, n):
    errs = []

    def wrap(i):
        try:
            target(i)
        except BaseException as e:  # noqa: BLE001 - reported below
            errs.append((i, repr(e)))

    ts = [threading.Thread(target=wrap, args=(i,)) for i in range(n)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in ts), "a thread did not finish"
    assert not errs, errs


def test_one_context_per_thread(ref_values):
    xs = pick(ref_values, 24, 11)
    got = {}

    def work(i):
        e = mdr_hip.Engine(0)
        try:
            for k, x in enumerate(xs[i::4]):
                kw = dict(params=mdr_hip.PARAMS_TUNED) if k % 2 else {}
                got[x] = e.mertens(x, **kw)
            got[("batch", i)] = e.batch(xs[i::4], params=mdr_hip.PARAMS_TUNED, use_graph=i % 2)
        finally:
            e.close()

    run_threads(work, 4)
    for x in xs:
        assert got[x] == ref_values[x], x
    for i in range(4):
        assert got[("batch", i)] == [ref_values[x] for x in xs[i::4]]


def test_shared_context_serialises_calls(engine, ref_values):
    xs = pick(ref_values, 16, 12)
    got = {}

    def work(i):
        for x in xs[i::2]:
            got[x] = engine.mertens(x, params=mdr_hip.PARAMS_TUNED, no_table_reuse=i)

    run_threads(work, 2)
    assert all(got[x] == ref_values[x] for x in xs)
