"""End-to-end M(X) on the GPU through the C ABI.

Bit-exact against: the reference tests' known values (test_dr.cpp:26-47,
test_cuda.cu:38-60, README.md:188), the compiled reference's outputs
(tests/golden/ref_values.json), the oracle at random X, and -- at full sizes --
the size-independent invariants of the method: M(X) does not depend on the
cutoff u, the split cx, the strategy, the chunking, the schedule, the
partition or on being evaluated alone or in a batch.
"""
import os
import random
import subprocess

import pytest

import mdr_hip
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def test_known_values_all_modes(engine, known):
    modes = [dict(), dict(strategy=mdr_hip.STRATEGY_DENSE), dict(strategy=mdr_hip.STRATEGY_STREAMING),
             dict(params=mdr_hip.PARAMS_TUNED), dict(sched=mdr_hip.SCHED_STATIC),
             dict(partition=mdr_hip.PARTITION_LEARNED)]
    for X, v in sorted(known.items()):
        if X > 10**16:
            continue
        for kw in modes:
            print("known", X, kw, flush=True)
            assert engine.mertens(X, **kw) == v, (X, kw)


def test_small_and_negative_inputs(engine):
    for X in [-5, 0, 1, 2, 3, 4, 5]:
        assert engine.mertens(X) == oracle.mertens(X, 1)
    rng = random.Random(1)
    xs = [rng.randrange(4, 5 * 10**7) for _ in range(40)] + [100000, 100001, 49999999, 50000000, 50000001]
    for X in xs:
        assert engine.mertens(X) == oracle.mertens(X, 8), X


def test_reference_values(engine, ref_values):
    # every value the compiled reference produced (up to ~1.3e16)
    for X in sorted(ref_values):
        assert engine.mertens(X) == ref_values[X], X


def test_reference_values_as_one_batch(engine, ref_values):
    xs = sorted(ref_values)
    assert engine.batch(xs) == [ref_values[x] for x in xs]


def test_oracle_random_points(engine):
    rng = random.Random(2)
    for _ in range(30):
        X = rng.randrange(5 * 10**7, 10**12)
        assert engine.mertens(X) == oracle.mertens(X, 8), X


def test_compute_compressed_boundaries(engine):
    # test_sieve.cpp:205-214 through the forced streaming strategy
    f = mdr_hip.STRATEGY_FORCE_STREAMING
    assert engine.mertens(99975680, strategy=f) == oracle.mertens_recurrence(99975680, 1 << 22)
    assert engine.mertens(0, strategy=f) == 0
    assert engine.mertens(10**8, strategy=f) == 1928
    assert engine.mertens(10**10, strategy=f) == -33722


def test_invariance_full_size(engine, known):
    # M(1e16) under every control: u caps, cx, strategy, chunk length, schedule
    X, v = 10**16, known[10**16]
    for kw in [dict(cx=0.6), dict(cx=2.4), dict(u_cap=2 * 10**9), dict(u_cap=10**10),
               dict(strategy=mdr_hip.STRATEGY_DENSE, u_cap_dense=4 * 10**9), dict(chunk=5000),
               dict(chunk=1 << 20), dict(sched=mdr_hip.SCHED_STATIC, chunk=20000), dict(params=mdr_hip.PARAMS_TUNED),
               dict(partition=mdr_hip.PARTITION_TERMS), dict(partition=mdr_hip.PARTITION_TERMS, chunk=777),
               dict(partition=mdr_hip.PARTITION_TERMS, sched=mdr_hip.SCHED_STATIC, params=mdr_hip.PARAMS_TUNED)]:
        print("invariance", kw, flush=True)
        assert engine.mertens(X, **kw) == v, kw


def test_mertens_1e17(engine, known):
    assert engine.mertens(10**17) == known[10**17]
    assert engine.mertens(10**17, params=mdr_hip.PARAMS_TUNED) == known[10**17]


def test_batch_config_sample(engine, xlists, ref_values):
    # batch workload: all values below 1e11 plus the 100 largest, checked
    xs = xlists["batch_4096_1e9_1e14_seed3"]
    sample = sorted(set([x for x in xs if x < 10**11] + sorted(xs)[-100:]))
    got = engine.batch(sample)
    assert got == [ref_values[x] for x in sample]


def test_batch_equals_single(engine, xlists):
    xs = xlists["batch_4096_1e9_1e14_seed3"][:64]
    single = [engine.mertens(x) for x in xs]
    assert engine.batch(xs) == single
    assert engine.batch(xs, partition=mdr_hip.PARTITION_TERMS) == single
    assert engine.last_stats.partition_used == mdr_hip.PARTITION_TERMS


def test_large_batch_smallest_values(engine, xlists, ref_values):
    xs = sorted(xlists["large_64_1e16_1e18_seed5"])[:4]
    assert engine.batch(xs) == [ref_values[x] for x in xs]


def test_dropin_cpp_program():
    exe = os.path.join(ROOT, "tests", "cpp", "test_dropin")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe, str(10**16)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_dropin_windowed_cpp_program():
    exe = os.path.join(ROOT, "tests", "cpp", "test_dropin_windowed")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe, str(10**16)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "[win]" in r.stderr


def test_largest_table_and_external_1e18():
    """M(1e18) = -46758740 is not a reference value (OEIS A084237): a
    cross-check at the largest X of the batch workloads, with the tuned cutoff,
    with a 3.9e11-value table (U - u_dense above 2^38) and with a table of
    2^31 + 1 compressed lines (line indices past 2^31, read as unsigned 32-bit
    integers)."""
    e = mdr_hip.Engine(0)
    try:
        X = 10**18
        assert e.mertens(X, mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, no_table_reuse=1)) == -46758740
        ud = 1048576
        U = 390_000_000_000 // 256 * 256
        assert U - ud > 2**38
        o = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, tune_u=U, tune_ud=ud, no_table_reuse=1)
        assert e.mertens(X, o) == -46758740
        assert e.last_stats.U == U
        U = ud + 256 * (2**31 + 1)
        o = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, tune_u=U, tune_ud=ud, no_table_reuse=1)
        assert e.mertens(X, o) == -46758740
        assert e.last_stats.U == U
    finally:
        e.close()


def test_cli_modes(known):
    exe = os.path.join(ROOT, "deleglise-rivat_amd", "mdr_bench")
    r = subprocess.run([exe, str(10**12)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"M({10**12}) = {known[10**12]}" in r.stdout and "Timing Breakdown" in r.stdout
    r = subprocess.run([exe, "--benchmark", "13", "0", "3"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = [l for l in r.stdout.splitlines() if l.startswith("| 10^")]
    assert len(rows) == 13
    for e, row in zip(range(1, 14), rows):
        cells = [c.strip() for c in row.strip("|").split("|")]
        assert int(cells[1]) == 10**e and int(cells[2]) == known[10**e], row
    for mode, top in (("--compare", 15), ("--compare-split", 14)):
        r = subprocess.run([exe, mode, str(top)], capture_output=True, text=True, timeout=300)
        print(r.stdout)
        assert r.returncode == 0, r.stderr
        rows = [l for l in r.stdout.splitlines() if l.startswith("| 10^")]
        assert [int(l.strip("|").split("|")[1]) for l in rows] == [known[10**e] for e in range(10, top + 1)]


def test_largest_int64_x_two_cutoffs_agree():
    """X = 2^63 - 1, the top of the int64 range (no reference value exists):
    the tuned plan (cutoff 1e12 above the one-table cap, so a windowed table)
    and a 3e11 one-table cutoff must give the same M(X); the 64-bit S2 path
    (j >= 2^31) and the largest quotients are exercised."""
    e = mdr_hip.Engine(0)
    try:
        X = 2**63 - 1
        a = e.mertens(X, params=mdr_hip.PARAMS_TUNED, no_table_reuse=1)
        assert e.last_stats.windows >= 2 and abs(e.last_stats.u - 10**12) < 256
        b = e.mertens(X, params=mdr_hip.PARAMS_TUNED, tune_u=3 * 10**11, tune_ud=524288, no_table_reuse=1)
        assert e.last_stats.windows == 0
        assert a == b == 580197744
    finally:
        e.close()
