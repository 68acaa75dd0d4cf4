"""The Dirichlet cross-check engine on the GPU.

dirichlet::DirichletEngine::compute_mertens is the engine the reference's
tests and benchmark use when dirichlet_engine.hpp is present (test_dr.cpp:8-11,
:71-88; bench_dr.cpp:9-12, :98).  Its algorithm (sum_{d <= v} M(v/d) = 1 over
the quotients of X, with a table to about X^(2/3)) shares nothing with the
Deleglise-Rivat split, so agreement with the known values, the compiled
reference's values and the DR engine is an independent check of both.
"""
import os
import random
import subprocess

import pytest

import mdr_hip
import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dirichlet_known_values(engine, known):
    for X, v in sorted(known.items()):
        if X > 10**15:
            continue
        print("dirichlet", X, flush=True)
        assert engine.dirichlet(X) == v, X


def test_dirichlet_cross_checks_of_test_dr(engine, ref_values):
    # the reference's cross-validation list (test_dr.cpp:73-75)
    for x in [54321, 987654, 12345678, 987654321, 5000000000, 123456789012, 1000000000000]:
        assert engine.dirichlet(x) == engine.mertens(x), x


def test_dirichlet_small_and_random(engine, ref_values):
    for X in [-3, 0, 1, 2, 3, 4, 5, 99999, 100000, 100001]:
        assert engine.dirichlet(X) == oracle.mertens(X, 1), X
    rng = random.Random(31)
    for X in rng.sample(sorted(ref_values), 25):
        print("dirichlet ref", X, flush=True)
        assert engine.dirichlet(X) == ref_values[X], X
    for _ in range(10):
        X = rng.randrange(10**6, 3 * 10**8)
        assert engine.dirichlet(X) == oracle.mertens(X, 8), X


def test_dirichlet_matches_dr_at_table_edges(engine):
    # X just above and below the table limit L (K = X // (L + 1) in {0, 1, 2})
    for X in [10**9 - 1, 10**9, 10**9 + 1, 8 * 10**9, 27 * 10**9 + 5]:
        assert engine.dirichlet(X) == engine.mertens(X, params=mdr_hip.PARAMS_TUNED), X


def test_dirichlet_at_1e17_and_its_range_limit(engine, known):
    assert engine.dirichlet(10**17) == known[10**17]
    with pytest.raises(mdr_hip.MdrError, match="RANGE"):
        engine.dirichlet(10**18 + 1)


def test_dirichlet_cpp_cross_validation_program():
    exe = os.path.join(ROOT, "tests", "cpp", "test_dropin_dirichlet")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
