"""Shared test set-up: markers, paths, the oracle and (for -m gpu) the engine."""
import json
import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path.insert(0, os.path.join(ROOT, "deleglise-rivat_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long GPU case (minutes)")


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def known():
    return {int(k): v for k, v in load_json("known.json").items()}


@pytest.fixture(scope="session")
def ref_values():
    return {int(k): v for k, v in load_json("ref_values.json").items()}


@pytest.fixture(scope="session")
def xlists():
    return load_json("xlists.json")


@pytest.fixture(scope="session")
def engine():
    import mdr_hip
    e = mdr_hip.Engine(0)
    yield e
    e.close()
