"""Tuning-cache reader (deleglise-rivat_amd/csrc/mdr_tune.cpp), host only.

The autotuner writes one line per decade of X: `decade fx cx chunks u_dense
sieve_shape`.  Every sieve shape the kernels instantiate must survive the
reader (shapes 4 and 5, the 49152/512 and 24576/512 workgroups, were once
dropped to the default); out-of-range shapes fall back to shape 0 and
malformed lines are skipped.
"""
import os
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
CSRC = os.path.join(ROOT, "deleglise-rivat_amd", "csrc")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("tune_probe") / "tune_probe")
    subprocess.run(["g++", "-O1", "-std=c++20", "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "tune_probe.cpp"), os.path.join(CSRC, "mdr_tune.cpp"),
                    "-ldl", "-o", exe], check=True)
    return exe


def run(probe, path, xs):
    env = dict(os.environ, MDR_TUNE_FILE=str(path))
    out = subprocess.run([probe] + [str(x) for x in xs], capture_output=True, text=True, env=env,
                         check=True).stdout.split("\n")
    rows = {}
    for line in out:
        if line.strip():
            f = line.split()
            rows[int(f[0])] = dict(shape=int(f[1]), chunks=int(f[2]), u=int(f[3]), ud=int(f[4]), cx=float(f[5]))
    return rows


def test_every_sieve_shape_is_kept(probe, tmp_path):
    # decades 8..13 carry shapes 0..5; 14 an out-of-range shape; 15 none
    lines = ["# decade fx cx chunks u_dense sieve_shape"]
    for d, sh in zip(range(8, 14), range(6)):
        lines.append(f"{d} 1.0 2.0 {1000 * d} 130944 {sh}")
    lines += ["14 1.0 2.0 0 130944 6", "15 1.0 2.0 0 130944", "16 not a line", "batch 2.0"]
    f = tmp_path / "tune.txt"
    f.write_text("\n".join(lines) + "\n")
    xs = [10**d for d in range(8, 17)]
    rows = run(probe, f, xs)
    for d, sh in zip(range(8, 14), range(6)):
        assert rows[10**d]["shape"] == sh, d
        assert rows[10**d]["chunks"] == 1000 * d
    assert rows[10**14]["shape"] == 0
    assert rows[10**15]["shape"] == 0
    assert rows[10**16]["shape"] == 0 and rows[10**16]["chunks"] == 0


def test_shipped_cache_shapes(probe):
    # the cache shipped beside the library: decades 9 and 11-14 choose
    # shape 5 (24576/512), which the reader must return as written
    path = os.path.join(ROOT, "deleglise-rivat_amd", "tune_cache.txt")
    want = {}
    for line in open(path):
        f = line.split()
        if f and f[0].isdigit() and len(f) >= 6:
            want[int(f[0])] = int(f[5])
    rows = run(probe, path, [10**d for d in sorted(want)])
    for d, sh in want.items():
        assert rows[10**d]["shape"] == sh, d
    assert any(sh in (4, 5) for sh in want.values())
