"""Writes the golden fixtures under tests/golden/.

  xlists.json      seeded log-uniform X lists of the batch / learned / large-batch
                   workloads (mt19937_64, restated below and checked against the
                   C++ standard's required 10000th output).
  known.json       M(X) values the reference's own tests and README hold:
                   test_dr.cpp:26-47, test_cuda.cu:38-60, README.md:188.
  ref_values.json  M(X) computed here by the compiled reference
                   (oracle/_ref/bench_dr, built from /root/reference by
                   oracle/Makefile) for the sampled workload values, the
                   test_cuda.cu:85-99 cross-check values and extra X.

Run from the repository root:  python tests/golden/make_golden.py [--skip-ref]
"""
import json
import math
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
BENCH_DR = os.path.join(ROOT, "oracle", "_ref", "bench_dr")

MASK = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne Twister (Matsumoto & Nishimura 2000), as std::mt19937_64."""

    def __init__(self, seed):
        self.mt = [0] * 312
        self.mt[0] = seed & MASK
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & MASK
        self.idx = 312

    def _twist(self):
        mt = self.mt
        for i in range(312):
            x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self.idx = 0

    def next(self):
        if self.idx >= 312:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & MASK


def log_uniform(seed, n, lo, hi):
    g = MT19937_64(seed)
    out = []
    a, b = math.log(lo), math.log(hi)
    for _ in range(n):
        u = (g.next() >> 11) * 2.0 ** -53
        x = int(math.floor(math.exp(a + u * (b - a))))
        out.append(min(max(x, lo), hi))
    return out


def check_generator():
    g = MT19937_64(5489)
    v = 0
    for _ in range(10000):
        v = g.next()
    assert v == 9981545732273789042, v   # [rand.predef] in the C++ standard


KNOWN = {  # test_dr.cpp:26-47
    1: 1, 2: 0, 3: -1, 4: -1, 5: -2, 10: -1, 100: 1, 1000: 2, 10**4: -23, 10**5: -48,
    10**6: 212, 10**7: 1037, 10**8: 1928, 10**9: -222, 10**10: -33722, 10**11: -87856,
    10**12: 62366, 10**13: 599582, 10**14: -875575, 10**15: -3216373,
    10**16: -3195437,        # test_cuda.cu:59
    10**17: -21830254,       # README.md:188
}
CROSS = [54321, 987654, 12345678, 987654321, 5000000000, 123456789012, 1000000000000]  # test_cuda.cu:85-99
BOUNDARY = [100000, 100001, 262143, 262144, 262145, 999983, 1000000, 49999999, 50000000,
            50000001, 50000003, 987654321, 10000000000, 99975680]   # test_sieve.cpp:183-208


def ref_mertens(X, threads=8):
    out = subprocess.run([BENCH_DR, str(X), str(threads)], capture_output=True, text=True, check=True).stdout
    for line in out.splitlines():
        if line.startswith("M(") and "=" in line:
            return int(line.split("=")[1])
    raise RuntimeError(out)


def main():
    check_generator()
    xl = {
        "batch_4096_1e9_1e14_seed3": log_uniform(3, 4096, 10**9, 10**14),
        "train_512_1e14_1e16_seed4": log_uniform(4, 512, 10**14, 10**16),
        "large_64_1e16_1e18_seed5": log_uniform(5, 64, 10**16, 10**18),
    }
    with open(os.path.join(HERE, "xlists.json"), "w") as f:
        json.dump(xl, f)
    with open(os.path.join(HERE, "known.json"), "w") as f:
        json.dump({str(k): v for k, v in KNOWN.items()}, f, indent=0)
    if "--skip-ref" in sys.argv:
        return
    path = os.path.join(HERE, "ref_values.json")
    vals = {}
    if os.path.exists(path):
        vals = {int(k): v for k, v in json.load(open(path)).items()}
    b = xl["batch_4096_1e9_1e14_seed3"]
    want = sorted(set([x for x in b if x < 10**11] + sorted(b)[-100:]))
    want += CROSS + BOUNDARY
    want += sorted(xl["large_64_1e16_1e18_seed5"])[:4]
    want += log_uniform(11, 24, 10**11, 10**15)     # extra full-size parity points
    for i, X in enumerate(sorted(set(want))):
        if X in vals:
            continue
        vals[X] = ref_mertens(X)
        if i % 100 == 0:
            print(i, X, vals[X], flush=True)
            with open(path, "w") as f:
                json.dump({str(k): v for k, v in sorted(vals.items())}, f)
    with open(path, "w") as f:
        json.dump({str(k): v for k, v in sorted(vals.items())}, f)


if __name__ == "__main__":
    main()
