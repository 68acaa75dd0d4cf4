"""Large-prime stage of the sieve (deleglise-rivat_amd/csrc/mdr_table.hip,
stage 4), host only.

The kernel finds each large prime's offsets in a segment with 32-bit
multiplies on per-prime records (large_prime_consts, mdr_common.hpp) instead
of a 64-bit division: a Shoup multiplication for the first multiple of p^2 or
p, a Barrett reduction mod p, and an inverse-multiplication test of p^2 | n;
and it finalises four accumulator bytes at once (mu_swar, mu_nibbles) and
packs each 64-value word's kept bits (the values not divisible by 4) into
the compressed lines (kept_index, kept48, expand48).  The probe compiles that
same constexpr code for the host and checks the first against plain
64/128-bit remainders for every prime from 256 to the largest sieving prime
of a table below SIEVE_MAX_U, the second against the per-byte rule mu_byte
for every byte value and threshold, the third against per-bit rules on
seeded mu words; the table tests on the GPU check the sieve as a whole
against Eratosthenes.
"""
import os
import subprocess

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
CSRC = os.path.join(ROOT, "deleglise-rivat_amd", "csrc")


def test_large_prime_arithmetic(tmp_path):
    exe = str(tmp_path / "sieve_arith_probe")
    subprocess.run(["g++", "-O2", "-std=c++20", "-I" + CSRC, os.path.join(ROOT, "tests", "cpp", "sieve_arith_probe.cpp"),
                    "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    assert out.strip().endswith("checks ok"), out
    primes = int(out.split()[0])
    assert primes == 681528   # primes in [256, isqrt(24576 * 2^32 - 1)]
