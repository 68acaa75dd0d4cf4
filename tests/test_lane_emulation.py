"""Lane-by-lane CPU emulation of the compressed S1 lookups (CPU only).

The quotient kernel sums M(floor(y/n)) over a run of cnt positions n = nlo,
nlo + ST, ... with the 64 lanes of a wave taking positions lane, lane + 64,
... (csrc/mdr_quot.hip, s1_comp).  Each lane runs its own trip count
ceil((cnt - lane) / 64), so no lane touches a position past the run.

The compressed tier is emulated in its stored form (csrc/mdr_common.hpp):
64-byte lines of 256 values starting at n = 1 (mod 4), whose 192 values not
divisible by 4 are kept as three 64-bit words of nonzero and negative bits
with a checkpoint each.  A lookup at tier offset off reads line off >> 8,
kept index k = kept_index(off mod 256) = r - (r + 1) // 4, word k >> 6 and
bit k & 63; the emulation builds the lines from mu and checks every lookup
against the plain prefix sum.

A switched-off variant gave every lane the wave's trip count ceil(cnt / 64)
and took the line from a 32-bit index ((q - 1) mod 2^32) >> 8 counted from an
origin ud / 256 lines before the line array.  The emulation walks whole runs
of both kinds with a bounds check on every line index and shows the cause of
that variant's wrong M(X) and of its fault: whenever cnt is not a multiple of
64 the last step's spare lanes evaluate positions past the run; where the run
ends at the dense/compressed boundary those positions have q <= ud, whose line
index lies before the line array (and q = 0, past n > y, wraps the 32-bit
index to a line far beyond it), and everywhere else they add terms that belong
to other runs, so the sum is wrong even when every read stays inside.
"""
import numpy as np

import oracle

LINE = 256      # values per line
KEPT = 192      # values per line not divisible by 4 (three 64-bit words)


def kept_index(r):
    return r - (r + 1) // 4


def table(U, ud):
    """M(0..U) and the compressed lines above ud: per line, three words of
    nonzero bits, three of negative bits and three checkpoints, built from mu
    in the stored layout (ud a multiple of 256)."""
    assert ud % LINE == 0
    mu = oracle.mobius(U).astype(np.int64)
    M = np.cumsum(mu)
    lines = -(-(U - ud) // LINE)
    nz = np.zeros((lines, 3), dtype=np.uint64)
    ng = np.zeros((lines, 3), dtype=np.uint64)
    ck = np.zeros((lines, 3), dtype=np.int64)
    for line in range(lines):
        n0 = ud + 1 + LINE * line
        k = 0
        for r in range(LINE):
            n = n0 + r
            if r % 4 == 3:
                assert n > U or mu[n] == 0          # multiples of 4
                continue
            if k % 64 == 0:
                ck[line, k // 64] = M[min(n, U + 1) - 1]
            if n <= U and mu[n] != 0:
                nz[line, k // 64] |= np.uint64(1 << (k % 64))
                if mu[n] < 0:
                    ng[line, k // 64] |= np.uint64(1 << (k % 64))
            k += 1
        assert k == KEPT
    return M, (nz, ng, ck), lines


def line_M(tab, line, sub, bit):
    """The kernel's line_M: checkpoint of kept word sub plus the plane bits
    0..bit of that word."""
    nz, ng, ck = tab
    mask = (1 << (bit + 1)) - 1
    z, g = int(nz[line, sub]) & mask, int(ng[line, sub]) & mask
    return int(ck[line, sub]) + bin(z).count("1") - 2 * bin(g).count("1")


def lookup(tab, ud, q):
    off = q - ud - 1
    k = kept_index(off % LINE)
    return line_M(tab, off >> 8, k >> 6, k & 63)


def run_shipped(M, tab, U, ud, lines, y, nlo, cnt, st):
    total, reads = 0, []
    for lane in range(64):
        iters = (cnt - lane + 63) // 64 if lane < cnt else 0
        for k in range(iters):
            n = nlo + st * (lane + 64 * k)
            q = y // n
            off = q - ud - 1
            line = off >> 8
            reads.append(line)
            assert 0 <= line < lines and 0 <= off < U - ud, (y, n, q)
            total += lookup(tab, ud, q)
    return total, reads


def run_uniform_trip(M, tab, U, ud, lines, y, nlo, cnt, st):
    """The switched-off variant: every lane runs ceil(cnt / 64) steps; the line
    is ((q - 1) mod 2^32) >> 8 from an origin ud / 256 lines before the line
    array.  Returns the sum over in-bounds reads and the line indices
    (relative to the array) of every read."""
    total, reads = 0, []
    steps = -(-cnt // 64)
    for lane in range(64):
        for k in range(steps):
            n = nlo + st * (lane + 64 * k)
            q = y // n
            line = (((q - 1) & 0xffffffff) >> 8) - ud // LINE
            reads.append(line)
            if 0 <= line < lines and q > ud:
                total += lookup(tab, ud, q)
    return total, reads


def direct(M, y, nlo, cnt, st):
    return int(sum(M[y // (nlo + st * i)] for i in range(cnt)))


def test_every_compressed_value_reads_its_prefix_sum():
    U, ud = 100_000, LINE * 30
    M, tab, lines = table(U, ud)
    for q in range(ud + 1, U + 1):
        assert lookup(tab, ud, q) == M[q], q


def test_shipped_lookup_walks_whole_runs_in_bounds():
    U, ud = 400_000, LINE * 30
    M, tab, lines = table(U, ud)
    rng = np.random.default_rng(3)
    for _ in range(120):
        y = int(rng.integers(U * 5, U * 400))
        st = int(rng.choice([1, 2]))
        nlo = y // U + 1
        split = y // (ud + 1)                    # n <= split: q > ud (compressed tier)
        cnt = (split - nlo) // st + 1
        cnt = int(rng.integers(1, cnt + 1)) if cnt > 1 else 1
        got, reads = run_shipped(M, tab, U, ud, lines, y, nlo, cnt, st)
        assert got == direct(M, y, nlo, cnt, st)
        assert min(reads) >= 0 and max(reads) < lines


def test_uniform_trip_count_reads_outside_the_lines_at_the_tier_boundary():
    # a run that ends exactly at the dense/compressed boundary, with a length
    # that is not a multiple of 64: the spare lanes of the last step take n
    # past the boundary, where q <= ud, and their line index is negative
    U, ud = 400_000, LINE * 30
    M, tab, lines = table(U, ud)
    bad_runs = 0
    for y in (9_000_000, 12_345_679, 30_000_001, 77_777_777):
        for st in (1, 2):
            nlo = y // U + 1
            split = y // (ud + 1)
            cnt = (split - nlo) // st + 1
            if cnt % 64 == 0:
                cnt -= 1
            nlo = split - st * (cnt - 1)         # the run ends at the boundary
            shipped, r1 = run_shipped(M, tab, U, ud, lines, y, nlo, cnt, st)
            assert min(r1) >= 0 and max(r1) < lines
            _, r2 = run_uniform_trip(M, tab, U, ud, lines, y, nlo, cnt, st)
            assert min(r2) < 0, (y, st)            # reads before the line array
            bad_runs += 1
    assert bad_runs == 8


def test_uniform_trip_count_wraps_past_y_and_double_counts_inside_runs():
    U, ud = 400_000, LINE * 30
    M, tab, lines = table(U, ud)
    # positions past n > y (q = 0): the 32-bit index of q - 1 wraps to
    # 2^24 - 1, a line far beyond the array
    y, nlo, cnt, st = 1_000_003, 999_990, 10, 1
    _, r = run_uniform_trip(M, tab, U, ud, lines, y, nlo, cnt, st)
    assert max(r) >= (0xffffffff >> 8) - ud // LINE > lines
    # a run cut inside the compressed tier (a chunk boundary): every read is
    # in bounds, but the spare lanes add the next chunk's terms
    y, st = 50_000_000, 1
    nlo = y // U + 1
    cnt = 1000                                   # not a multiple of 64
    shipped, _ = run_shipped(M, tab, U, ud, lines, y, nlo, cnt, st)
    uniform, r = run_uniform_trip(M, tab, U, ud, lines, y, nlo, cnt, st)
    assert min(r) >= 0 and max(r) < lines
    assert shipped == direct(M, y, nlo, cnt, st)
    extra = direct(M, y, nlo, 64 * -(-cnt // 64), st) - shipped
    assert uniform - shipped == extra != 0
