#!/usr/bin/env python3
"""Benchmark of the MI355X Deleglise-Rivat engine.

Metric: M(X) wall time at X = 1e17, learned versus static S2 balance;
batched M(X) evaluations per second.

One step = one exact M(1e17) through the C ABI, table build included (the
table is rebuilt every step; X is the only input and lives on the host as an
integer argument, so nothing needs staging).  `value` is evaluations per
second over the K timed steps.  Beside it:
  - the partitions of the same X: equal quotient positions per chunk, equal
    predicted cycles (learned) and whole terms per chunk (the reference GPU
    engine's unit), under the dynamic chunk queue and under a one-chunk-per-
    wave round robin, with each launch's wave busy fraction (the idle tail
    bounds what any partition can gain),
  - the other configurations: M(1e12) dense and M(1e15) streaming with the
    reference's parameters (warm median of 3 and cold), the 4096-X batch in
    [1e9, 1e14] against one-at-a-time calls, the 64-X batch in [1e16, 1e18],
    each batch under the learned and the static partition,
  - the quotient kernel's roofline against its binding resource: VALU issue
    (a lower bound on instructions from the inner loops' instruction counts)
    and DRAM bytes (counter profile committed under profiles/),
  - the reference CPU engine, compiled from the reference sources, timed on
    this host's cores at the same configurations (cpu_baseline).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import re
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "deleglise-rivat_amd"))

import mdr_hip  # noqa: E402

METRIC = "M(X) wall time at X=1e17 learned vs static S2 balance; batched M(X) evals/sec"
HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every
# 4 cycles (16 lanes per cycle; the unit SQ_ACTIVE_INST_VALU counts in), at
# the 2.4 GHz peak engine clock.
SIMDS, CLOCK_HZ, CYCLES_PER_INST = 1024, 2.4e9, 4
ISSUE_PEAK_GINST = SIMDS * CLOCK_HZ / CYCLES_PER_INST / 1e9
# VALU instructions per quotient position in the quotient kernel's inner loops
# (gfx950 build of csrc/mdr_quot.hip, counted per loop iteration and divided by
# the positions an iteration takes): dense S1 23 per 2 positions, compressed
# S1 66 per 2, S2 48..80 per 4 depending on the piece (the cheapest is used,
# so the issue time below is a lower bound).
VALU_PER_S1_DENSE, VALU_PER_S1_COMP, VALU_PER_S2_MIN = 11.5, 33.0, 12.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "traffic_mertens_1e17.json")
# Random 64-byte-line read rate of the device by buffer size (committed
# summary of deleglise-rivat_amd/mdr_lineprobe): the ceiling of the quotient
# kernel's scattered table lookups, each L2 miss a 128-byte line.
LINE_PROFILE = os.path.join(ROOT, "profiles", "random_lines_mi355x.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_json(name):
    with open(os.path.join(ROOT, "tests", "golden", name)) as f:
        return json.load(f)


# ---- CPU baseline ---------------------------------------------------------------
def host_info():
    info = {"nproc": os.cpu_count()}
    try:
        info["nproc_affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for key, name in (("Model name", "model"), ("Socket(s)", "sockets"), ("Core(s) per socket", "cores_per_socket"),
                          ("Thread(s) per core", "threads_per_core")):
            m = re.search(rf"^{re.escape(key)}:\s*(.+)$", out, re.M)
            if m:
                info[name] = m.group(1).strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_threads():
    # the host share this process may use: OMP_NUM_THREADS when the host sets
    # it (a one-GPU host share is 16 cores), else every CPU of the affinity set
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def ref_single(exe, X, threads, env=None):
    out = subprocess.run([exe, str(X), str(threads)], capture_output=True, text=True, timeout=900,
                         env=dict(os.environ, **(env or {}))).stdout
    return (int(re.search(r"M\(\d+\) = (-?\d+)", out).group(1)),
            float(re.search(r"Elapsed time: ([0-9.]+)", out).group(1)))


def ref_median(exe, exp, threads, repeats, env=None):
    # the reference's own method: decades 10^1 .. 10^exp in one process, median of `repeats`
    out = subprocess.run([exe, "--benchmark", str(exp), str(threads), str(repeats)], capture_output=True, text=True,
                         timeout=900, env=dict(os.environ, **(env or {}))).stdout
    row = [l for l in out.splitlines() if l.startswith(f"| 10^{exp} ")][0]
    cells = [c.strip() for c in row.strip("|").split("|")]
    return int(cells[2]), float(cells[3].split()[0])


def cpu_baseline(threads, expect):
    """The reference's OpenMP engine, compiled from its sources by oracle/Makefile,
    timed on this host: M(1e17) single run (the headline configuration), M(1e12)
    dense and M(1e15) streaming as medians of 3 (benchmarks/streaming-2026-10-19.md
    :15-27, bench_dr.cpp:48-73)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "bench_dr")
    if not os.path.exists(exe):
        return None
    base = {"unit": "M(X)/s", "cores": threads, "kind": "reference", "host": host_info()}
    ref_single(exe, 10**8, threads)   # warm-up
    m12, t12 = ref_median(exe, 12, threads, 3, {"MERTENS_USE_CHMT": "0"})
    m15, t15 = ref_median(exe, 15, threads, 3)
    m17, t17 = ref_single(exe, 10**17, threads)
    for X, m in ((10**12, m12), (10**15, m15), (10**17, m17)):
        if X in expect:
            assert m == expect[X], (X, m)
    base.update({
        "value": 1.0 / t17,
        "sample": f"the whole M(1e17) = {m17} on {threads} threads: {t17:.2f} s (single run, as the reference "
                  f"times 1e16 and above); M(1e12) dense and M(1e15) streaming as medians of 3",
        "seconds": {"1e17": round(t17, 3), "1e15_streaming_median3": round(t15, 4),
                    "1e12_dense_median3": round(t12, 5)},
    })
    return base


# ---- rooflines --------------------------------------------------------------------
def random_line_ceiling():
    """Best random-line rate (G lines/s) over the buffers beyond the L2s
    (64 MiB and up) from the committed probe summary, or None."""
    if not os.path.exists(LINE_PROFILE):
        return None
    best = None
    with open(LINE_PROFILE) as f:
        for line in f:
            m = re.match(r"\|\s*(\d+) MiB\s*\|\s*[\d.]+\s*\|\s*([\d.]+)\s*\|", line)
            if m and int(m.group(1)) >= 64:
                v = float(m.group(2))
                best = v if best is None else max(best, v)
    return best


def traffic_profile():
    if os.path.exists(TRAFFIC_PROFILE):
        with open(TRAFFIC_PROFILE) as f:
            return json.load(f)
    return None


def quotient_roofline(stats, X, prof):
    """k_quot over the timed steps, against VALU issue and DRAM bandwidth.
    issue: lower-bound wave instructions per launch (positions of each kind x
    the inner loop's instructions per position / 64 lanes) / mean launch time
    (HIP events around the launch on the engine's stream) against 1024 SIMDs
    issuing one wave64 instruction per 4 cycles.  dram: bytes per launch from
    the committed counter profile (TCC_EA0_RDREQ_DRAM_32B and
    TCC_EA0_WRREQ_WRITE_DRAM_32B, 32-byte units, calibrated on a device copy)
    / the same launch time against 8 TB/s.  The larger fraction binds."""
    st = stats[-1]
    dense = st.s1_terms - st.s1_comp_terms
    insts = (VALU_PER_S1_DENSE * dense + VALU_PER_S1_COMP * st.s1_comp_terms + VALU_PER_S2_MIN * st.s2_terms) / 64.0
    ms = sum(s.quotient_ms for s in stats) / len(stats)
    sec = ms / 1000.0
    issue_ach = insts / sec / 1e9
    issue = {"achieved": round(issue_ach, 2), "peak": ISSUE_PEAK_GINST, "unit": "Ginst/s",
             "frac": round(issue_ach / ISSUE_PEAK_GINST, 4), "insts_per_launch": insts,
             "per_position": {"s1_dense": VALU_PER_S1_DENSE, "s1_comp": VALU_PER_S1_COMP, "s2_min": VALU_PER_S2_MIN}}
    dram = None
    k = (prof or {}).get("kernels", {}).get("k_quot") if prof and prof.get("X") == X else None
    if k:
        byts = k["dram_read_bytes"] + k["dram_write_bytes"]
        ach = byts / sec / 1e9
        dram = {"achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "bytes": byts, "l2_hit_rate": k.get("l2_hit_rate")}
        ceil = random_line_ceiling()
        if ceil:
            lines = k["dram_read_bytes"] / 128.0 / sec / 1e9
            dram["random_lines"] = {"achieved": round(lines, 2), "peak": ceil, "unit": "G lines/s",
                                    "frac": round(lines / ceil, 4),
                                    "note": "128-byte L2 fills per second against the device's best random "
                                            "64-byte-line read rate (profiles/random_lines_mi355x.md)"}
    binding = "valu" if dram is None or issue["frac"] >= dram["frac"] else "hbm"
    top = issue if binding == "valu" else dram
    return {"bound": binding, "achieved": top["achieved"], "peak": top["peak"], "unit": top["unit"],
            "frac": top["frac"], "traffic": dram["bytes"] if dram else None, "kernel": "k_quot",
            "kernel_ms": round(ms, 3), "launches": len(stats), "issue": issue, "hbm": dram,
            "valu_busy_counter": (k or {}).get("valu_busy"), "valu_issue_counter": (k or {}).get("valu_issue"),
            "algorithmic_bytes_gather_model": st.quotient_bytes,
            "note": "issue = lower-bound VALU instructions from the inner loops / live kernel time; hbm = DRAM-side "
                    "bytes by counters (an upper bound on HBM bytes: the MALL may serve part) / live kernel time; "
                    "valu_busy_counter = SQ_ACTIVE_INST_VALU x 4 / (SIMDs x cycles) and valu_issue_counter = "
                    "SQ_INSTS_VALU x 4 / (SIMDs x cycles) from the same profile"}


def table_roofline(stats, X, prof):
    """Table build (k_sieve + scan + fix-up) over the timed steps: table bytes
    written / build time against HBM, and the sieve kernel's VALU issue from
    the committed counter profile (the binding resource): SQ_INSTS_VALU x 4
    cycles / (SIMDs x the kernel's cycles), the share of issue slots at one
    wave64 instruction per 4 cycles.  The profile's cycle count (GRBM_GUI_ACTIVE
    over the dispatch) carries about 1 % of error, so a saturated kernel can
    read slightly above 1: frac is that share capped at 1, the reading is kept
    beside it."""
    st = stats[-1]
    ms = sum(s.table_ms for s in stats) / len(stats)
    ach = st.table_bytes / (ms / 1000) / 1e9
    k = (prof or {}).get("kernels", {}).get("k_sieve") if prof and prof.get("X") == X else None
    out = {"bound": "valu" if k else "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(ach / HBM_PEAK_GBS, 4), "bytes": st.table_bytes, "ms": round(ms, 3)}
    if k:
        issue = k.get("valu_issue")
        out.update({"frac": round(min(1.0, issue), 4) if issue is not None else None, "achieved": None, "peak": None,
                    "unit": "VALU issue slots", "valu_issue_measured": issue, "valu_busy_counter": k.get("valu_busy"),
                    "hbm_frac": round(ach / HBM_PEAK_GBS, 4), "hbm_achieved_gbs": round(ach, 2),
                    "valu_insts_per_mark": k.get("valu_lane_insts_per_mark"), "marks": k.get("marks"),
                    "dram_bytes": k["dram_read_bytes"] + k["dram_write_bytes"]})
    return out


# ---- GPU legs ---------------------------------------------------------------------
def config_times(expect):
    """M(1e12) dense (C1) and M(1e15) streaming (C2) with the reference's
    parameters and the static partition: cold (fresh context, first call) and
    warm (median of 3, table rebuilt every call), as wall time and M(X)/s."""
    rows = {}
    for name, X, strat in (("C1_1e12_dense", 10**12, mdr_hip.STRATEGY_DENSE),
                           ("C2_1e15_streaming", 10**15, mdr_hip.STRATEGY_STREAMING)):
        o = mdr_hip.make_opts(strategy=strat, partition=mdr_hip.PARTITION_STATIC, no_table_reuse=1)
        t0 = time.perf_counter()
        with mdr_hip.Engine(0) as e:
            r = e.mertens(X, o)
            t_cold = time.perf_counter() - t0
            ts, st = [], None
            for _ in range(3):
                t1 = time.perf_counter()
                assert e.mertens(X, o) == r
                ts.append(time.perf_counter() - t1)
                st = e.last_stats
        if X in expect:
            assert r == expect[X], (X, r)
        warm = statistics.median(ts)
        rows[name] = {"X": X, "result": r, "cold_s": round(t_cold, 4), "warm_s_median3": round(warm, 5),
                      "evals_per_s": round(1.0 / warm, 2), "U": st.U, "cx": st.cx,
                      "table_ms": round(st.table_ms, 3), "quotient_ms": round(st.quotient_ms, 3)}
        log("config", name, rows[name])
    return rows


def batch_times(E, ref, one_at_a_time=True):
    """The 4096-X batch workload (log-uniform in [1e9, 1e14], seed 3): one call
    per launch form (eager, graph, static partition), checked against the
    compiled reference's values and each other, and 256 of the X one call
    each (own table) for the batching gain."""
    xs = load_json("xlists.json")["batch_4096_1e9_1e14_seed3"]
    kw = dict(params=mdr_hip.PARAMS_TUNED, sched=mdr_hip.SCHED_DYNAMIC, no_table_reuse=1)
    bo = mdr_hip.make_opts(partition=mdr_hip.PARTITION_LEARNED, **kw)
    bg = mdr_hip.make_opts(partition=mdr_hip.PARTITION_LEARNED, use_graph=1, **kw)
    bs = mdr_hip.make_opts(partition=mdr_hip.PARTITION_STATIC, **kw)
    E.batch(xs[:64], bo)
    E.batch(xs[:64], bg)
    t0 = time.perf_counter()
    out = E.batch(xs, bo)
    tb = time.perf_counter() - t0
    bst = E.last_stats
    t0 = time.perf_counter()
    outg = E.batch(xs, bg)
    tg = time.perf_counter() - t0
    t0 = time.perf_counter()
    outs = E.batch(xs, bs)
    ts = time.perf_counter() - t0
    checked = [i for i, x in enumerate(xs) if x in ref]
    bad = sum(out[i] != ref[xs[i]] for i in checked) + sum(a != b for a, b in zip(out, outg))
    bad += sum(a != b for a, b in zip(out, outs))
    batch = {"workload": "4096 X log-uniform in [1e9, 1e14], seed 3", "n": len(xs), "seconds": round(tb, 4),
             "evals_per_s": round(len(xs) / tb, 1),
             "graph_seconds": round(tg, 4), "graph_evals_per_s": round(len(xs) / tg, 1),
             "static_partition_seconds": round(ts, 4), "static_over_learned": round(ts / tb, 3),
             "partition_used": "learned" if bst.partition_used == mdr_hip.PARTITION_LEARNED else
                               "static (every X below 1e14, the cost model's training range)",
             "checked_vs_reference": len(checked),
             "table_U": bst.U, "table_ms": round(bst.table_ms, 2), "quotient_ms": round(bst.quotient_ms, 2)}
    if one_at_a_time:
        # the same X values, one independent call each (own table, eager and
        # graph launches)
        one = xs[:256]
        so = mdr_hip.make_opts(partition=mdr_hip.PARTITION_LEARNED, **kw)
        sg = mdr_hip.make_opts(partition=mdr_hip.PARTITION_LEARNED, use_graph=1, **kw)
        t0 = time.perf_counter()
        single = [E.mertens(x, so) for x in one]
        t1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        single_g = [E.mertens(x, sg) for x in one]
        t1g = time.perf_counter() - t0
        bad += sum(single[i] != out[i] for i in range(len(one))) + sum(single_g[i] != out[i] for i in range(len(one)))
        batch["one_at_a_time_evals_per_s"] = round(len(one) / t1, 1)
        batch["one_at_a_time_graph_evals_per_s"] = round(len(one) / t1g, 1)
        batch["speedup_vs_one_at_a_time"] = round((len(xs) / min(tb, tg)) / (len(one) / min(t1, t1g)), 2)
    batch["mismatches"] = int(bad)
    return batch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--x", type=int, default=10**17)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-static", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the M(1e12) / M(1e15) configuration timings")
    ap.add_argument("--no-large", action="store_true", help="skip the 64-X batch in [1e16, 1e18]")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 4096-X batch, --steps times (for kernel profiles of the batch)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path would shard by "
              "outer-term chunks with one 16-byte all-reduce per X, which is not implemented.", file=sys.stderr)
        sys.exit(2)
    known = {int(k): v for k, v in load_json("known.json").items()}
    ref = {int(k): v for k, v in load_json("ref_values.json").items()}
    if args.batch_only:
        E = mdr_hip.Engine(0)
        for i in range(args.steps):
            b = batch_times(E, ref, one_at_a_time=False)
            log("batch", i, b)
            assert b["mismatches"] == 0
        E.close()
        print(json.dumps({"batch": b}), flush=True)
        return
    expect_all = dict(ref)
    expect_all.update(known)
    X = args.x
    expect = expect_all.get(X)
    prof = traffic_profile()

    # cold: a fresh context's first call (device buffers allocated, sieving
    # primes uploaded, table built); the timed steps below are warm
    t0 = time.perf_counter()
    E = mdr_hip.Engine(0)
    cold_r = E.mertens(X, mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, partition=mdr_hip.PARTITION_LEARNED))
    t_cold = time.perf_counter() - t0
    if expect is not None:
        assert cold_r == expect, (cold_r, expect)
    log(f"cold: context + first M({X}) in {t_cold:.3f} s")
    E.close()
    E = mdr_hip.Engine(0)
    tuned = dict(params=mdr_hip.PARAMS_TUNED, no_table_reuse=1)
    # production path: learned cuts fed to the dynamic chunk queue
    learned = mdr_hip.make_opts(partition=mdr_hip.PARTITION_LEARNED, sched=mdr_hip.SCHED_DYNAMIC, **tuned)
    nchunk = E.waves()

    def timed(opts, k, tag):
        res, stats = [], []
        t0 = time.perf_counter()
        for i in range(k):
            r = E.mertens(X, opts)
            stats.append(E.last_stats)
            res.append(r)
            log(f"{tag} step {i}: M({X}) = {r}  table {E.last_stats.table_ms:.1f} ms  partition "
                f"{E.last_stats.partition_ms:.1f} ms  quotient {E.last_stats.quotient_ms:.1f} ms")
        t1 = time.perf_counter()
        if expect is not None:
            assert all(r == expect for r in res), (res, expect)
        return (t1 - t0) / k, stats

    for i in range(args.warmup):
        E.mertens(X, learned)
        log(f"warmup {i} done")
    t_learned, st_l = timed(learned, args.steps, "learned")
    used = "learned" if st_l[-1].partition_used == mdr_hip.PARTITION_LEARNED else "static (no cost model)"

    cmp = None
    if not args.no_static:
        def mode(partition, sched, chunk=0, k=args.steps, tag=""):
            o = mdr_hip.make_opts(partition=partition, sched=sched, chunk=chunk, **tuned)
            t, _ = timed(o, k, tag)
            ow = mdr_hip.make_opts(partition=partition, sched=sched, chunk=chunk, wave_span=1, **tuned)
            E.mertens(X, ow)
            s = E.last_stats
            return t, {"wall_s": round(t, 4), "quotient_ms": round(s.quotient_ms, 2),
                       "wave_busy_frac": round(s.wave_busy_frac, 4), "wave_span_ms": round(s.wave_span_ms, 2),
                       "wave_busy_ms": round(s.wave_busy_ms, 2)}
        t_sq, m_sq = mode(mdr_hip.PARTITION_STATIC, mdr_hip.SCHED_DYNAMIC, tag="static+queue")
        _, m_lq = mode(mdr_hip.PARTITION_LEARNED, mdr_hip.SCHED_DYNAMIC, k=1, tag="learned+queue (stamped)")
        t_lrr, m_lrr = mode(mdr_hip.PARTITION_LEARNED, mdr_hip.SCHED_STATIC, nchunk, tag="learned+round-robin")
        t_srr, m_srr = mode(mdr_hip.PARTITION_STATIC, mdr_hip.SCHED_STATIC, nchunk, tag="static+round-robin")
        t_tq, m_tq = mode(mdr_hip.PARTITION_TERMS, mdr_hip.SCHED_DYNAMIC, k=1, tag="terms+queue")
        m_lq["wall_s"] = round(t_learned, 4)
        cmp = {"schedule_round_robin_chunks": nchunk,
               "learned_queue": m_lq, "static_queue": m_sq, "learned_round_robin": m_lrr,
               "static_round_robin": m_srr, "terms_queue": m_tq,
               "static_over_learned_queue": round(t_sq / t_learned, 3),
               "static_over_learned_round_robin": round(t_srr / t_lrr, 3),
               "terms_over_learned_queue": round(t_tq / t_learned, 3),
               # waves are busy this fraction of the static+queue quotient launch:
               # no partition of the same work can shorten the launch by more
               "partition_gain_bound_quotient": round(1.0 / m_sq["wave_busy_frac"], 4)
               if m_sq["wave_busy_frac"] > 0 else None}
        log("partition comparison", cmp)

    configs = None if args.no_configs else config_times(expect_all)

    batch = None
    if not args.no_batch:
        batch = batch_times(E, ref)
        log("batch", batch)
        assert batch["mismatches"] == 0

    large = None
    if not args.no_large:
        # 64 X log-uniform in [1e16, 1e18], seed 5: one call sharing one table
        xs = load_json("xlists.json")["large_64_1e16_1e18_seed5"]
        kw = dict(params=mdr_hip.PARAMS_TUNED, sched=mdr_hip.SCHED_DYNAMIC, no_table_reuse=1)
        t0 = time.perf_counter()
        out = E.batch(xs, mdr_hip.make_opts(partition=mdr_hip.PARTITION_LEARNED, **kw))
        tl = time.perf_counter() - t0
        lst = E.last_stats
        t0 = time.perf_counter()
        outs = E.batch(xs, mdr_hip.make_opts(partition=mdr_hip.PARTITION_STATIC, **kw))
        tls = time.perf_counter() - t0
        checked = [i for i, x in enumerate(xs) if x in ref]
        bad = sum(out[i] != ref[xs[i]] for i in checked) + sum(a != b for a, b in zip(out, outs))
        # invariance: the median X alone, with its own table and cutoff
        mid = sorted(range(len(xs)), key=lambda i: xs[i])[len(xs) // 2]
        alone = E.mertens(xs[mid], mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, no_table_reuse=1))
        bad += int(alone != out[mid])
        large = {"workload": "64 X log-uniform in [1e16, 1e18], seed 5, one batch", "n": len(xs),
                 "seconds": round(tl, 3), "evals_per_s": round(len(xs) / tl, 2),
                 "static_partition_seconds": round(tls, 3), "static_over_learned": round(tls / tl, 3),
                 "table_U": lst.U, "table_ms": round(lst.table_ms, 1), "quotient_ms": round(lst.quotient_ms, 1),
                 "checked_vs_reference": len(checked), "checked_vs_static": len(xs), "checked_alone": 1,
                 "mismatches": bad}
        log("large batch", large)
        assert bad == 0

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cpu_threads(), expect_all)
        log("cpu", cpu)

    st = st_l[-1]
    line = {
        "metric": METRIC, "value": round(1.0 / t_learned, 5), "unit": "M(X)/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_learned * 1000, 2),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": f"single exact M({X:.0e}) incl. table build, {used} partition", "X": X,
                   "U": st.U, "u_dense": st.u_dense, "cx": st.cx, "N": st.N, "chunks": st.chunks,
                   "parallelism": "1 GPU"},
        "result": E.mertens(X, mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED)) if expect is None else expect,
        "wall_s_learned": round(t_learned, 4),
        "cold_s": round(t_cold, 4),
        "partition_used": used,
        "partition_comparison": cmp,
        "phase_ms": {"table": round(st.table_ms, 2), "partition": round(st.partition_ms, 2),
                     "quotient": round(st.quotient_ms, 2)},
        "quotient_positions": {"s1_dense": st.s1_terms - st.s1_comp_terms, "s1_comp": st.s1_comp_terms,
                               "s2": st.s2_terms},
        "cost_model": {"predicted_cycles": st.predicted_cycles, "measured_wave_cycles": st.measured_cycles,
                       "predicted_over_measured": (round(st.predicted_cycles / st.measured_cycles, 3)
                                                   if st.measured_cycles > 0 else None)},
        "roofline": quotient_roofline(st_l, X, prof),
        "roofline_table": table_roofline(st_l, X, prof),
        "configs": configs,
        "batch": batch,
        "large_batch": large,
        "cpu_baseline": cpu,
    }
    E.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
