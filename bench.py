#!/usr/bin/env python3
"""Benchmark of the MI355X Deleglise-Rivat engine.

Metric: M(X) wall time at X = 1e17, learned versus static S2 balance;
batched M(X) evaluations per second.

One step = one exact M(1e17) through the C ABI, table build included (the
table is rebuilt every step; X is the only input and lives on the host as an
integer argument, so nothing needs staging).  `value` is evaluations per
second over the K timed steps.  Beside it:
  - the static-partition times of the same X: equal quotient positions per
    chunk, and whole terms per chunk (the reference GPU engine's unit), under
    the dynamic chunk queue and under a one-chunk-per-wave round robin,
  - batched evals/s for the 4096-X workload (seed 3, [1e9, 1e14]) against
    one-at-a-time calls,
  - the quotient kernel's roofline (algorithmic bytes / kernel time),
  - the compiled reference CPU engine timed on this host (cpu_baseline).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "deleglise-rivat_amd"))

import mdr_hip  # noqa: E402

METRIC = "M(X) wall time at X=1e17 learned vs static S2 balance; batched M(X) evals/sec"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per quotient position: S1 dense-tier lookup 2 B (int16),
# S1 compressed-tier lookup 20 B (two 8-byte planes + a 4-byte checkpoint),
# S2 position 3 B (the reference's mu byte stream, A bytes per A/3 positions).
BYTES_S1_DENSE, BYTES_S1_COMP, BYTES_S2 = 2.0, 20.0, 3.0
# Reference CPU engine: t(M(1e17)) / t(M(1e16)) measured on 8 threads of the
# build container (174.93 s / 20.99 s); used to scale the bounded 1e16 sample.
REF_RATIO_17_16 = 174.93 / 20.99


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_json(name):
    with open(os.path.join(ROOT, "tests", "golden", name)) as f:
        return json.load(f)


def cpu_baseline(threads):
    exe = os.path.join(ROOT, "oracle", "_ref", "bench_dr")
    X = 10**16
    if os.path.exists(exe):
        out = subprocess.run([exe, str(X), str(threads)], capture_output=True, text=True, timeout=600).stdout
        t = float(re.search(r"Elapsed time: ([0-9.]+)", out).group(1))
        m = int(re.search(r"M\(\d+\) = (-?\d+)", out).group(1))
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        m, t = oracle.time_mertens(X, threads)
        kind = "port"
    est17 = t * REF_RATIO_17_16
    return {"value": 1.0 / est17, "unit": "M(X)/s", "cores": threads, "kind": kind,
            "sample": f"M(1e16) = {m} in {t:.2f} s on {threads} threads, scaled to 1e17 by the reference's "
                      f"own 1e17/1e16 time ratio {REF_RATIO_17_16:.3f} (est. {est17:.1f} s per M(1e17))",
            "seconds_1e16": t, "est_seconds_1e17": est17}


def roofline(stats):
    """k_quot over the timed steps: algorithmic bytes per launch / mean launch
    time (hipEvents recorded around the launch on the engine's stream)."""
    st = stats[-1]
    byts = (BYTES_S1_DENSE * (st.s1_terms - st.s1_comp_terms) + BYTES_S1_COMP * st.s1_comp_terms
            + BYTES_S2 * st.s2_terms)
    ms = sum(s.quotient_ms for s in stats) / len(stats)
    sec = ms / 1000.0
    ach = byts / sec / 1e9
    traffic, issue = None, {}
    tf = os.path.join(ROOT, "profiles", "traffic_k_quot_1e17.json")
    if os.path.exists(tf):
        prof = json.load(open(tf))
        traffic = prof.get("hbm_bytes_per_launch")
        issue = prof.get("issue", {}).get("k_quot", {})
    return {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "k_quot",
            "algorithmic_bytes": byts, "kernel_ms": round(ms, 3), "launches": len(stats),
            "valu_issue_frac": issue.get("valu_issue_frac"), "valu_busy": issue.get("valu_busy"),
            "l2_hit_rate": issue.get("l2_hit_rate"),
            "note": "per-unit bytes count every compressed lookup as a 20-byte HBM fetch; L2/MALL serve part of "
                    "them (traffic, from FETCH/WRITE_SIZE, is below the algorithmic bytes), so frac can exceed 1; "
                    "the binding resource is VALU issue: valu_busy (SQ_ACTIVE_INST_VALU over the SIMDs' cycles) is the measured "
                    "fraction of SIMD cycles spent issuing VALU work, valu_issue_frac the same from SQ_INSTS_VALU at 4 cycles each"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--x", type=int, default=10**17)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-static", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the 64-X batch in [1e16, 1e18]")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path would shard by "
              "outer-term chunks with one 16-byte all-reduce per X, which this build does not run.", file=sys.stderr)
        sys.exit(2)
    known = {int(k): v for k, v in load_json("known.json").items()}
    ref = {int(k): v for k, v in load_json("ref_values.json").items()}
    X = args.x
    expect = known.get(X, ref.get(X))

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
    # production path: learned cuts fed to the dynamic chunk queue
    learned = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, partition=mdr_hip.PARTITION_LEARNED,
                                sched=mdr_hip.SCHED_DYNAMIC, no_table_reuse=1)
    # partition comparison under the static round-robin schedule (one chunk per wave)
    nchunk = E.waves()
    learned_rr = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, partition=mdr_hip.PARTITION_LEARNED,
                                   sched=mdr_hip.SCHED_STATIC, chunk=nchunk, no_table_reuse=1)
    static_rr = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, partition=mdr_hip.PARTITION_STATIC,
                                  sched=mdr_hip.SCHED_STATIC, chunk=nchunk, no_table_reuse=1)
    static_q = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, partition=mdr_hip.PARTITION_STATIC,
                                 sched=mdr_hip.SCHED_DYNAMIC, no_table_reuse=1)
    # the reference GPU engine's partition: whole terms (G evaluations) per chunk
    terms_q = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, partition=mdr_hip.PARTITION_TERMS,
                                sched=mdr_hip.SCHED_DYNAMIC, no_table_reuse=1)

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
        t_sq, _ = timed(static_q, args.steps, "static+queue")
        t_lrr, _ = timed(learned_rr, args.steps, "learned+round-robin")
        t_srr, _ = timed(static_rr, args.steps, "static+round-robin")
        t_tq, _ = timed(terms_q, 1, "terms+queue")
        cmp = {"schedule_round_robin_chunks": nchunk,
               "wall_s_terms_queue": round(t_tq, 4), "terms_over_learned_queue": round(t_tq / t_learned, 3),
               "wall_s_learned_round_robin": round(t_lrr, 4), "wall_s_static_round_robin": round(t_srr, 4),
               "static_over_learned_round_robin": round(t_srr / t_lrr, 3),
               "wall_s_static_queue": round(t_sq, 4), "static_over_learned_queue": round(t_sq / t_learned, 3)}
        log("partition comparison", cmp)

    batch = None
    if not args.no_batch:
        xs = load_json("xlists.json")["batch_4096_1e9_1e14_seed3"]
        kw = dict(params=mdr_hip.PARAMS_TUNED, partition=mdr_hip.PARTITION_LEARNED, sched=mdr_hip.SCHED_DYNAMIC)
        bo = mdr_hip.make_opts(no_table_reuse=1, **kw)
        bg = mdr_hip.make_opts(no_table_reuse=1, use_graph=1, **kw)
        E.batch(xs[:64], bo)
        E.batch(xs[:64], bg)
        t0 = time.perf_counter()
        out = E.batch(xs, bo)
        tb = time.perf_counter() - t0
        bst = E.last_stats
        t0 = time.perf_counter()
        outg = E.batch(xs, bg)
        tg = time.perf_counter() - t0
        checked = [i for i, x in enumerate(xs) if x in ref]
        bad = sum(out[i] != ref[xs[i]] for i in checked) + sum(a != b for a, b in zip(out, outg))
        # one-at-a-time: the same X values, one independent call each (own table,
        # eager and graph launches)
        one = xs[:256]
        so = mdr_hip.make_opts(no_table_reuse=1, **kw)
        sg = mdr_hip.make_opts(no_table_reuse=1, use_graph=1, **kw)
        t0 = time.perf_counter()
        single = [E.mertens(x, so) for x in one]
        t1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        single_g = [E.mertens(x, sg) for x in one]
        t1g = time.perf_counter() - t0
        bad += sum(single[i] != out[i] for i in range(len(one))) + sum(single_g[i] != out[i] for i in range(len(one)))
        batch = {"workload": "4096 X log-uniform in [1e9, 1e14], seed 3", "n": len(xs), "seconds": round(tb, 4),
                 "evals_per_s": round(len(xs) / tb, 1),
                 "graph_seconds": round(tg, 4), "graph_evals_per_s": round(len(xs) / tg, 1),
                 "one_at_a_time_evals_per_s": round(len(one) / t1, 1),
                 "one_at_a_time_graph_evals_per_s": round(len(one) / t1g, 1),
                 "speedup_vs_one_at_a_time": round((len(xs) / min(tb, tg)) / (len(one) / min(t1, t1g)), 2),
                 "checked_vs_reference": len(checked), "mismatches": int(bad),
                 "table_U": bst.U, "table_ms": round(bst.table_ms, 2), "quotient_ms": round(bst.quotient_ms, 2)}
        log("batch", batch)
        assert bad == 0

    large = None
    if not args.no_large:
        # 64 X log-uniform in [1e16, 1e18], seed 5: one call sharing one table
        xs = load_json("xlists.json")["large_64_1e16_1e18_seed5"]
        lo = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, partition=mdr_hip.PARTITION_LEARNED,
                               sched=mdr_hip.SCHED_DYNAMIC, no_table_reuse=1)
        t0 = time.perf_counter()
        out = E.batch(xs, lo)
        tl = time.perf_counter() - t0
        lst = E.last_stats
        checked = [i for i, x in enumerate(xs) if x in ref]
        bad = sum(out[i] != ref[xs[i]] for i in checked)
        # invariance: the median X alone, with its own table and cutoff
        mid = sorted(range(len(xs)), key=lambda i: xs[i])[len(xs) // 2]
        alone = E.mertens(xs[mid], mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, no_table_reuse=1))
        bad += int(alone != out[mid])
        large = {"workload": "64 X log-uniform in [1e16, 1e18], seed 5, one batch", "n": len(xs),
                 "seconds": round(tl, 3), "evals_per_s": round(len(xs) / tl, 2), "table_U": lst.U,
                 "table_ms": round(lst.table_ms, 1), "quotient_ms": round(lst.quotient_ms, 1),
                 "checked_vs_reference": len(checked), "checked_alone": 1, "mismatches": bad}
        log("large batch", large)
        assert bad == 0

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(threads)
        log("cpu", cpu)

    st = st_l[-1]
    t_table = sum(s.table_ms for s in st_l) / len(st_l)
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
        "roofline": roofline(st_l),
        "roofline_table": {"bound": "hbm", "achieved": round(st.table_bytes / (t_table / 1000) / 1e9, 2),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(st.table_bytes / (t_table / 1000) / 1e9 / HBM_PEAK_GBS, 4),
                           "bytes": st.table_bytes, "ms": round(t_table, 3)},
        "batch": batch,
        "large_batch": large,
        "cpu_baseline": cpu,
    }
    E.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
