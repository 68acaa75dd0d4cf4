#!/usr/bin/env python3
"""Autotuner of the GPU split: writes tune_cache.txt beside libmdr_hip.so.

For every decade of X it measures exact M(X) calls (table build included,
nothing cached) at two points, 10^d and 10^(d+1/2), and picks by coordinate
descent
  fx      the cutoff factor, u = fx (X / ln ln X)^(2/3),
  cx      the S1/S2 split factor (A = cx sqrt(y)),
  u_dense the dense int16 tier (everything above it in compressed lines),
  chunks  the quotient-phase chunk count of the dynamic queue (0 = automatic),
  sieve   the sieve's segment length and workgroup shape,
minimising the summed time of the two points.  Then it picks the batch cutoff
factor on a log-uniform batch of X.  Every configuration must give the same
M(X) (the result does not depend on the split); a mismatch aborts the run.

Cache format (read by csrc/mdr_tune.cpp): '# ...' comments, one line per
decade 'decade fx cx chunks u_dense sieve_shape', and one line 'batch fb'.

Usage: python autotune.py [--decades 8-18] [--reps 2] [--out PATH] [--quick]
"""
import argparse
import json
import math
import os
import random
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import mdr_hip  # noqa: E402

U_CAP = 500_000_000_000   # the library's one-table cap (45 % of a 288 GB device would hold 5.2e11)
LINE = 256


def cutoff(X, fx, cx, cap=True):
    """The cutoff the library derives from fx (csrc/mdr_tune.cpp tune_lookup),
    capped at the one-table limit (the sweep measures one table there)."""
    ll = math.log(max(2.0, math.log(X)))
    u = int(fx * (X / ll) ** (2.0 / 3.0))
    u = max(u, 3 * math.isqrt(X), int(cx * math.sqrt(X)) + 1000 + 64)
    if cap:
        u = min(u, U_CAP)
    return (u + LINE - 1) // LINE * LINE


def clamp_fx(X, cfg):
    """The largest fx (3 decimals) whose cutoff at X stays within the one-table
    cap, when cfg's would exceed it.  The sweep measured the capped table for
    such an fx, while the library windows every table above the cap: the
    recorded fx then gives the library the table that was measured at 10^d."""
    fx = cfg["fx"]
    if cutoff(X, fx, cfg["cx"], cap=False) <= U_CAP:
        return fx
    ll = math.log(max(2.0, math.log(X)))
    f = math.floor(U_CAP / (X / ll) ** (2.0 / 3.0) * 1000) / 1000
    while f > 0.001 and cutoff(X, f, cfg["cx"], cap=False) > U_CAP:
        f = round(f - 0.001, 3)
    return min(fx, f)


def log(*a):
    print(*a, flush=True)


class Tuner:
    def __init__(self, E, reps):
        self.E, self.reps = E, reps
        self.expect = {}

    def time(self, X, cfg):
        # GPU time of the call (table, partition, quotient phase: the bench's
        # steady state); host-side allocation when the sweep alternates table
        # sizes is left out.  The least of the repetitions.
        reps = self.reps
        o = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, no_table_reuse=1, sched=mdr_hip.SCHED_DYNAMIC,
                              partition=mdr_hip.PARTITION_LEARNED, tune_u=cutoff(X, cfg["fx"], cfg["cx"]),
                              tune_ud=cfg["ud"], tune_cx=cfg["cx"], chunk=cfg["chunks"],
                              sieve_shape=cfg.get("sieve", 0) + 1)
        ts = []
        for _ in range(reps):
            r = self.E.mertens(X, o)
            st = self.E.last_stats
            ts.append((st.table_ms + st.partition_ms + st.quotient_ms) / 1000.0)
            if X in self.expect and self.expect[X] != r:
                raise SystemExit(f"M({X}) changed with {cfg}: {r} != {self.expect[X]}")
            self.expect[X] = r
        return min(ts)

    def cost(self, xs, cfg):
        return sum(self.time(X, cfg) for X in xs)

    def sweep(self, xs, cfg, key, values):
        best, best_t = cfg[key], None
        for v in values:
            c = dict(cfg, **{key: v})
            if key == "fx" and cutoff(max(xs), v, c["cx"]) == U_CAP and v > cfg["fx"] and \
                    cutoff(max(xs), cfg["fx"], c["cx"]) == U_CAP:
                continue   # capped: the same cutoff again
            t = self.cost(xs, c)
            log(f"  {key}={v}: {t * 1000:.2f} ms")
            if best_t is None or t < best_t:
                best, best_t = v, t
        cfg[key] = best
        return best_t


def tune_decade(T, d, waves, quick):
    xs = [10**d] + ([int(round(10 ** (d + 0.5)))] if d < 18 else [])
    cfg = {"fx": 0.7, "cx": 2.0 if d >= 13 else 0.95 if d >= 11 else 0.7, "ud": 1048576, "chunks": 0, "sieve": 0}
    log(f"decade {d}: X = {xs}")
    fxs = [0.35, 0.5, 0.7, 1.0, 1.4, 2.0] if not quick else [0.5, 0.7, 1.0]
    cxs = [0.7, 1.0, 1.35, 2.0, 2.8, 4.0] if not quick else [1.35, 2.0, 2.8]
    uds = [LINE * (2**k // LINE) for k in range(16, 24)] if not quick else [262144, 1048576, 4194304]
    chs = [0, waves * 4, waves * 8, waves * 16, waves * 32, waves * 64] if not quick else [0, waves * 32]
    T.sweep(xs, cfg, "fx", fxs)
    T.sweep(xs, cfg, "cx", cxs)
    T.sweep(xs, cfg, "ud", uds)
    # second passes around the choices (the best cutoff moves with cx and the
    # tier, and the cx grid is coarse)
    f = cfg["fx"]
    T.sweep(xs, cfg, "fx", sorted({round(f * m, 3) for m in (0.8, 1.0, 1.25)}))
    c = cfg["cx"]
    T.sweep(xs, cfg, "cx", sorted({round(min(4.9, c * m), 2) for m in (0.87, 1.0, 1.15)}))
    # sieve segment length / workgroup (0 = 49152/1024, 1 = 24576/256,
    # 2 = 98304/1024, 3 = 49152/256, 4 = 49152/512, 5 = 24576/512): the table is
    # the same, only its build time moves
    T.sweep(xs, cfg, "sieve", [0, 1, 2, 3, 4, 5] if not quick else [0, 5])
    if d >= 14:
        t = T.sweep(xs, cfg, "chunks", chs)
    else:   # small calls: the automatic count (it adapts to the work) wins
        t = T.cost(xs, cfg)
    cfg["fx"] = clamp_fx(xs[0], cfg)
    log(f"decade {d}: {cfg} -> {t * 1000:.2f} ms")
    return cfg, t


def tune_batch(E, rng_seed, n, lo, hi, quick):
    rng = random.Random(rng_seed)
    xs = [int(math.exp(math.log(lo) + rng.random() * (math.log(hi) - math.log(lo)))) for _ in range(n)]
    best, best_t, ref = 1.0, None, None
    for fb in ([0.25, 0.5, 1.0, 2.0, 4.0] if not quick else [0.5, 1.0, 2.0]):
        o = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, no_table_reuse=1, tune_batch=fb,
                              partition=mdr_hip.PARTITION_LEARNED)
        E.batch(xs[:32], o)
        t0 = time.perf_counter()
        out = E.batch(xs, o)
        t = time.perf_counter() - t0
        if ref is None:
            ref = out
        elif out != ref:
            raise SystemExit(f"batch results changed with fb={fb}")
        log(f"  batch fb={fb}: {t * 1000:.1f} ms ({n / t:.0f} evals/s)")
        if best_t is None or t < best_t:
            best, best_t = fb, t
    return best, best_t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--decades", default="8-18")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=os.path.join(HERE, "tune_cache.txt"))
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--report", default=None, help="also write a JSON report here")
    ap.add_argument("--merge", default=None, help="keep this cache's lines for decades not tuned now")
    ap.add_argument("--no-batch", action="store_true", help="keep the batch factor of --merge")
    args = ap.parse_args()
    a, b = (int(v) for v in args.decades.split("-"))
    E = mdr_hip.Engine(0)
    waves = E.waves()
    T = Tuner(E, args.reps)
    rows, report = [], {"decades": {}}
    for d in range(a, b + 1):
        cfg, t = tune_decade(T, d, waves, args.quick)
        rows.append((d, cfg))
        report["decades"][d] = dict(cfg, ms=round(t * 1000, 3))
    kept, fb = {}, None
    if args.merge and os.path.exists(args.merge):
        for line in open(args.merge):
            f = line.split()
            if not f or f[0].startswith("#"):
                continue
            if f[0] == "batch":
                fb = float(f[1])
            elif int(f[0]) not in range(a, b + 1):
                kept[int(f[0])] = {"fx": float(f[1]), "cx": float(f[2]), "chunks": int(f[3]), "ud": int(f[4]),
                                   "sieve": int(f[5]) if len(f) > 5 else 0}
    rows = sorted(list(kept.items()) + rows)
    if fb is None or not args.no_batch:
        fb, tb = tune_batch(E, 3, 4096, 10**9, 10**14, args.quick)
        report["batch"] = {"fb": fb, "ms": round(tb * 1000, 2)}
    lines = [f"# written by autotune.py on {time.strftime('%Y-%m-%d %H:%M:%S')}, {waves} resident quotient waves",
             "# decade fx cx chunks u_dense sieve_shape"]
    lines += [f"{d} {c['fx']} {c['cx']} {c['chunks']} {c['ud']} {c.get('sieve', 0)}" for d, c in rows]
    lines.append(f"batch {fb}")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    log("\n".join(lines))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    E.close()


if __name__ == "__main__":
    main()
