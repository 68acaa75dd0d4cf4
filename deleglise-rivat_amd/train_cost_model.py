#!/usr/bin/env python3
"""Trains the quotient-range cost model on the GPU and writes cost_model.bin.

1. Runs the plain kernels (static partition) over the 512 seeded log-uniform
   X in [1e14, 1e16] (tests/golden/xlists.json, train_512_1e14_1e16_seed4) as
   one batch with the cycle log on: every cost piece of a sample of G
   evaluations records its shader-clock cycles.
2. Features/targets log2(cycles per position) per piece; 10% held out.
3. Trains the 5-32-32-1 network with the engine's own GPU kernels (Adam).
4. Reports the held-out error and writes the weights beside the library,
   where the engine loads them at start-up.

Usage: python deleglise-rivat_amd/train_cost_model.py [--out PATH] [--epochs N]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
import mdr_hip  # noqa: E402


def collect(engine, xs, stride):
    opts = mdr_hip.make_opts(params=mdr_hip.PARAMS_TUNED, partition=mdr_hip.PARTITION_STATIC,
                             sched=mdr_hip.SCHED_STATIC, record_cycles=stride)
    t0 = time.time()
    engine.batch(xs, opts)
    log = engine.cycle_log()
    print(f"  batch of {len(xs)} X: {time.time() - t0:.2f} s, {len(log)} cycle records", flush=True)
    return log


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "cost_model.bin"))
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--stride", type=int, default=4)
    ap.add_argument("--min-count", type=int, default=64)
    ap.add_argument("--report", default=None, help="write a JSON summary here")
    args = ap.parse_args()
    xs = json.load(open(os.path.join(ROOT, "tests", "golden", "xlists.json")))["train_512_1e14_1e16_seed4"]
    with mdr_hip.Engine(0) as e:
        e.cost_model_set(None)
        log = collect(e, xs, args.stride)
        keep = [s for s in log if s.count >= args.min_count and s.cycles > 0]
        arr = (mdr_hip.Stamp * len(keep))(*keep)
        F, T = mdr_hip.mlp_features(arr)
        rng = np.random.default_rng(0)
        perm = rng.permutation(len(F))
        nval = len(F) // 10
        val, tr = perm[:nval], perm[nval:]
        t0 = time.time()
        loss = e.mlp_train(F[tr], T[tr], epochs=args.epochs, batch=1024, lr=2e-3, seed=1)
        w = e.cost_model_get()
        t_train = time.time() - t0
        pv = e.mlp_predict(w, F[val])
        rmse = float(np.sqrt(np.mean((pv - T[val]) ** 2)))
        base = float(np.sqrt(np.mean((T[val] - T[tr].mean()) ** 2)))
        # error of predicted total cycles per held-out piece
        cnt = np.array([keep[i].count for i in val], dtype=np.float64)
        cyc = np.array([keep[i].cycles for i in val], dtype=np.float64)
        rel = np.abs(cnt * np.exp2(pv) - cyc) / cyc
        e.cost_model_save(args.out)
        rep = {"records": len(log), "used": len(F), "train": len(tr), "val": int(nval),
               "train_loss_log2": loss, "val_rmse_log2": rmse, "val_rmse_log2_mean_predictor": base,
               "val_median_rel_err_cycles": float(np.median(rel)), "train_seconds": t_train, "out": args.out}
        print(json.dumps(rep), flush=True)
        if args.report:
            with open(args.report, "w") as f:
                json.dump(rep, f, indent=1)


if __name__ == "__main__":
    main()
