"""Train the learned prefilter's encoder and save its weights.

``python -m sseqs_amd.train_encoder [--pairs-db 5000000 --queries 5000
--randoms 1500 --steps 50000 --time-budget 300 --out PATH]`` builds the
synthetic training set (every second sequence of a --pairs-db database is a
planted homolog of one of --queries queries with lengths uniform in
[50, 1000]; each query is paired with all its homologs and --randoms
background sequences, about 10 M pairs with the defaults, scored by exact
affine SW), trains with the HIP forward/backward kernels and AdamW, and
writes the weights as an .npz of float32 arrays (E, W1, b1, W2, b2, W3, b3).
The training queries and database use seeds distinct from the benchmark's.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch


def main(argv=None) -> int:
    from .learned import WEIGHTS_PATH, Encoder, make_training_set, train
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--pairs-db", type=int, default=5_000_000)
    ap.add_argument("--queries", type=int, default=5000)
    ap.add_argument("--randoms", type=int, default=1500)
    ap.add_argument("--steps", type=int, default=50000)
    ap.add_argument("--queries-per-step", type=int, default=32)
    ap.add_argument("--pairs-per-query", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--time-budget", type=float, default=300.0)
    ap.add_argument("--homolog-weight", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--out", default=str(WEIGHTS_PATH))
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    t0 = time.time()
    ts = make_training_set(args.pairs_db, args.queries, args.randoms, seed=args.seed, device=dev)
    log(f"[train] {ts.y.size} pairs over {len(ts.queries)} queries in {time.time() - t0:.1f} s; "
        f"mean y {float(ts.y.mean()):.3f}, homolog-like (y > 0.3) {float((ts.y > 0.3).mean()):.3f}")
    enc = Encoder(dev, seed=args.seed)
    t1 = time.time()
    losses = train(enc, ts, args.steps, args.queries_per_step, args.pairs_per_query, args.lr,
                   seed=args.seed, log=log, time_budget_s=args.time_budget,
                   homolog_weight=args.homolog_weight)
    secs = time.time() - t1
    enc.save(args.out)
    print(json.dumps({"pairs": int(ts.y.size), "queries": len(ts.queries),
                      "steps": enc.step_count, "train_s": round(secs, 1),
                      "pairs_seen": enc.step_count * args.queries_per_step * args.pairs_per_query,
                      "loss_first": losses[0], "loss_last": losses[-1], "out": args.out}),
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
