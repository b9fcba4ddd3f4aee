"""Train the learned prefilter's encoder and save its weights.

``python -m sseqs_amd.train_encoder [--pairs-db 5000000 --queries 5000
--randoms 1500 --steps 40000 --time-budget 330 --out PATH]`` builds the
synthetic training set (every second sequence of a --pairs-db database is a
planted homolog of one of --queries queries with lengths uniform in
[50, 1000]; each query is paired with all its homologs and --randoms
background sequences, about 10 M pairs with the defaults, scored by exact
affine SW), trains with the HIP forward/backward kernels and AdamW, and
writes the weights as an .npz of float32 arrays (E, W1, b1, W2, b2, W3, b3).
The training queries and database use seeds distinct from the benchmark's.
Every --mine-every steps the encoder's own hardest non-homologs of each
query are mined and fill --hard-frac of the background pairs.  After
training, ``learned.recall_eval`` measures the proposal's recall on a scaled
copy of the batched workload's database (--eval-db sequences, its own
seeds); ``--eval-only W1.npz [W2.npz ...]`` only evaluates saved weights.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch


def main(argv=None) -> int:
    from .learned import WEIGHTS_PATH, Encoder, make_training_set, recall_eval, train
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--pairs-db", type=int, default=5_000_000)
    ap.add_argument("--queries", type=int, default=5000)
    ap.add_argument("--randoms", type=int, default=1500)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--queries-per-step", type=int, default=32)
    ap.add_argument("--pairs-per-query", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--time-budget", type=float, default=330.0)
    ap.add_argument("--homolog-weight", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--mse-weight", type=float, default=1.0)
    ap.add_argument("--nce-weight", type=float, default=1.0)
    ap.add_argument("--tau", type=float, default=0.05)
    ap.add_argument("--tau-y", type=float, default=0.1)
    ap.add_argument("--mine-every", type=int, default=5000)
    ap.add_argument("--hard-per-query", type=int, default=64)
    ap.add_argument("--hard-frac", type=float, default=0.5)
    ap.add_argument("--eval-db", type=int, default=40_000_000)
    ap.add_argument("--eval-queries", type=int, default=32)
    ap.add_argument("--eval-only", nargs="+", default=None, metavar="WEIGHTS")
    ap.add_argument("--init", default=None, metavar="WEIGHTS",
                    help="start from saved weights instead of a random initialisation")
    ap.add_argument("--out", default=str(WEIGHTS_PATH))
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")

    def evaluate(enc):
        if args.eval_db <= 0:
            return None
        r = recall_eval(enc, args.eval_db, args.eval_queries, log=log)
        return {k: v for k, v in r.items() if k != "per_query"}

    if args.eval_only:
        for w in args.eval_only:
            res = evaluate(Encoder.load(w, device=dev))
            print(json.dumps({"weights": w, "eval": res}), flush=True)
        return 0
    t0 = time.time()
    ts = make_training_set(args.pairs_db, args.queries, args.randoms, seed=args.seed, device=dev)
    log(f"[train] {ts.y.size} pairs over {len(ts.queries)} queries in {time.time() - t0:.1f} s; "
        f"mean y {float(ts.y.mean()):.3f}, homolog-like (y > 0.3) {float((ts.y > 0.3).mean()):.3f}")
    enc = Encoder.load(args.init, device=dev) if args.init else Encoder(dev, seed=args.seed)
    t1 = time.time()
    losses = train(enc, ts, args.steps, args.queries_per_step, args.pairs_per_query, args.lr,
                   seed=args.seed, log=log, time_budget_s=args.time_budget,
                   homolog_weight=args.homolog_weight, mine_every=args.mine_every,
                   hard_per_query=args.hard_per_query, hard_frac=args.hard_frac,
                   mse_weight=args.mse_weight, nce_weight=args.nce_weight, tau=args.tau,
                   tau_y=args.tau_y)
    secs = time.time() - t1
    enc.save(args.out)
    ev = evaluate(enc)
    print(json.dumps({"pairs": int(ts.y.size), "queries": len(ts.queries),
                      "steps": enc.step_count, "train_s": round(secs, 1),
                      "pairs_seen": enc.step_count * args.queries_per_step * args.pairs_per_query,
                      "loss_first": losses[0], "loss_last": losses[-1], "out": args.out,
                      "eval": ev}),
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
