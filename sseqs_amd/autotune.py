"""Autotuner of the batched sweep's launch parameters.

Times, on the GPU it runs on, the choices that the results do not depend on:

* the SW work queue (``sseqs_db_sw_batch``) per stripe-width class
  (1-8, 9-16, 17-32 columns per lane): ``item`` (candidates per work item,
  the queue's tile width) and ``blocks_per_cu`` (persistent blocks per
  compute unit, i.e. how many queries' items a CU works on at once);
* the prefilter's dense kernels (queries whose band does not fit the band
  kernel) per correlation size N: ``SSEQS_CORR_MT`` (m-tiles per pass, i.e.
  how many shift tiles reuse each decoded database fragment).

Each candidate is timed with HIP events on the launch stream over the
workload it is tuned for (queries of that class against the database), the
fastest is kept, and the table is written to ``sseqs_amd/tuning.json``,
which ``sseqs_amd.multi`` reads.  Run as
``python -m sseqs_amd.autotune [--n-seqs N]`` on the target GPU.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import torch

from .multi import SW_CLASSES, TUNING_PATH, MultiSearch

ITEMS = (64, 128, 256, 512, 1024, 2048, 4096)
BLOCKS_PER_CU = (1, 2, 4, 8)
CORR_MT = (2, 4, 8)


def _time(fn, reps: int = 3) -> float:
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_sw(db, queries_by_class: dict, top_prefilter: int, log=print) -> dict:
    """Best (item, blocks_per_cu) per stripe-width class: the candidates of
    each class's queries are selected once, then only the SW stage is
    timed for every choice."""
    out, table = {}, {}
    for (lo, hi), qs in queries_by_class.items():
        if not qs:
            continue
        key = f"{lo}-{hi}"
        ms = MultiSearch(db, qs, top_prefilter, 16, backtrack=False, tuning={})
        ms._prefilter_select()
        ms._sw()
        ref = ms.cs.clone()
        best = None
        for item in ITEMS:
            for bpc in BLOCKS_PER_CU:
                ms.set_sw_tuning({"sw": {key: {"item": item, "blocks_per_cu": bpc}}})
                ms.cs.fill_(-2)
                ms_time = _time(ms._sw)
                if not torch.equal(ms.cs, ref):
                    raise RuntimeError(f"SW scores depend on the queue parameters ({key})")
                table[f"{key} item={item} bpc={bpc}"] = round(ms_time, 3)
                log(f"[autotune] sw {key} item={item} blocks/CU={bpc}: {ms_time:.3f} ms")
                if best is None or ms_time < best[0]:
                    best = (ms_time, item, bpc)
        out[key] = {"item": best[1], "blocks_per_cu": best[2], "ms": round(best[0], 3)}
    return out, table


def tune_prefilter(db, queries: list[str], log=print) -> dict:
    """Best SSEQS_CORR_MT per correlation size for the queries whose
    prefilter runs on the dense kernels."""
    import os

    from .query import prepare
    from .search import prefilter_kernel_name
    out, table = {}, {}
    scores = torch.empty(db.n, dtype=torch.float32, device=db.device)
    from ._lib import call, ptr, stream_of
    for query in queries:
        q = prepare(query, device=db.device)
        for N, b, e in db.length_ranges(q.length):
            if prefilter_kernel_name(q.length, q.k, N).startswith("k_corr_band"):
                continue
            if str(N) in out:
                continue
            res = {}
            for mt in CORR_MT:
                os.environ["SSEQS_CORR_MT"] = str(mt)

                def run():
                    call("sseqs_db_prefilter", ptr(db.packed), ptr(db.seq_off), ptr(db.seq_len),
                         b, e, ptr(q.qrev), q.length, ptr(q.lut_a), ptr(q.lut_b_flat), 21, q.k,
                         N, N, scores.data_ptr() + 4 * b, stream_of(db.device))
                res[mt] = _time(run)
                table[f"N={N} Lq={q.length} MT={mt}"] = round(res[mt], 3)
                log(f"[autotune] prefilter N={N} Lq={q.length} MT={mt}: {res[mt]:.3f} ms")
            os.environ.pop("SSEQS_CORR_MT", None)
            mt = min(res, key=res.get)
            out[str(N)] = {"SSEQS_CORR_MT": mt}
    return out, table


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n-seqs", type=int, default=20_000_000)
    ap.add_argument("--top-prefilter", type=int, default=100_000)
    ap.add_argument("--queries-per-class", type=int, default=8)
    ap.add_argument("--out", default=str(TUNING_PATH))
    args = ap.parse_args(argv)
    from . import synth
    dev = torch.device("cuda:0")
    t0 = time.time()
    db, _, _, batch = synth.bench_db(args.n_seqs, device=dev)
    print(f"[autotune] database {db.n} sequences in {time.time() - t0:.1f} s", file=sys.stderr)
    by_class = {c: [] for c in SW_CLASSES}
    for qtext in batch:
        tpt = (len(qtext) + 31) // 32
        for lo, hi in SW_CLASSES:
            if lo <= tpt <= hi and len(by_class[(lo, hi)]) < args.queries_per_class:
                by_class[(lo, hi)].append(qtext)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    sw, sw_table = tune_sw(db, by_class, args.top_prefilter, log)
    long_q = sorted(batch, key=len)[-2:]
    pf, pf_table = tune_prefilter(db, long_q, log)
    res = {"gpu": torch.cuda.get_device_name(dev), "n_seqs": db.n,
           "top_prefilter": args.top_prefilter, "sw": sw, "prefilter": pf,
           "timings_ms": {**sw_table, **pf_table}}
    with open(args.out, "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)
    print(json.dumps({"sw": sw, "prefilter": pf}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
