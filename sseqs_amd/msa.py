"""``msa()`` -- the drop-in MSA search call (mirror of sseqs/msa.py:71-446).

Same arguments, defaults, file naming and return values as the reference.
Differences: the database is HBM-resident and loaded on first use instead of
at import (``db=`` selects another one); ``cascade="global"`` runs the
whole-database cascade (global top ``top_prefilter`` by prefilter, affine SW,
top ``num_msas``) instead of the reference's per-batch one; a batch whose
cascade would leave fewer than ``top_sw_affine`` affine candidates is
skipped (the reference raises inside torch.topk there); ``hit_width``
(default 500, as the reference) is the number of residues of each hit that
is aligned into the a3m, None for whole hits.
"""
from __future__ import annotations

import hashlib
import os
import pickle
import threading
import time

import torch

from .db import SeqDB, default_db
from .query import prepare
from .search import batch_search, global_search, rank_hits
from .to_a3m import to_a3m

_last: dict[int, threading.Thread] = {}


def msa(queries, filenames=None, num_msas=4096, save_raw="", verbose=True, bs=50_000_000,
        fft_rank=4, top_fft=200, top_sw=10, top_sw_affine=2, return_timings=False,
        sync_time=False, dlength=1e6, gpu: int = 0, blocking: bool = True, *,
        db: SeqDB | None = None, cascade: str = "batch", top_prefilter: int = 100_000,
        hit_width: int | None = 500):
    if not blocking:  # sseqs/msa.py:88-101
        t = _last.get(gpu)
        if t is not None and t.is_alive():
            t.join()

        def _target():
            with torch.cuda.device(gpu):
                msa(queries, filenames, num_msas, save_raw, verbose, bs, fft_rank, top_fft,
                    top_sw, top_sw_affine, return_timings, sync_time, dlength, gpu, True,
                    db=db, cascade=cascade, top_prefilter=top_prefilter, hit_width=hit_width)

        nt = threading.Thread(target=_target)
        _last[gpu] = nt
        nt.start()
        return nt

    with torch.cuda.device(gpu):
        if isinstance(queries, str):
            queries = [queries]
        if isinstance(filenames, str):
            filenames = [filenames]
        if filenames is None:  # sseqs/msa.py:113-115
            filenames = [hashlib.sha256(q.encode()).hexdigest()[:8] + ".a3m" for q in queries]
        for fn in filenames:
            if fn is not None and os.path.exists(fn):
                os.remove(fn)
        if db is None:
            db = default_db()
        timings: dict = {}
        for qi, query in enumerate(queries):
            q = prepare(query, fft_rank, device=db.device)
            t0 = time.time()
            if cascade == "batch":
                bar = None
                progress = None
                if verbose:
                    from tqdm import tqdm
                    bar = tqdm(total=db.nbytes / 1e9, unit="GB")

                    def progress(lo, nb, maxl, _bar=bar, _q=query):
                        _bar.update(nb / 1e9)
                        _bar.set_description(f"db_len={int(maxl)} q_len={len(_q)}")
                hits = batch_search(db, query, q, bs=bs, top_fft=top_fft, top_sw=top_sw,
                                    hit_width=hit_width,
                                    top_sw_affine=top_sw_affine, dlength=dlength,
                                    timings=timings, progress=progress)
                if bar is not None:
                    bar.close()
                ranked = rank_hits(hits, num_msas)
                seqs = [h.seq for h in ranked]
                raw = [query, [h.seq for h in hits], [h.linear for h in hits],
                       [h.fft for h in hits], [h.norm for h in hits]]
            elif cascade == "global":
                gh = global_search(db, q, top_prefilter=top_prefilter, top_hits=num_msas)
                width = hit_width if hit_width is not None else max(
                    1, int(db.seq_len[gh.ids].max().item()) if gh.ids.numel() else 1)
                seqs = db.strings(gh.ids, width)
                raw = [query, seqs, gh.scores.cpu().tolist(), [], gh.scores.cpu().tolist()]
            else:
                raise ValueError(f"unknown cascade {cascade!r}")
            if sync_time:
                torch.cuda.synchronize()
            timings["t_search"] = timings.get("t_search", 0.0) + time.time() - t0
            if save_raw != "":
                if verbose:
                    print(f"Saving to `{save_raw}`.")
                with open(save_raw, "wb") as fh:
                    pickle.dump(raw, fh)
            t0 = time.time()
            if filenames[qi] is not None:
                if verbose:
                    print(f"Saving result to {filenames[qi]}")
                to_a3m(query, seqs, filenames[qi])
            timings["t_a3m"] = timings.get("t_a3m", 0.0) + time.time() - t0
        if return_timings:
            return timings
