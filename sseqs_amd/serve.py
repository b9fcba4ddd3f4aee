"""Continuous-batching search engine for the MSA server.

The reference server runs one ``msa()`` call per HTTP request inside its
event loop (server.py:31-110), so requests wait for each other whole.  Here
the database stays resident in HBM and a device loop admits queued queries
into batched sweeps (``sseqs_amd.multi.MultiSearch``: prefilter or learned
proposal, SW, hits and backtrack of every query of the sweep):

* admission: a sweep takes the oldest waiting queries while their device
  buffers (candidate lists and scores, hit lists, backtrack rows and DP
  workspace) fit the HBM budget of an in-flight sweep, up to ``max_batch``;
* overlap: a sweep runs in two stages on two streams (of the same, default
  priority) -- candidate selection on the first, SW + hits + backtrack on
  the second -- so the next sweep's selection starts while the previous
  sweep's SW and backtrack still run, and the host formats finished
  sweeps' a3m text (the library's host formatter, which releases the
  interpreter lock) on two worker threads while the device works on the
  next;
  at most ``in_flight`` sweeps hold buffers at once, so a sweep is admitted
  when its predecessor frees them;
* a query's hits, alignments and a3m text do not depend on which other
  queries share its sweep (every kernel computes each query independently
  and ties are broken by index), so a request gets the same answer alone
  and in the mix.

``submit(query)`` returns a ``concurrent.futures.Future`` of the a3m text
(the format ``to_a3m`` writes); ``latencies`` records submit-to-result
times for reporting p50/p99.
"""
from __future__ import annotations

import concurrent.futures as cf
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import torch

from . import _lib
from .db import SeqDB
from .multi import BT_STREAMS, MultiSearch


@dataclass
class _Request:
    query: str
    future: cf.Future
    t_submit: float
    t_start: float = 0.0
    t_done: float = 0.0


@dataclass
class EngineStats:
    sweeps: int = 0
    queries: int = 0
    batch_sizes: list = field(default_factory=list)


def query_bytes(db: SeqDB, q_len: int, top_prefilter: int, top_hits: int,
                hit_width: int = 500) -> int:
    """Device bytes one query holds in a sweep: candidate ids and scores,
    hit ids/scores/positions, and its backtrack rows (aligned query and
    target, right-aligned, and the alignment info)."""
    kc = min(top_prefilter, db.n)
    kh = min(top_hits, kc)
    A = q_len + hit_width + 1
    return kc * 12 + kh * 20 + kh * A * 2 + kh * 24


def sweep_bytes(db: SeqDB, lq_max: int, top_hits: int, hit_width: int = 500) -> int:
    """Device bytes a sweep holds besides its queries: one prefilter score
    vector, the select workspace and the backtrack workspaces (one per
    backtrack stream)."""
    lib = _lib.load()
    kh = top_hits
    bt = (int(lib.sseqs_backtrack_ws_bytes(lq_max, lq_max * kh * hit_width, kh)) +
          kh * (hit_width + lq_max + 1) * 2 + kh * (hit_width + 8))
    return db.n * 4 + int(lib.sseqs_topk_ws_bytes(db.n)) + BT_STREAMS * bt


class SearchEngine:
    """Device loop of the continuous-batching server (see module doc)."""

    def __init__(self, db: SeqDB, top_prefilter: int = 100_000, top_hits: int = 4096,
                 max_batch: int = 64, hbm_budget: int | None = None, in_flight: int = 2,
                 proposer=None, hit_width: int = 500, max_wait_s: float = 0.0):
        self.db = db
        # candidates per query: a count, or "calibrated" for the proposer's
        # operating point by query length (Proposer.calibrate)
        if top_prefilter == "calibrated":
            if proposer is None or getattr(proposer, "k_by_len", None) is None:
                raise ValueError("top_prefilter='calibrated' needs a calibrated proposer")
            self.kc = "calibrated"
        else:
            self.kc = int(top_prefilter)
        self.kh = int(top_hits)
        self.max_batch = int(max_batch)
        self.in_flight = max(1, int(in_flight))
        self.proposer = proposer
        self.hit_width = int(hit_width)
        self.max_wait_s = float(max_wait_s)
        if proposer is not None and hasattr(proposer, "_buffers"):
            proposer._buffers(db.device)  # its score buffers before the budget is taken
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(db.device)
            hbm_budget = int(0.5 * free) // self.in_flight
        self.budget = int(hbm_budget)
        self.stats = EngineStats()
        self.latencies: list[float] = []
        self._q: deque[_Request] = deque()
        self._cv = threading.Condition()
        self._slots = threading.Semaphore(self.in_flight)
        self._stop = False
        self._streams = [torch.cuda.Stream(db.device), torch.cuda.Stream(db.device)]
        self._fin = cf.ThreadPoolExecutor(max_workers=2, thread_name_prefix="sseqs-finish")
        self._loop = threading.Thread(target=self._run, name="sseqs-device-loop", daemon=True)
        self._loop.start()

    # ------------------------------------------------------------ client --
    def submit(self, query: str) -> cf.Future:
        fut: cf.Future = cf.Future()
        if not query or not any(c in "ARNDCQEGHILKMFPSTWYV" for c in query):
            fut.set_exception(ValueError("query has no residues"))
            return fut
        with self._cv:
            self._q.append(_Request(query, fut, time.perf_counter()))
            self._cv.notify()
        return fut

    def search(self, query: str, timeout: float | None = None) -> str:
        return self.submit(query).result(timeout)

    def close(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify()
        self._loop.join()
        self._fin.shutdown(wait=True)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -------------------------------------------------------- admission --
    def _admit(self) -> list[_Request]:
        """The oldest waiting requests whose buffers fit the budget."""
        batch: list[_Request] = []
        used = 0
        lq_max = 0
        while self._q and len(batch) < self.max_batch:
            r = self._q[0]
            lq = max(lq_max, len(r.query))
            qb = query_bytes(self.db, len(r.query), self._kc(r.query), self.kh, self.hit_width)
            need = sweep_bytes(self.db, lq, self.kh, self.hit_width) + used + qb
            if batch and need > self.budget:
                break
            batch.append(self._q.popleft())
            used += qb
            lq_max = lq
        return batch

    def _kc(self, query: str) -> int:
        """Candidates passed to SW for one query."""
        if self.kc == "calibrated":
            return int(self.proposer.top_prefilter_for([query])[0])
        return self.kc

    # ------------------------------------------------------------- loop --
    def _run(self) -> None:
        dev = self.db.device
        torch.cuda.set_device(dev)
        while True:
            with self._cv:
                while not self._q and not self._stop:
                    self._cv.wait()
                if self._stop and not self._q:
                    return
            if self.max_wait_s > 0:
                time.sleep(self.max_wait_s)  # let a burst gather into one sweep
            self._slots.acquire()  # a free set of sweep buffers
            with self._cv:
                batch = self._admit()
            if not batch:
                self._slots.release()
                continue
            t0 = time.perf_counter()
            for r in batch:
                r.t_start = t0
            try:
                ms = MultiSearch(self.db, [r.query for r in batch],
                                 [self._kc(r.query) for r in batch], self.kh,
                                 hit_width=self.hit_width, proposer=self.proposer)
                s1, s2 = self._streams
                # the sweep's inputs were copied to the device on this stream
                s1.wait_stream(torch.cuda.current_stream(dev))
                with torch.cuda.stream(s1):
                    ms._prefilter_select()
                    ev1 = torch.cuda.Event()
                    ev1.record(s1)
                with torch.cuda.stream(s2):
                    s2.wait_event(ev1)
                    ms._sw()
                    ms._hits()
                    ms._backtrack()
                    done = torch.cuda.Event()
                    done.record(s2)
            except Exception as e:  # noqa: BLE001 - delivered to the callers
                for r in batch:
                    r.future.set_exception(e)
                self._slots.release()
                continue
            self.stats.sweeps += 1
            self.stats.queries += len(batch)
            self.stats.batch_sizes.append(len(batch))
            self._fin.submit(self._finish, ms, batch, done)

    def _finish(self, ms: MultiSearch, batch: list[_Request], done) -> None:
        try:
            done.synchronize()
            res = ms.results()
            texts = []
            for r, h in zip(batch, res):
                texts.append(h.a3m_text())
        except Exception as e:  # noqa: BLE001 - delivered to the callers
            for r in batch:
                r.future.set_exception(e)
            return
        finally:
            del ms
            self._slots.release()
        now = time.perf_counter()
        for r, t in zip(batch, texts):
            r.t_done = now
            self.latencies.append(now - r.t_submit)
            r.future.set_result(t)
