"""Batched sweep: many queries of mixed lengths against the resident database.

Each query runs the whole-database cascade of ``search.global_search``
(prefilter of every sequence, top ``top_prefilter``, affine SW 11/1, top
``top_hits``, device backtrack), with the stages of all queries issued
together:

* the prefilter and the candidate select of every query, one after the
  other into one reused score buffer (a 200 M-sequence score vector per
  query in flight would not fit);
* affine SW of every query's candidates in one persistent launch per
  length bucket (stripe width ceil(Lq / 32), ``sseqs_db_sw_batch``): work
  items of ``item`` candidates of one query, queued by bucket and query;
* the top ``top_hits`` of every query in one segmented select;
* the backtrack of every query's hits.

``MultiSearch`` holds every buffer the sweep needs, so the sweep allocates
nothing after set-up and can be captured once in a HIP graph and replayed
(``capture()`` / ``replay()``); an eager ``run()`` issues the same launches.
Scores, candidates, hits and alignments of a query equal those of a
single-query ``global_search`` + ``backtrack_hits`` (same kernels, same tie
rules).  Launch parameters of the SW queue and of the prefilter come from
``tuning.json`` (written by ``sseqs_amd.autotune``) or defaults.
"""
from __future__ import annotations

import json
import os
from contextlib import contextmanager
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_of
from .db import SeqDB
from .query import PreparedQuery, prepare
from .search import BacktrackWorkspace, DeviceAlignments, alignments_to_dicts
from .to_a3m import a3m_text, a3m_text_rows

TUNING_PATH = Path(__file__).resolve().parent / "tuning.json"

# stripe-width classes that share SW queue parameters
SW_STREAMS = 2  # streams the SW queues of different stripe widths alternate over
BT_STREAMS = 2  # streams the queries' backtracks alternate over (a workspace each)
SW_CLASSES = ((1, 8), (9, 16), (17, 32))
DEFAULT_SW = {"item": 1024, "blocks_per_cu": 2}
N_CU = 256


def load_tuning(path: str | os.PathLike | None = None) -> dict:
    """Launch parameters picked by the autotuner (empty when absent)."""
    p = Path(path) if path is not None else TUNING_PATH
    if not p.exists():
        return {}
    with open(p) as fh:
        return json.load(fh)


def sw_params(tuning: dict, tpt: int) -> dict:
    """SW queue parameters (item, blocks_per_cu) for a stripe width."""
    for lo, hi in SW_CLASSES:
        if lo <= tpt <= hi:
            return dict(DEFAULT_SW, **tuning.get("sw", {}).get(f"{lo}-{hi}", {}))
    return dict(DEFAULT_SW)


@contextmanager
def prefilter_env(tuning: dict, N: int):
    """The prefilter's tuning knobs for correlation size N (results do not
    depend on them), set around the launches that read them."""
    knobs = tuning.get("prefilter", {}).get(str(N), {})
    old = {}
    try:
        for k, v in knobs.items():
            if k not in os.environ:
                old[k] = None
                os.environ[k] = str(v)
        yield
    finally:
        for k in old:
            os.environ.pop(k, None)


@dataclass
class QueryHits:
    """One query's results of a sweep (views into the sweep's buffers)."""

    query: str
    candidates: torch.Tensor   # int64 ids passed to SW (ascending)
    cand_scores: torch.Tensor  # int32 affine score of every candidate
    ids: torch.Tensor          # int64 hit ids, best first
    scores: torch.Tensor       # int32 hit scores, best first
    alignments: DeviceAlignments | None

    def alignment_dicts(self) -> list[dict]:
        return alignments_to_dicts(self.alignments) if self.alignments is not None else []

    def a3m_text(self) -> str:
        """The a3m text of the hits (``to_a3m.a3m_text_rows`` of the rows)."""
        al = self.alignments
        if al is None:
            return a3m_text(self.query, [])
        return a3m_text_rows(self.query, [(al.out_q.cpu().numpy(), al.out_t.cpu().numpy(),
                                           al.info.cpu().numpy())])


class MultiSearch:
    """Buffers and launch plan of the batched sweep of ``queries`` over ``db``."""

    def __init__(self, db: SeqDB, queries: list[str], top_prefilter: int = 100_000,
                 top_hits: int = 4096, gap_open: int = 11, gap_extend: int = 1,
                 backtrack: bool = True, hit_width: int = 500, tuning: dict | None = None,
                 proposer=None):
        if not queries:
            raise ValueError("no queries")
        self.db, self.dev = db, db.device
        self.texts = list(queries)
        self.qs: list[PreparedQuery] = [prepare(q, device=self.dev) for q in queries]
        for q in self.qs:
            L = int(q.codes.numel())
            if L == 0 or L > 2048:
                raise ValueError("every query needs 1..2048 residues")
        self.nq = len(self.qs)
        # candidates per query: one count for all, or one per query (then
        # the lists are laid out back to back, query i's at c_off[i])
        kq = np.asarray(top_prefilter, dtype=np.int64).reshape(-1)
        if kq.size == 1:
            kq = np.full(self.nq, int(kq[0]), np.int64)
        if kq.size != self.nq or np.any(kq < 1):
            raise ValueError("top_prefilter: a positive count, or one per query")
        self.kc_q = np.minimum(kq, db.n)
        self.kc = int(self.kc_q.max())
        self.kh = min(int(top_hits), int(self.kc_q.min()))
        if self.kh > 4096:
            raise ValueError("top_hits is limited to 4096 per query")
        self.c_off_h = np.concatenate([[0], np.cumsum(self.kc_q)]).astype(np.int64)
        self.go, self.ge = int(gap_open), int(gap_extend)
        # candidate selection: the plain prefilter's top kc, or a proposer's
        # (an object with select(ms) filling ms.cand, e.g. learned.Proposer)
        self.proposer = proposer
        self.tuning = load_tuning() if tuning is None else tuning
        dev = self.dev
        lib = _lib.load()
        self.scores = torch.empty(db.n, dtype=torch.float32, device=dev)
        self.sel_ws = _lib.workspace(lib.sseqs_topk_ws_bytes(db.n), dev)
        self.cand = torch.empty(int(self.c_off_h[-1]), dtype=torch.int64, device=dev)
        self.cs = torch.empty(int(self.c_off_h[-1]), dtype=torch.int32, device=dev)
        codes = [q.codes.cpu().numpy() for q in self.qs]
        q_off = np.concatenate([[0], np.cumsum([c.size for c in codes])]).astype(np.int64)
        self.qcodes = torch.from_numpy(np.concatenate(codes)).to(dev)
        self.q_off = torch.from_numpy(q_off).to(dev)
        self.c_off = torch.from_numpy(self.c_off_h).to(dev)
        self.q_len = np.diff(q_off)
        self.long_queries = [i for i in range(self.nq) if self.q_len[i] > 1024]
        self.set_sw_tuning(self.tuning)
        # hits: one segmented select over all queries
        self.seg_begin = self.c_off_h[:-1].copy()
        self.hit_pos = torch.empty(self.nq * self.kh, dtype=torch.int64, device=dev)
        self.meta64 = torch.from_numpy(np.concatenate(
            [self.seg_begin, np.arange(self.nq, dtype=np.int64) * self.kh])).to(dev)
        self.meta32 = torch.from_numpy(np.concatenate(
            [self.kc_q.astype(np.int32), np.full(self.nq, self.kh, np.int32)])).to(dev)
        self.hit_ids = torch.empty(self.nq * self.kh, dtype=torch.int64, device=dev)
        self.hit_sc = torch.empty(self.nq * self.kh, dtype=torch.int32, device=dev)
        self.backtrack = bool(backtrack)
        if self.backtrack:
            lq_max = max(len(t) for t in self.texts)
            nbt = BT_STREAMS if self.nq > 1 else 1
            self.bt_wss = [BacktrackWorkspace(self.kh, lq_max, int(hit_width), dev)
                           for _ in range(nbt)]
            self.bt_ws = self.bt_wss[0]
            A = self.bt_ws.A
            self.out_q = torch.zeros((self.nq, self.kh, A), dtype=torch.uint8, device=dev)
            self.out_t = torch.zeros((self.nq, self.kh, A), dtype=torch.uint8, device=dev)
            self.info = torch.zeros((self.nq, self.kh, 6), dtype=torch.int32, device=dev)
        for q in self.qs:  # length ranges are computed (with a host read) once
            db.length_ranges(q.length)
        self.learned_groups = None
        if proposer is not None and hasattr(proposer, "prepare"):
            proposer.prepare(self)  # query-side work with host copies, before any capture
        # streams of the SW queues (default priority) and their fork/join events
        self._sw_streams = [torch.cuda.Stream(device=dev) for _ in range(SW_STREAMS)]
        self._sw_fork = torch.cuda.Event()
        self._sw_join = [torch.cuda.Event() for _ in self._sw_streams]
        # the backtracks run after SW has joined: they reuse the SW streams
        self._bt_streams = [self._sw_streams[i % SW_STREAMS] for i in range(BT_STREAMS)]
        self._bt_fork = torch.cuda.Event()
        self._bt_join = [torch.cuda.Event() for _ in self._bt_streams]
        self.graph: torch.cuda.CUDAGraph | None = None

    def set_sw_tuning(self, tuning: dict) -> None:
        """Plan the SW queues: one per stripe width (widest first), items
        (query, chunk of ``item`` candidates) in query order."""
        self.tuning = dict(self.tuning, sw=tuning.get("sw", {}))
        self.buckets = []
        lens = self.q_len
        for tpt in sorted({int((L + 31) // 32) for L in lens if L <= 1024}, reverse=True):
            qi = [i for i in range(self.nq) if lens[i] <= 1024 and (lens[i] + 31) // 32 == tpt]
            prm = sw_params(self.tuning, tpt)
            item = int(prm["item"])
            per = [-(-int(self.kc_q[i]) // item) for i in qi]
            iq = np.repeat(np.array(qi, np.int32), per)
            ij = np.concatenate([np.arange(p, dtype=np.int32) for p in per])
            grid = int(min(N_CU * int(prm["blocks_per_cu"]), iq.size))
            self.buckets.append(dict(
                tpt=tpt, item=item, grid=grid, n=int(iq.size),
                item_q=torch.from_numpy(iq).to(self.dev), item_j=torch.from_numpy(ij).to(self.dev),
                counter=torch.zeros(1, dtype=torch.int32, device=self.dev)))

    # ----------------------------------------------------------- stages --
    def _prefilter_select(self) -> None:
        if self.proposer is not None:
            self.proposer.select(self)
            return
        for i in range(self.nq):
            self.plain_select(i)

    def plain_select(self, i: int) -> None:
        """Query i's candidates from the plain prefilter: its scores of every
        sequence, then their top kc_q[i]."""
        st = stream_of(self.dev)
        q = self.qs[i]
        for N, b, e in self.db.length_ranges(q.length):
            with prefilter_env(self.tuning, N):
                call("sseqs_db_prefilter", ptr(self.db.packed), ptr(self.db.seq_off),
                     ptr(self.db.seq_len), b, e, ptr(q.qrev), q.length, ptr(q.lut_a),
                     ptr(q.lut_b_flat), 21, q.k, N, N, self.scores.data_ptr() + 4 * b, st)
        call("sseqs_topk", ptr(self.scores), _lib.F32, self.db.n, int(self.kc_q[i]), 0,
             self.cand.data_ptr() + 8 * int(self.c_off_h[i]), ptr(self.sel_ws),
             self.sel_ws.numel(), st)

    def _sw(self) -> None:
        """The SW queues, one per stripe width, alternating over SW_STREAMS
        streams forked from and joined back to the current stream, so that
        one queue's tail overlaps the next queue's start; then the queries
        past 1024 residues on the current stream."""
        main = torch.cuda.current_stream(self.dev)
        st = main.cuda_stream
        db = self.db
        side = self._sw_streams if len(self.buckets) > 1 else []
        if side:
            self._sw_fork.record(main)
            for s_ in side:
                s_.wait_event(self._sw_fork)
        for idx, b in enumerate(self.buckets):
            call("sseqs_db_sw_batch", ptr(db.packed), ptr(db.seq_off), ptr(db.seq_len),
                 ptr(self.qcodes), ptr(self.q_off), ptr(self.cand), ptr(self.c_off),
                 ptr(b["item_q"]), ptr(b["item_j"]), b["n"], b["item"], b["tpt"], b["grid"],
                 self.go, self.ge, ptr(self.cs), ptr(b["counter"]),
                 side[idx % len(side)].cuda_stream if side else st)
        if side:
            for s_, e in zip(side, self._sw_join):
                e.record(s_)
                main.wait_event(e)
        for i in self.long_queries:  # past 1024 residues: the 32-bit kernel
            q = self.qs[i]
            c0 = int(self.c_off_h[i])
            call("sseqs_db_sw_affine", ptr(db.packed), ptr(db.seq_off), ptr(db.seq_len),
                 self.cand.data_ptr() + 8 * c0, int(self.kc_q[i]), ptr(q.codes),
                 int(q.codes.numel()), self.go, self.ge, self.cs.data_ptr() + 4 * c0, st)

    def _hits(self) -> None:
        st = stream_of(self.dev)
        m = self.nq
        call("sseqs_topk_segments", ptr(self.cs), _lib.I32, ptr(self.meta64), ptr(self.meta32),
             self.meta32.data_ptr() + 4 * m, self.meta64.data_ptr() + 8 * m, m, self.kh, 1, 0, 0,
             1, ptr(self.hit_pos), st)
        torch.index_select(self.cand, 0, self.hit_pos, out=self.hit_ids)
        torch.index_select(self.cs, 0, self.hit_pos, out=self.hit_sc)

    def _backtrack(self) -> None:
        """Gather + backtrack of every query's hits.  With several queries
        the queries alternate over BT_STREAMS streams forked from and joined
        back to the current stream, each with its own workspace: one
        query's launch holds four waves per SIMD, and its longest hits are
        its tail, so two queries in flight fill the GPU better."""
        if not self.backtrack:
            return
        from .search import backtrack_hits
        main = torch.cuda.current_stream(self.dev)
        side = self._bt_streams[: len(self.bt_wss)] if len(self.bt_wss) > 1 else []
        if side:
            self._bt_fork.record(main)
            for s_ in side:
                s_.wait_event(self._bt_fork)
        for i, q in enumerate(self.qs):
            k = i % len(side) if side else 0
            with torch.cuda.stream(side[k] if side else main):
                al = backtrack_hits(self.db, q, self.hit_ids[i * self.kh:(i + 1) * self.kh],
                                    self.bt_wss[k], self.go, self.ge)
                self.out_q[i, :, : al.width].copy_(al.out_q)
                self.out_t[i, :, : al.width].copy_(al.out_t)
                self.info[i].copy_(al.info)
        if side:
            for s_, e in zip(side, self._bt_join):
                e.record(s_)
                main.wait_event(e)

    def run(self) -> None:
        """One sweep, launched eagerly on the current stream."""
        self._prefilter_select()
        self._sw()
        self._hits()
        self._backtrack()

    # ----------------------------------------------------------- graphs --
    def capture(self) -> None:
        """Capture one sweep as a HIP graph (after an eager sweep has set up
        every lazily created state)."""
        self.run()
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        torch.cuda.synchronize(self.dev)
        self.graph = g

    def replay(self) -> None:
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()

    # ---------------------------------------------------------- results --
    def results(self) -> list[QueryHits]:
        out = []
        for i, t in enumerate(self.texts):
            s = slice(int(self.c_off_h[i]), int(self.c_off_h[i + 1]))
            h = slice(i * self.kh, (i + 1) * self.kh)
            al = None
            if self.backtrack:
                A = self.bt_ws.A
                al = DeviceAlignments(self.out_q[i, :, :A], self.out_t[i, :, :A], self.info[i], A)
            out.append(QueryHits(t, self.cand[s], self.cs[s], self.hit_ids[h], self.hit_sc[h],
                                 al))
        return out


def multi_search(db: SeqDB, queries: list[str], top_prefilter: int = 100_000,
                 top_hits: int = 4096, **kw) -> list[QueryHits]:
    """One eager batched sweep (see MultiSearch)."""
    ms = MultiSearch(db, queries, top_prefilter, top_hits, **kw)
    ms.run()
    return ms.results()
