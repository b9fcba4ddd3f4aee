"""Learned prefilter: a convolutional sequence encoder, an FP8 embedding table
of the whole database and an MFMA proposal of candidates for exact SW.

The encoder (kernels in csrc/learned.hip): residue embedding 20 -> 64, three
conv1d layers of width 9 (64 -> 128 -> 128 -> 256) with GELU, mean pool over
positions and L2 normalisation; the similarity of a query and a database
sequence is the cosine of their embeddings.  It is trained on (query,
target, normalised SW score) pairs -- targets are planted homologs and
background sequences of a synthetic database, the score is the affine SW
score (reference stripe rule, 11/1) divided by the query's score against
itself -- with hand-written forward and backward kernels (bf16 operands,
fp32 accumulation) and AdamW.  The loss is the squared error between cosine
and score plus a soft-target InfoNCE term over the step's sequences (each
query against every target of the step, its own pairs weighted by
exp(score / 0.1)), with hard negatives mined from the training database by
the encoder itself.

Every database sequence is embedded once into an HBM-resident table of
256 FP8 (e4m3) values per sequence (51.2 GB for 200 M sequences).  A batch
of queries is embedded the same way and the proposal GEMM
(``sseqs_propose`` / ``sseqs_propose_max``, MFMA fp8) scores it against the
whole table; each query's top ``top_prefilter`` go to exact affine SW, which
alone decides the hits.
``Proposer`` plugs this into ``multi.MultiSearch`` (and so into the serving
engine); its ``calibrate`` measures Recall@4096 against exhaustive SW on
sample queries and sets the candidate count per query length that reaches
the gate (0.99); a query given fewer candidates than that falls back to the
plain prefilter.

There is no counterpart in the reference, which only has the FFT prefilter
(sseqs/msa.py:270-310); the candidates this proposes replace that
prefilter's stage-1 selection.
"""
from __future__ import annotations

import json
import math
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_of
from .ops import topk_indices
from .query import AA_MAP

WEIGHTS_PATH = Path(__file__).resolve().parent / "encoder_weights.npz"
C0, C1, C2, C3 = 64, 128, 128, 256
DIM = C3
FP8_SCALE = 16.0
PAD_CODE = 21
# parameter blocks of the flat fp32 vector: name -> shape
SHAPES = [("E", (22, C0)), ("W1", (9, C0, C1)), ("b1", (C1,)), ("W2", (9, C1, C2)),
          ("b2", (C2,)), ("W3", (9, C2, C3)), ("b3", (C3,))]


def codes_of(seq: str) -> np.ndarray:
    """Encoder codes: residues 0..19, 20 for any other letter."""
    return np.array([AA_MAP.get(c, 20) for c in seq], dtype=np.uint8)


@dataclass
class Rows:
    """Flat row layout of a batch of sequences (four pad rows before each
    sequence and after the last, total rounded up to a multiple of 128)."""

    M: int
    row0: torch.Tensor   # int64 [nseq] first row of each sequence
    length: torch.Tensor  # int32 [nseq]
    row0_h: np.ndarray
    len_h: np.ndarray

    @staticmethod
    def plan(lengths: np.ndarray, device) -> "Rows":
        ln = np.asarray(lengths, dtype=np.int64)
        row0 = (4 + np.concatenate([[0], np.cumsum(ln + 4)[:-1]]) if ln.size
                else np.zeros(0, np.int64))
        M = int((row0[-1] + ln[-1] + 4) if ln.size else 128)
        M = (M + 127) // 128 * 128
        return Rows(M, torch.from_numpy(row0.astype(np.int64)).to(device),
                    torch.from_numpy(ln.astype(np.int32)).to(device), row0.astype(np.int64),
                    ln.astype(np.int32))


class Encoder:
    """Encoder weights on the device, with forward (and, for training,
    backward and AdamW) over batches of sequences."""

    def __init__(self, device="cuda", seed: int = 0, weights: dict | None = None):
        self.dev = torch.device(device)
        sizes = [int(np.prod(s)) for _, s in SHAPES]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.n_params = int(self.offsets[-1])
        self.p = torch.zeros(self.n_params, dtype=torch.float32, device=self.dev)
        if weights is None:
            rng = np.random.default_rng(seed)
            init = {}
            for name, shape in SHAPES:
                if name == "E":
                    v = rng.normal(0.0, 1.0, shape)
                    v[20:] = 0.0
                elif name.startswith("W"):
                    fan_in = shape[0] * shape[1]
                    v = rng.normal(0.0, math.sqrt(2.0 / fan_in), shape)
                else:
                    v = np.zeros(shape)
                init[name] = v
            weights = init
        self.load_state(weights)
        self.packed = {}
        self._pack()
        self.g = self.m = self.v = None
        self.step_count = 0

    # ----------------------------------------------------------- params --
    def view(self, t: torch.Tensor, name: str) -> torch.Tensor:
        i = [n for n, _ in SHAPES].index(name)
        return t[self.offsets[i]: self.offsets[i + 1]].view(SHAPES[i][1])

    def load_state(self, weights: dict) -> None:
        for name, shape in SHAPES:
            v = np.asarray(weights[name], dtype=np.float32).reshape(shape)
            self.view(self.p, name).copy_(torch.from_numpy(v))

    def state(self) -> dict:
        return {name: self.view(self.p, name).cpu().numpy() for name, _ in SHAPES}

    def save(self, path=WEIGHTS_PATH) -> None:
        np.savez(path, **self.state())

    @staticmethod
    def load(path=WEIGHTS_PATH, device="cuda") -> "Encoder":
        with np.load(path) as f:  # arrays only (allow_pickle stays off)
            return Encoder(device, weights={k: f[k] for k in f.files})

    def _pack(self) -> None:
        """bf16 weight layouts of the forward and backward GEMMs."""
        st = stream_of(self.dev)
        for name, cin, cout in (("W1", C0, C1), ("W2", C1, C2), ("W3", C2, C3)):
            if name not in self.packed:
                self.packed[name] = (
                    torch.empty(9 * cin * cout, dtype=torch.bfloat16, device=self.dev),
                    torch.empty(9 * cin * cout, dtype=torch.bfloat16, device=self.dev))
            f, b = self.packed[name]
            call("sseqs_enc_pack_weights", ptr(self.view(self.p, name)), cin, cout, ptr(f), ptr(b),
                 st)

    # ---------------------------------------------------------- buffers --
    def _buffers(self, M: int, train: bool) -> dict:
        """Activations with four zero margin rows at both ends (the conv
        windows read rows -4 .. M + 3)."""
        dev = self.dev
        z = lambda C: torch.zeros((M + 8, C), dtype=torch.bfloat16, device=dev)
        b = {"x0": z(C0), "h1": z(C1), "h2": z(C2),
             "h3": torch.empty((M, C3), dtype=torch.bfloat16, device=dev)}
        if train:
            b.update({"u1": torch.empty((M, C1), dtype=torch.bfloat16, device=dev),
                      "u2": torch.empty((M, C2), dtype=torch.bfloat16, device=dev),
                      "u3": torch.empty((M, C3), dtype=torch.bfloat16, device=dev),
                      "du3": z(C3), "du2": z(C2), "du1": z(C1),
                      "dx0": torch.empty((M, C0), dtype=torch.float32, device=dev)})
        return b

    @staticmethod
    def _row(t: torch.Tensor) -> int:
        """Pointer to row 0 of a buffer with four margin rows."""
        return t.data_ptr() + 4 * t.shape[1] * t.element_size()

    def _conv(self, A, lda, Bt, M, N, K, bias, mask, U, Y, Uprev, Yf, epi, st) -> None:
        call("sseqs_enc_conv", A, lda, ptr(Bt), M, N, K, bias, ptr(mask), U, Y, Uprev, Yf, epi,
             st)

    # ---------------------------------------------------------- forward --
    def forward(self, code: torch.Tensor, mask: torch.Tensor, rows: Rows, train: bool = False,
                buf: dict | None = None):
        """Embeddings e (float32 [nseq, 256], unit norm) of the sequences laid
        out in ``rows`` (codes and mask per row).  Returns (e, znorm, buf)."""
        st = stream_of(self.dev)
        M = rows.M
        b = buf if buf is not None else self._buffers(M, train)
        call("sseqs_enc_embed", ptr(code), M, ptr(self.view(self.p, "E")), self._row(b["x0"]), st)
        layers = (("W1", "b1", "x0", "h1", "u1", C0, C1), ("W2", "b2", "h1", "h2", "u2", C1, C2),
                  ("W3", "b3", "h2", "h3", "u3", C2, C3))
        for wn, bn, xin, yout, uout, cin, cout in layers:
            yp = ptr(b[yout]) if yout == "h3" else self._row(b[yout])
            self._conv(self._row(b[xin]), cin, self.packed[wn][0], M, cout, 9 * cin,
                       ptr(self.view(self.p, bn)), mask, ptr(b[uout]) if train else None, yp, None,
                       None, 0, st)
        n = int(rows.row0.numel())
        e = torch.empty((n, DIM), dtype=torch.float32, device=self.dev)
        znorm = torch.empty(n, dtype=torch.float32, device=self.dev)
        call("sseqs_enc_pool", ptr(b["h3"]), ptr(rows.row0), ptr(rows.length), n, ptr(e),
             ptr(znorm), st)
        return e, znorm, b

    # --------------------------------------------------------- backward --
    def backward(self, de: torch.Tensor, e: torch.Tensor, znorm: torch.Tensor, code, mask,
                 rows: Rows, b: dict) -> None:
        """Gradients of the loss whose gradient with respect to e is ``de``,
        accumulated into self.g."""
        st = stream_of(self.dev)
        M = rows.M
        if self.g is None:
            self.g = torch.zeros_like(self.p)
        g = self.g
        n = int(rows.row0.numel())
        b["du3"].zero_()
        call("sseqs_enc_pool_grad", ptr(de), ptr(e), ptr(znorm), ptr(b["u3"]), ptr(rows.row0),
             ptr(rows.length), n, self._row(b["du3"]), st)
        # layer 3
        call("sseqs_enc_bias_grad", self._row(b["du3"]), C3, M, ptr(self.view(g, "b3")), st)
        call("sseqs_enc_wgrad", self._row(b["h2"]), C2, self._row(b["du3"]), C3, M,
             ptr(self.view(g, "W3")), st)
        self._conv(self._row(b["du3"]), C3, self.packed["W3"][1], M, C2, 9 * C3, None, mask, None,
                   self._row(b["du2"]), ptr(b["u2"]), None, 1, st)
        # layer 2
        call("sseqs_enc_bias_grad", self._row(b["du2"]), C2, M, ptr(self.view(g, "b2")), st)
        call("sseqs_enc_wgrad", self._row(b["h1"]), C1, self._row(b["du2"]), C2, M,
             ptr(self.view(g, "W2")), st)
        self._conv(self._row(b["du2"]), C2, self.packed["W2"][1], M, C1, 9 * C2, None, mask, None,
                   self._row(b["du1"]), ptr(b["u1"]), None, 1, st)
        # layer 1 and the embedding
        call("sseqs_enc_bias_grad", self._row(b["du1"]), C1, M, ptr(self.view(g, "b1")), st)
        call("sseqs_enc_wgrad", self._row(b["x0"]), C0, self._row(b["du1"]), C1, M,
             ptr(self.view(g, "W1")), st)
        self._conv(self._row(b["du1"]), C1, self.packed["W1"][1], M, C0, 9 * C1, None, mask, None,
                   None, None, ptr(b["dx0"]), 2, st)
        call("sseqs_enc_embed_grad", ptr(code), M, ptr(b["dx0"]), ptr(self.view(g, "E")), st)

    def adam_step(self, lr: float, wd: float = 0.0, b1: float = 0.9, b2: float = 0.999,
                  eps: float = 1e-8) -> None:
        if self.m is None:
            self.m = torch.zeros_like(self.p)
            self.v = torch.zeros_like(self.p)
        self.step_count += 1
        call("sseqs_enc_adam", ptr(self.p), ptr(self.g), ptr(self.m), ptr(self.v), self.n_params,
             float(lr), float(b1), float(b2), float(eps), float(wd), self.step_count,
             stream_of(self.dev))
        self._pack()
        self.g.zero_()

    # ---------------------------------------------------------- helpers --
    def rows_of_strings(self, seqs: list[str]):
        """(code, mask, rows) of host strings."""
        lens = np.array([len(s) for s in seqs], dtype=np.int64)
        rows = Rows.plan(lens, self.dev)
        code = np.full(rows.M, PAD_CODE, dtype=np.uint8)
        mask = np.zeros(rows.M, dtype=np.uint8)
        for s, r0 in zip(seqs, rows.row0_h):
            c = codes_of(s)
            code[r0: r0 + c.size] = c
            mask[r0: r0 + c.size] = 1
        return (torch.from_numpy(code).to(self.dev), torch.from_numpy(mask).to(self.dev), rows)

    def rows_of_db(self, db, ids: np.ndarray, max_len: int = 2048):
        """(code, mask, rows) of database sequences ``ids`` (host int64)."""
        ids = np.asarray(ids, dtype=np.int64)
        lens = np.minimum(db.seq_len[torch.from_numpy(ids).to(db.device)].cpu().numpy(), max_len)
        rows = Rows.plan(lens, self.dev)
        code = torch.full((rows.M,), PAD_CODE, dtype=torch.uint8, device=self.dev)
        mask = torch.zeros(rows.M, dtype=torch.uint8, device=self.dev)
        d_ids = torch.from_numpy(ids).to(db.device)
        off = db.seq_off[d_ids].contiguous()
        ln = db.seq_len[d_ids].contiguous()
        call("sseqs_enc_db_rows", ptr(db.packed), ptr(off), ptr(ln), 0, ids.size, ptr(rows.row0),
             int(max_len), ptr(code), ptr(mask), stream_of(self.dev))
        return code, mask, rows

    def embed_strings(self, seqs: list[str]) -> torch.Tensor:
        code, mask, rows = self.rows_of_strings(seqs)
        return self.forward(code, mask, rows)[0]


# ------------------------------------------------------------ training --

@dataclass
class TrainingSet:
    """(query, target, y) pairs over a synthetic training database."""

    db: object
    queries: list[str]
    q_ids: np.ndarray     # int32 [npair] query index
    t_ids: np.ndarray     # int64 [npair] database id of the target
    y: np.ndarray         # float32 [npair] SW(q, t) / SW(q, q)
    homolog: np.ndarray   # bool [npair] the target is a planted homolog of the query
    selfs: np.ndarray | None = None  # float32 [nquery] SW(q, q)


def make_training_set(n_db: int, n_queries: int, randoms_per_query: int, seed: int = 11,
                      device="cuda", lo: int = 50, hi: int = 1000,
                      hom_period: int = 2) -> TrainingSet:
    """Synthetic training pairs: a database of n_db sequences in which every
    hom_period-th is a planted homolog of one of n_queries queries (lengths
    uniform in [lo, hi]; 30-90 % identity, indels), each query paired with
    all its homologs and with randoms_per_query background sequences, scored
    by affine SW (11/1, reference stripe rule) and normalised by the
    query's score against itself."""
    from . import synth
    from .db import SeqDB
    from .multi import MultiSearch
    from .sw import sw
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, size=n_queries)
    queries = [synth.make_query(int(L), 100_000 + i, seed=seed) for i, L in enumerate(lens)]
    db = SeqDB.synthetic(n_db, queries, seed=seed, device=device, hom_period=hom_period)
    hq = db.extra["hom_query"].cpu().numpy()
    bg = np.nonzero(hq < 0)[0]
    order = np.argsort(hq, kind="stable")
    hq_sorted = hq[order]
    starts = np.searchsorted(hq_sorted, np.arange(n_queries))
    ends = np.searchsorted(hq_sorted, np.arange(n_queries), side="right")
    cand, qidx, ishom = [], [], []
    for i in range(n_queries):
        hom = order[starts[i]:ends[i]]
        rnd = rng.choice(bg, size=randoms_per_query, replace=False)
        c = np.concatenate([hom, rnd])
        srt = np.argsort(c, kind="stable")
        cand.append(c[srt])
        ishom.append((np.arange(c.size) < hom.size)[srt])
        qidx.append(np.full(c.size, i, np.int32))
    selfs = np.array([float(sw(q, [q])[0]) for q in queries], dtype=np.float32)
    scores = sw_pairs(db, queries, cand)
    q_all = np.concatenate(qidx)
    y = np.clip(scores / np.maximum(selfs[q_all], 1.0), 0.0, 1.0).astype(np.float32)
    ts = TrainingSet(db, queries, q_all, np.concatenate(cand), y, np.concatenate(ishom))
    ts.selfs = selfs
    return ts


def sw_pairs(db, queries: list[str], cand: list[np.ndarray]) -> np.ndarray:
    """Affine SW (11/1, reference stripe rule) of query i against database
    sequences cand[i] (ascending ids), for every i, with the batched SW
    queue; the scores of all pairs concatenated (float32)."""
    from .multi import MultiSearch
    ms = MultiSearch(db, queries, top_prefilter=1, top_hits=1, backtrack=False, tuning={})
    counts = np.array([c.size for c in cand], dtype=np.int64)
    c_all = np.concatenate(cand).astype(np.int64)
    ms.kc, ms.kc_q = int(counts.max()), counts
    ms.c_off_h = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    ms.cand = torch.from_numpy(c_all).to(db.device)
    ms.cs = torch.empty(c_all.size, dtype=torch.int32, device=db.device)
    ms.c_off = torch.from_numpy(np.concatenate([[0], np.cumsum(counts)])).to(db.device)
    ms.set_sw_tuning({})
    # per-query candidate counts differ: items past a query's list are empty
    ms._sw()
    return ms.cs.cpu().numpy().astype(np.float32)


def mine_hard(enc: Encoder, ts: TrainingSet, per_query: int, chunk: int = 128,
              log=None) -> tuple[list[np.ndarray], list[np.ndarray]]:
    """Hard negatives: for every training query, the ``per_query`` database
    sequences that are not its planted homologs and have the highest
    proposal score under the current encoder, with their normalised SW
    scores.  Returns (ids per query, y per query)."""
    from .ops import topk_segments
    t0 = time.time()
    tab = EmbeddingTable(enc, ts.db)
    prop = Proposer(tab)
    nq, n = len(ts.queries), ts.db.n
    lo = np.searchsorted(ts.q_ids, np.arange(nq))
    hi = np.searchsorted(ts.q_ids, np.arange(nq), side="right")
    ids = []
    for c0 in range(0, nq, chunk):
        c1 = min(nq, c0 + chunk)
        sc = prop.scores_for(ts.queries[c0:c1])
        for j, i in enumerate(range(c0, c1)):
            own = ts.t_ids[lo[i]:hi[i]][ts.homolog[lo[i]:hi[i]]]
            if own.size:
                sc[j, torch.from_numpy(own).to(sc.device)] = float("-inf")
        sel, base = topk_segments(sc, np.arange(c1 - c0) * n, np.full(c1 - c0, n),
                                  np.full(c1 - c0, per_query))
        sel = sel.cpu().numpy()
        ids += [sel[base[j]: base[j] + per_query] for j in range(c1 - c0)]
        del sc
    del prop, tab
    torch.cuda.empty_cache()
    scores = sw_pairs(ts.db, ts.queries, ids)
    ys = np.clip(scores.reshape(nq, per_query) / np.maximum(ts.selfs[:, None], 1.0), 0.0, 1.0)
    if log:
        log(f"[mine] {per_query} hard negatives per query in {time.time() - t0:.1f} s; "
            f"mean y {float(ys.mean()):.3f}, max y {float(ys.max()):.3f}")
    return ids, [ys[i].astype(np.float32) for i in range(nq)]


def train(enc: Encoder, ts: TrainingSet, steps: int, queries_per_step: int = 32,
          pairs_per_query: int = 32, lr: float = 1e-3, seed: int = 0, log=None,
          time_budget_s: float | None = None, homolog_weight: float = 4.0,
          lr_floor: float = 0.05, mine_every: int = 0, hard_per_query: int = 64,
          hard_frac: float = 0.5, mse_weight: float = 1.0, nce_weight: float = 0.0,
          tau: float = 0.05, tau_y: float = 0.1) -> list[float]:
    """AdamW on ``mse_weight`` times the squared error between cosine and
    normalised SW score, homolog pairs weighted ``homolog_weight`` (the
    proposal must rank the homologs of a query among themselves, background
    pairs only need to stay low), plus ``nce_weight`` times a soft-target
    InfoNCE loss: each query's logits are its cosines / ``tau`` with every
    target of the step (the other queries' targets are its negatives), and
    its target distribution weights its own pairs by exp(y / ``tau_y``), so
    that the strongest homologs are pushed above everything else in the
    step.  Each step takes queries_per_step queries and pairs_per_query
    of each query's pairs, half homologs and half background where
    available.  With ``mine_every`` > 0, every mine_every steps the
    ``hard_per_query`` highest-scoring non-homologs of each query under the
    current encoder are mined (``mine_hard``) and then fill ``hard_frac`` of
    the background slots.  The learning rate warms up over 200 steps, then
    follows a cosine to lr_floor * lr at the last step."""
    rng = np.random.default_rng(seed)
    nq = len(ts.queries)
    # pairs are stored query by query: each query's range, split by kind
    lo = np.searchsorted(ts.q_ids, np.arange(nq))
    hi = np.searchsorted(ts.q_ids, np.arange(nq), side="right")
    by_q_h = [lo[i] + np.nonzero(ts.homolog[lo[i]:hi[i]])[0] for i in range(nq)]
    by_q_b = [lo[i] + np.nonzero(~ts.homolog[lo[i]:hi[i]])[0] for i in range(nq)]
    hard_ids = hard_y = None
    losses = []
    t0 = time.time()
    half = pairs_per_query // 2

    def pick(i):
        """(target ids, y, weights) of query i's pairs this step."""
        h = rng.choice(by_q_h[i], size=min(half, by_q_h[i].size), replace=False)
        nb = pairs_per_query - h.size
        nh = int(round(nb * hard_frac)) if hard_ids is not None else 0
        b = rng.choice(by_q_b[i], size=min(nb - nh, by_q_b[i].size), replace=False)
        t = [ts.t_ids[h], ts.t_ids[b]]
        y = [ts.y[h], ts.y[b]]
        w = [np.full(h.size, homolog_weight, np.float32), np.ones(b.size, np.float32)]
        if nh:
            k = rng.choice(hard_ids[i].size, size=min(nh, hard_ids[i].size), replace=False)
            t.append(hard_ids[i][k])
            y.append(hard_y[i][k])
            w.append(np.ones(k.size, np.float32))
        return np.concatenate(t), np.concatenate(y), np.concatenate(w)

    for step in range(steps):
        if mine_every > 0 and step > 0 and step % mine_every == 0:
            hard_ids, hard_y = mine_hard(enc, ts, hard_per_query, log=log)
        qs = rng.choice(nq, size=min(queries_per_step, nq), replace=False)
        picks = [pick(i) for i in qs]
        tids = np.concatenate([p[0] for p in picks])
        qpos = np.concatenate([np.full(p[0].size, k, np.int32) for k, p in enumerate(picks)])
        # sequences of the step: the queries, then the targets
        qcode, qmask, qrows = enc.rows_of_strings([ts.queries[i] for i in qs])
        tcode, tmask, trows = enc.rows_of_db(ts.db, tids)
        code = torch.cat([qcode, tcode])
        mask = torch.cat([qmask, tmask])
        rows = Rows(qrows.M + trows.M, torch.cat([qrows.row0, trows.row0 + qrows.M]),
                    torch.cat([qrows.length, trows.length]),
                    np.concatenate([qrows.row0_h, trows.row0_h + qrows.M]),
                    np.concatenate([qrows.len_h, trows.len_h]))
        e, znorm, buf = enc.forward(code, mask, rows, train=True)
        qi = torch.from_numpy(qpos).to(enc.dev)
        ti = torch.arange(len(qs), len(qs) + tids.size, dtype=torch.int32, device=enc.dev)
        y = torch.from_numpy(np.concatenate([p[1] for p in picks])).to(enc.dev)
        wt = torch.from_numpy(np.concatenate([p[2] for p in picks])).to(enc.dev)
        de = torch.zeros_like(e)
        loss = torch.zeros(1, dtype=torch.float32, device=enc.dev)
        st = stream_of(enc.dev)
        if mse_weight > 0:
            wt *= mse_weight
            call("sseqs_enc_pair_loss", ptr(e), int(e.shape[0]), ptr(qi), ptr(ti), ptr(y),
                 ptr(wt), tids.size, ptr(de), None, ptr(loss), st)
        if nce_weight > 0:
            P = np.zeros((len(qs), tids.size), np.float32)
            c0 = 0
            for k, p in enumerate(picks):
                w = np.exp((p[1] - p[1].max()) / tau_y)
                P[k, c0: c0 + w.size] = w / w.sum()
                c0 += w.size
            Pd = torch.from_numpy(P).to(enc.dev)
            G = torch.empty_like(Pd)
            call("sseqs_enc_infonce", ptr(e), len(qs), tids.size, ptr(Pd), float(tau),
                 nce_weight / len(qs), ptr(G), ptr(de), ptr(loss), st)
        enc.backward(de, e, znorm, code, mask, rows, buf)
        warm = min(1.0, (step + 1) / 200)
        cos = lr_floor + (1 - lr_floor) * 0.5 * (1 + math.cos(math.pi * step / max(1, steps - 1)))
        enc.adam_step(lr * warm * cos)
        if step % 50 == 0 or step == steps - 1:
            losses.append(float(loss.item()))
            if log:
                log(f"[train] step {step} loss {losses[-1]:.5f} ({time.time() - t0:.1f} s)")
        if time_budget_s is not None and time.time() - t0 > time_budget_s:
            break
    return losses


def recall_eval(enc: Encoder, n_db: int = 40_000_000, n_eval: int = 32, seed: int = 5,
                query_seed: int = 7, log=None) -> dict:
    """Recall of the learned proposal on a scaled copy of the batched
    workload: the benchmark database construction (synth.bench_db) with
    n_db sequences and its own seeds, so that each batched query owns the
    same fraction of homologs, and top-k and the candidate count scaled by
    s = n_db / 200 M (k = 4096 s, K = 100 000 s).  For n_eval batched queries
    (lengths uniform in [50, 1000]): recall of the exhaustive SW top-k among
    the top K proposals, the candidate count that reaches 0.99 recall, and
    what the top K hold (own homologs, homologs of the two large families of
    the database, other homologs, background)."""
    from . import synth
    s = n_db / 200e6
    k, K = max(64, int(round(4096 * s))), int(round(100_000 * s))
    db, _, _, batch = synth.bench_db(n_db, device=enc.dev, seed=seed, query_seed=query_seed)
    tab = EmbeddingTable(enc, db)
    prop = Proposer(tab)
    owner = db.extra["hom_owner"]
    per = []
    for i in np.unique(np.linspace(0, len(batch) - 1, n_eval).round().astype(int)):
        q = batch[i]
        kth, allsc = exhaustive_top(db, q, k)
        sc = prop.scores_for([q])[0]
        curve, k99 = recall_curve(sc, allsc >= kth, k, (K,))
        o = owner[topk_indices(sc, K, sorted=False)]
        per.append({"q_len": len(q), "recall": curve[K], "k99_frac": k99 / db.n,
                    "own": int((o == 2 + i).sum()), "family": int(((o == 0) | (o == 1)).sum()),
                    "other_hom": int(((o >= 2) & (o != 2 + i)).sum()), "background": int((o < 0).sum())})
        del allsc, sc, o
        if log:
            log(f"[eval] {json.dumps(per[-1])}")
    del prop, tab, db
    torch.cuda.empty_cache()
    rec = np.array([p["recall"] for p in per])
    kf = np.array([p["k99_frac"] for p in per])
    return {"n_db": n_db, "k": k, "K": K, "queries": len(per),
            "recall_mean": round(float(rec.mean()), 4), "recall_min": round(float(rec.min()), 4),
            "recall_p10": round(float(np.quantile(rec, 0.1)), 4),
            "k99_frac_median": float(np.median(kf)), "k99_frac_mean": float(kf.mean()),
            "k99_frac_max": float(kf.max()), "per_query": per}


# --------------------------------------------------------------- table --

class EmbeddingTable:
    """FP8 embeddings of every sequence of a resident database (index
    order), built once by the encoder in chunks of at most ``rows`` rows
    (residues plus four pad rows per sequence)."""

    def __init__(self, enc: Encoder, db, rows: int = 48_000_000, max_len: int = 2048, log=None):
        self.enc, self.db = enc, db
        self.n = db.n
        self.table = torch.empty((db.n, DIM), dtype=torch.uint8, device=db.device)
        t0 = time.time()
        st = stream_of(db.device)
        # chunk boundaries: consecutive sequences whose rows fit `rows`
        need = np.minimum(db.seq_len.cpu().numpy().astype(np.int64), max_len) + 4
        cum = np.cumsum(need)
        bounds = [0]
        while bounds[-1] < db.n:
            base = cum[bounds[-1] - 1] if bounds[-1] else 0
            nxt = int(np.searchsorted(cum, base + rows, side="right"))
            bounds.append(max(nxt, bounds[-1] + 1))
        bounds[-1] = db.n
        Mcap = int(max(cum[e - 1] - (cum[b - 1] if b else 0) for b, e in zip(bounds, bounds[1:])))
        Mcap = (Mcap + 4 + 127) // 128 * 128
        bufs = enc._buffers(Mcap, train=False)
        self.rows_total = int(cum[-1]) if db.n else 0
        for ci, (b, e_) in enumerate(zip(bounds, bounds[1:])):
            code, mask, rws = enc.rows_of_db(db, np.arange(b, e_, dtype=np.int64), max_len)
            view = {}
            for k, t in bufs.items():
                if k in ("x0", "h1", "h2"):
                    view[k] = t[: rws.M + 8]
                    t[rws.M + 4: rws.M + 8].zero_()  # the margin after this chunk's rows
                else:
                    view[k] = t[: rws.M]
            e, _, _ = enc.forward(code, mask, rws, buf=view)
            call("sseqs_enc_to_fp8", ptr(e), e.numel(), FP8_SCALE,
                 self.table.data_ptr() + b * DIM, st)
            if log and ci % 50 == 0:
                log(f"[table] {e_}/{db.n} sequences ({time.time() - t0:.1f} s)")
        torch.cuda.synchronize(db.device)
        self.build_s = time.time() - t0
        del bufs

    @property
    def nbytes(self) -> int:
        return self.table.numel()

    @property
    def flops(self) -> float:
        """Encoder MFMA flops of the build: 2 * 9 * (64*128 + 128*128 +
        128*256) per row."""
        return 2.0 * 9 * (C0 * C1 + C1 * C2 + C2 * C3) * self.rows_total


# ------------------------------------------------------------ proposer --

def _ld(n: int) -> int:
    """Row stride of a proposal score matrix: n rounded up to 8 scores, so
    that every row starts 16-byte aligned for the kernel's vector stores."""
    return (n + 7) // 8 * 8


class Proposer:
    """Candidate selection for multi.MultiSearch from the embedding table:
    per query, the top kc sequences by its proposal score, in ascending
    index order.  A query's score of a sequence is the largest e_r . e_t
    (FP8 MFMA) over its embedding rows: the whole query and, past ``window``
    residues, windows of ``window`` residues every ``stride`` (a mean-pooled
    embedding of a long query is diluted against a homolog that covers only
    part of it; a window is not).  After ``calibrate``, a query whose
    candidate count does not reach the recall gate (``meets_gate``) is
    selected by the plain prefilter instead."""

    def __init__(self, table: EmbeddingTable, group: int = 64, gate: float = 0.99,
                 window: int = 384, stride: int = 192):
        if group % 16 or not 16 <= group <= 128:
            raise ValueError("group: a multiple of 16 in 16..128 rows")
        self.table = table
        self.group = group
        self.gate = gate
        self.window, self.stride = int(window), int(stride)
        self.use_learned = True
        self.calibration: dict | None = None
        self.k_by_len: list[tuple[int, int, int]] | None = None
        self.enforce_gate = True
        self._scores = None  # [group, ld] scores of one launch's rows
        self._best = None    # [1, ld] a query's maximum over its rows

    # queries of a launch group whose maxima the proposal kernel keeps itself
    MAX_FUSED_QUERIES = 32

    def rows_of(self, queries: list[str]) -> tuple[list[str], np.ndarray]:
        """(embedding row texts, rows per query): each query whole, then,
        when longer than ``window``, its windows of ``window`` residues
        starting every ``stride`` (the last one ending at the query's end)."""
        rows, count = [], np.zeros(len(queries), np.int64)
        for i, q in enumerate(queries):
            rows.append(q)
            count[i] = 1
            if self.window > 0 and len(q) > self.window:
                starts = list(range(0, len(q) - self.window + 1, self.stride))
                if starts[-1] != len(q) - self.window:
                    starts.append(len(q) - self.window)
                rows += [q[a: a + self.window] for a in starts]
                count[i] += len(starts)
        return rows, count

    def _groups(self, queries: list[str]) -> list[tuple]:
        """Launch groups: (query indices, q8 rows [Gp, 256], first row of
        each query in the group, its row count, and for a fused launch the
        row offsets of its queries (int32, one more than queries) or None).  Queries are packed first
        fit by decreasing row count, so that few launches (each a read of the
        whole table) hold all rows; a query's rows stay in one group."""
        tab = self.table
        dev = tab.table.device
        st = stream_of(dev)
        rows, count = self.rows_of(queries)
        if int(count.max(initial=1)) > self.group:
            raise ValueError("a query has more window rows than a group holds")
        e = tab.enc.embed_strings(rows) if rows else None
        first = np.concatenate([[0], np.cumsum(count)[:-1]]).astype(np.int64)
        bins: list[list[int]] = []
        fill: list[int] = []
        for i in sorted(range(len(queries)), key=lambda i: (-int(count[i]), i)):
            for g in range(len(bins)):
                if fill[g] + count[i] <= self.group:
                    bins[g].append(i)
                    fill[g] += int(count[i])
                    break
            else:
                bins.append([i])
                fill.append(int(count[i]))
        groups = []
        for members in bins:
            qidx = np.array(sorted(members), dtype=np.int64)
            cnt = count[qidx].copy()
            loc = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
            src = np.concatenate([np.arange(first[i], first[i] + count[i]) for i in qidx])
            Gp = (int(cnt.sum()) + 15) // 16 * 16
            q = torch.zeros((Gp, DIM), dtype=torch.float32, device=dev)
            q[: src.size] = e[torch.from_numpy(src).to(dev)]
            q8 = torch.empty((Gp, DIM), dtype=torch.uint8, device=dev)
            call("sseqs_enc_to_fp8", ptr(q), q.numel(), FP8_SCALE, ptr(q8), st)
            # a group with a multi-row query that the fused launch holds
            # writes one row per query (row j: the maximum over query j's rows)
            off = None
            if cnt.max() > 1 and qidx.size <= self.MAX_FUSED_QUERIES and Gp <= 64:
                off = np.append(loc, cnt.sum()).astype(np.int32)
            groups.append((qidx, q8, loc, cnt, off))
        return groups

    def _buffers(self, dev) -> int:
        """Allocate the score buffers (group rows of n scores, 25.6 GB for 64
        rows of 200 M) once; returns their row stride."""
        ld = _ld(self.table.n)
        if self._scores is None or self._scores.shape[1] != ld:
            self._scores = torch.empty((self.group, ld), dtype=torch.float16, device=dev)
            self._best = torch.empty((1, ld), dtype=torch.float16, device=dev)
        return ld

    def _propose(self, grp, st) -> None:
        """Scores of one launch group into the score buffer: every row's, or
        (fused groups) every query's maximum over its rows."""
        tab = self.table
        qidx, q8, _, _, off = grp
        ld = self._scores.shape[1]
        if off is None:
            call("sseqs_propose", ptr(tab.table), tab.n, ptr(q8), q8.shape[0],
                 ptr(self._scores), ld, st)
        else:
            call("sseqs_propose_max", ptr(tab.table), tab.n, ptr(q8), q8.shape[0],
                 off.ctypes.data, int(qidx.size), ptr(self._scores), ld, st)

    def _query_row(self, grp, j: int, out: int | None, st) -> int:
        """Address of query j's score row of the group after ``_propose``:
        its row in the score buffer, or the maximum over its rows written to
        ``out`` (default: the scratch row)."""
        _, _, first, count, off = grp
        ld = self._scores.shape[1]
        if off is not None:
            src, nrows = self._scores.data_ptr() + 2 * j * ld, 1
        else:
            src, nrows = self._scores.data_ptr() + 2 * int(first[j]) * ld, int(count[j])
        if nrows == 1 and out is None:
            return src
        dst = out if out is not None else self._best.data_ptr()
        call("sseqs_rows_max", src, ld, nrows, self.table.n, dst, st)
        return dst

    def scores_for(self, queries: list[str]) -> torch.Tensor:
        """Proposal scores (fp16 [len(queries), n], a view of a padded
        matrix) of the queries: each one's maximum over its rows."""
        tab = self.table
        dev = tab.table.device
        st = stream_of(dev)
        ld = self._buffers(dev)
        out = torch.empty((len(queries), ld), dtype=torch.float16, device=dev)
        for grp in self._groups(queries):
            self._propose(grp, st)
            for j, i in enumerate(grp[0]):
                self._query_row(grp, j, out[int(i)].data_ptr(), st)
        return out[:, : tab.n]

    # query-length classes of the calibrated operating point
    LEN_EDGES = (0, 100, 200, 350, 500, 700, 2049)

    def calibrate(self, queries: list[str], top_prefilter: int, k: int = 4096, log=None,
                  margin: float = 1.25) -> dict:
        """Recall@k of the proposal against exhaustive SW of the whole
        database for sample queries: per query, the recall among the top
        ``top_prefilter`` proposals and the candidate count that reaches the
        gate.  ``use_learned`` becomes whether every sample query reaches
        the gate at ``top_prefilter``.  The calibrated operating point (``top_prefilter_for``)
        gives a query of each length class ``margin`` times the largest
        count that class's sample queries needed (every sample query's when
        the class has none), at least k.  A query whose candidate count
        meets neither is selected by the plain prefilter (``meets_gate``)."""
        db = self.table.db
        per = []
        for q in queries:
            kth, allsc = exhaustive_top(db, q, k)
            curve, kg = recall_curve(self.scores_for([q])[0], allsc >= kth, k,
                                     (top_prefilter,), self.gate)
            per.append({"q_len": len(q), "kth_exhaustive_score": kth,
                        "recall": curve.get(top_prefilter, 1.0), "top_prefilter_for_gate": kg})
            del allsc
            if log:
                log(f"[calibrate] {json.dumps(per[-1])}")
        rec = [p["recall"] for p in per]
        self.use_learned = bool(per) and min(rec) >= self.gate
        kg = np.array([p["top_prefilter_for_gate"] for p in per], np.float64)
        ql = np.array([p["q_len"] for p in per])
        edges = self.LEN_EDGES
        self.k_by_len = []
        for lo, hi in zip(edges, edges[1:]):
            inside = kg[(ql >= lo) & (ql < hi)]
            top = inside.max() if inside.size else (kg.max() if kg.size else db.n)
            self.k_by_len.append((lo, hi, int(min(db.n, max(k, math.ceil(margin * top))))))
        self.calibration = {"queries": len(per), "top_prefilter": top_prefilter, "k": k,
                            "gate": self.gate, "recall_min": min(rec, default=None),
                            "recall_mean": float(np.mean(rec)) if rec else None,
                            "use_learned": self.use_learned, "k_by_len": self.k_by_len,
                            "per_query": per}
        return self.calibration

    def top_prefilter_for(self, queries: list[str]) -> np.ndarray:
        """Candidate counts of the calibrated operating point (int64, one
        per query); ``calibrate`` first."""
        if self.k_by_len is None:
            raise RuntimeError("calibrate() first")
        out = np.empty(len(queries), np.int64)
        for i, q in enumerate(queries):
            L = len(q)
            out[i] = next((K for lo, hi, K in self.k_by_len if lo <= L < hi),
                          self.k_by_len[-1][2])
        return out

    def prepare(self, ms) -> None:
        """Set-up of a sweep (before any capture): the fp8 embedding rows of
        its queries in launch groups, and the score buffers."""
        t0 = time.perf_counter()
        self._buffers(ms.dev)
        ms.learned_groups = self._groups(ms.texts)
        torch.cuda.synchronize(ms.dev)
        ms.query_embed_s = time.perf_counter() - t0

    def meets_gate(self, texts: list[str], counts) -> np.ndarray:
        """Per query, whether the learned proposal reaches the recall gate
        with ``counts`` candidates, as calibrated: the count reaches the
        query's calibrated operating point, or the gate held for every
        calibration query at a count no larger.  Before any calibration (or
        with ``enforce_gate`` off) every query takes the learned proposal."""
        counts = np.asarray(counts, np.int64).reshape(-1)
        if not self.enforce_gate or self.calibration is None:
            return np.ones(len(texts), bool)
        ok = np.zeros(len(texts), bool)
        if self.k_by_len is not None:
            ok |= counts >= self.top_prefilter_for(texts)
        if self.use_learned:
            ok |= counts >= self.calibration["top_prefilter"]
        return ok

    def select(self, ms) -> None:
        """Candidates of every query of a sweep: the learned proposal's top
        kc_q[i] where that meets the gate, the plain prefilter's elsewhere."""
        ok = self.meets_gate(ms.texts, ms.kc_q)
        for i in np.nonzero(~ok)[0]:
            ms.plain_select(int(i))
        if not ok.any():
            return
        tab = self.table
        st = stream_of(ms.dev)
        if getattr(ms, "learned_groups", None) is None:
            self.prepare(ms)
        for grp in ms.learned_groups:
            qidx = grp[0]
            if not ok[qidx].any():
                continue
            self._propose(grp, st)
            for j, i in enumerate(qidx):
                i = int(i)
                if ok[i]:
                    src = self._query_row(grp, j, None, st)
                    call("sseqs_topk", src, _lib.F16, tab.n, int(ms.kc_q[i]), 0,
                         ms.cand.data_ptr() + 8 * int(ms.c_off_h[i]), ptr(ms.sel_ws),
                         ms.sel_ws.numel(), st)


def recall_curve(scores: torch.Tensor, good: torch.Tensor, k: int, grid=(),
                 gate: float = 0.99) -> tuple[dict, int]:
    """Recall@k of a candidate ranking: ``good`` marks the sequences whose
    exhaustive SW score reaches the k-th best (ties count as found), the
    candidates are the top K by ``scores`` (stable, index order among equal
    scores).  Returns ({K: recall} for K in grid, the smallest K whose
    recall reaches ``gate``)."""
    order = torch.sort(scores.float(), descending=True, stable=True).indices
    cum = torch.cumsum(good[order].to(torch.int32), 0)
    del order
    curve = {int(K): round(min(k, int(cum[K - 1].item())) / k, 4) for K in grid
             if 0 < K <= cum.numel()}
    need = -(-int(round(gate * 10000)) * k // 10000)
    kg = int(torch.searchsorted(cum, torch.tensor([need], dtype=cum.dtype,
                                                  device=cum.device)).item()) + 1
    return curve, min(kg, cum.numel())


def exhaustive_top(db, query: str, k: int, chunk: int = 25_000_000) -> tuple[int, torch.Tensor]:
    """k-th best affine SW score of the query over the whole database, and
    every score (int32, index order)."""
    from .query import prepare
    from .search import sw_affine_ids
    q = prepare(query, device=db.device)
    allsc = torch.empty(db.n, dtype=torch.int32, device=db.device)
    for b in range(0, db.n, chunk):
        ids = torch.arange(b, min(db.n, b + chunk), device=db.device)
        allsc[b: b + ids.numel()] = sw_affine_ids(db, q, ids)
    kth = int(allsc[topk_indices(allsc, k, sorted=False)].min().item())
    return kth, allsc
