"""Device search pipelines over an HBM-resident database.

Two cascades share the kernels:

* ``global_search`` -- the whole-database cascade: prefilter every sequence
  (correlation size N = nextpow2(max(Lq, L)) per length range), keep the
  global top ``top_prefilter``, score them with affine Smith-Waterman (the
  reference's stripe rule, open 11 / extend 1), keep the top ``top_hits``.
  No host synchronisation happens until the hit ids are read back.
* ``batch_search`` -- the reference's own per-batch cascade of
  sseqs/msa.py:196-370 (bs-byte batches of the packed stream, per-batch N,
  top 1/top_fft by prefilter, linear SW -1 on those, top 1/top_sw, affine
  SW on segment[1:]+'@', top top_sw_affine, normalised by
  sqrt(Lq * batch maxlen)).

Ties are resolved by the lower index everywhere (torch.topk's order is
unspecified in the reference).
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, ops
from ._lib import call, ptr, stream_of
from .db import SeqDB
from .query import PreparedQuery


# ------------------------------------------------------------------ global --

@dataclass
class GlobalHits:
    ids: torch.Tensor          # int64 index-order ids of the hits, best first
    scores: torch.Tensor       # int32 affine scores, best first
    prefilter: torch.Tensor    # float32 prefilter score of every sequence
    candidates: torch.Tensor   # int64 ids passed to SW (ascending)
    cand_scores: torch.Tensor  # int32 affine score of every candidate
    timings: dict = field(default_factory=dict)


def prefilter_scores(db: SeqDB, q: PreparedQuery, out: torch.Tensor | None = None,
                     events: list | None = None) -> torch.Tensor:
    """Prefilter score of every database sequence (float32, index order).
    ``events``: optional list of len(length ranges) + 1 torch events, recorded
    on the launch stream before the first and after every range's launch."""
    n = db.n
    if out is None:
        out = torch.empty(n, dtype=torch.float32, device=db.device)
    st = stream_of(db.device)
    if events is not None:
        events[0].record()
    for i, (N, b, e) in enumerate(db.length_ranges(q.length)):
        call("sseqs_db_prefilter", ptr(db.packed), ptr(db.seq_off), ptr(db.seq_len), b, e,
             ptr(q.qrev), q.length, ptr(q.lut_a), ptr(q.lut_b_flat), 21, q.k, N, N,
             out.data_ptr() + 4 * b, st)
        if events is not None:
            events[i + 1].record()
    return out


def prefilter_kernel_name(q_len: int, k: int, N: int, maxlen: int | None = None) -> str:
    """Kernel that sseqs_db_prefilter launches for one length range."""
    buf = ctypes.create_string_buffer(64)
    call("sseqs_db_prefilter_kernel", int(q_len), int(k), int(N),
         int(N if maxlen is None else maxlen), buf, 64)
    return buf.value.decode()


def sw_affine_ids(db: SeqDB, q: PreparedQuery, ids: torch.Tensor, gap_open: int = 11,
                  gap_extend: int = 1) -> torch.Tensor:
    """Affine SW score of database sequences ``ids`` (reference stripe rule)."""
    ids = ids.to(torch.int64).contiguous()
    out = torch.empty(ids.numel(), dtype=torch.int32, device=db.device)
    if ids.numel():
        call("sseqs_db_sw_affine", ptr(db.packed), ptr(db.seq_off), ptr(db.seq_len), ptr(ids),
             ids.numel(), ptr(q.codes), int(q.codes.numel()), int(gap_open), int(gap_extend),
             ptr(out), stream_of(db.device))
    return out


def sw_linear_ids(db: SeqDB, q: PreparedQuery, ids: torch.Tensor, gap: int = -4) -> torch.Tensor:
    """Linear-gap SW score (``gap`` per gap position, reference stripe rule)
    of database sequences ``ids``."""
    ids = ids.to(torch.int64).contiguous()
    out = torch.empty(ids.numel(), dtype=torch.int32, device=db.device)
    if ids.numel():
        call("sseqs_db_sw_linear", ptr(db.packed), ptr(db.seq_off), ptr(db.seq_len), ptr(ids),
             ids.numel(), ptr(q.codes), int(q.codes.numel()), int(gap), ptr(out),
             stream_of(db.device))
    return out


def global_search(db: SeqDB, q: PreparedQuery, top_prefilter: int = 100_000,
                  top_hits: int = 4096, gap_open: int = 11, gap_extend: int = 1,
                  scores_buf: torch.Tensor | None = None,
                  linear_gap: int | None = None, events: list | None = None) -> GlobalHits:
    """The whole-database cascade.  With ``linear_gap`` set, the candidates
    are scored with linear-gap SW (that penalty per gap position) instead of
    affine SW."""
    if q.codes.numel() == 0:
        raise ValueError("query has no residues")
    if q.codes.numel() > 2048:
        raise ValueError("queries longer than 2048 residues are not supported")
    pre = prefilter_scores(db, q, scores_buf, events)
    kc = min(int(top_prefilter), db.n)
    cand = ops.topk_indices(pre, kc, sorted=False)
    if linear_gap is None:
        cs = sw_affine_ids(db, q, cand, gap_open, gap_extend)
    else:
        cs = sw_linear_ids(db, q, cand, linear_gap)
    kh = min(int(top_hits), kc)
    top = ops.topk_indices(cs, kh, sorted=True)
    return GlobalHits(cand[top], cs[top], pre, cand, cs)


@dataclass
class DeviceAlignments:
    """Backtrack output of ``n`` hits, all on the device."""

    out_q: torch.Tensor   # uint8 (n, A), right-aligned aligned query rows
    out_t: torch.Tensor   # uint8 (n, A), right-aligned aligned target rows
    info: torch.Tensor    # int32 (n, 6): len, q_start, q_end, t_start, t_end, score
    width: int            # A


class BacktrackWorkspace:
    """Reusable device buffers for backtracking up to ``n_max`` hits of at
    most ``width`` residues against queries of up to ``lq_max`` residues."""

    def __init__(self, n_max: int, lq_max: int, width: int, device):
        self.n_max, self.lq_max, self.width = n_max, lq_max, width
        self.A = lq_max + width + 1
        dev = torch.device(device)
        self.rows = torch.zeros((n_max, width), dtype=torch.uint8, device=dev)
        self.lens = torch.zeros(n_max, dtype=torch.int32, device=dev)
        self.t_off = torch.arange(n_max, dtype=torch.int64, device=dev) * width
        self.out_q = torch.zeros((n_max, self.A), dtype=torch.uint8, device=dev)
        self.out_t = torch.zeros((n_max, self.A), dtype=torch.uint8, device=dev)
        self.info = torch.zeros((n_max, 6), dtype=torch.int32, device=dev)
        total = lq_max * n_max * width
        nbytes = _lib.load().sseqs_backtrack_ws_bytes(lq_max, total, n_max)
        self.ws = _lib.workspace(nbytes, dev)


def backtrack_hits(db: SeqDB, q: PreparedQuery, ids: torch.Tensor,
                   ws: BacktrackWorkspace, gap_open: int = 11,
                   gap_extend: int = 1) -> DeviceAlignments:
    """Gather hits ``ids`` (truncated to ``ws.width`` residues, as the
    reference truncates a3m hits to 500) and align them to the query on the
    device, without a host round trip.  The targets sit at a fixed stride of
    ``width`` so no host-side offsets are needed."""
    n = int(ids.numel())
    Lq = len(q.text)
    if n > ws.n_max or Lq > ws.lq_max:
        raise ValueError("backtrack workspace too small")
    if n == 0:
        return DeviceAlignments(ws.out_q[:0], ws.out_t[:0], ws.info[:0], ws.A)
    W = ws.width
    dev = db.device
    st = stream_of(dev)
    call("sseqs_db_gather", ptr(db.packed), ptr(db.seq_off), ptr(db.seq_len),
         ptr(ids.to(torch.int64).contiguous()), n, W, ptr(ws.rows), ptr(ws.lens), st)
    cell_off = ws.t_off[:n] * Lq
    total = Lq * n * W
    call("sseqs_backtrack", ptr(q.qbytes), Lq, ptr(ws.rows), ptr(ws.t_off),
         ptr(ws.lens), ptr(cell_off), total, n, int(gap_open), int(gap_extend), ws.A,
         ptr(ws.out_q), ptr(ws.out_t), ptr(ws.info), ptr(ws.ws), ws.ws.numel(), st)
    return DeviceAlignments(ws.out_q[:n], ws.out_t[:n], ws.info[:n], ws.A)


def alignments_to_dicts(al: DeviceAlignments) -> list[dict]:
    """Host dicts in the format of sseqs/sw.py:207-217."""
    qh, th_, ih = al.out_q.cpu().numpy(), al.out_t.cpu().numpy(), al.info.cpu().numpy()
    A = al.width
    out = []
    for b in range(ih.shape[0]):
        n = int(ih[b, 0])
        out.append({
            "q_aligned": qh[b, A - n:].tobytes().decode("latin-1") if n else "",
            "t_aligned": th_[b, A - n:].tobytes().decode("latin-1") if n else "",
            "q_start_orig_0based": int(ih[b, 1]),
            "q_end_orig_0based": int(ih[b, 2]),
            "t_start_orig_0based": int(ih[b, 3]),
            "t_end_orig_0based": int(ih[b, 4]),
            "score": float(ih[b, 5]),
        })
    return out


# ------------------------------------------------------------------- batch --

@dataclass
class BatchHit:
    chunk: int
    batch: int
    norm: float         # affine / sqrt(Lq * batch maxlen), float32
    affine: int
    linear: int
    fft: float          # the reference logs scores[position] here (sic)
    seq: str


def _nextpow2(x: int) -> int:
    return 1 << (x - 1).bit_length()


def batch_search(db: SeqDB, query: str, q: PreparedQuery, bs: int = 50_000_000,
                 top_fft: int = 200, top_sw: int = 10, top_sw_affine: int = 2,
                 dlength: float = 1e6, chunk_bytes: int = 1_000_000_000,
                 timings: dict | None = None, progress=None,
                 hit_width: int | None = 500) -> list[BatchHit]:
    """The reference cascade of sseqs/msa.py:196-370 on the resident stream.
    Hits keep their first ``hit_width`` residues (500 in the reference,
    sseqs/msa.py:183,367; None keeps them whole)."""
    dev = db.device
    hits: list[BatchHit] = []
    Lq = len(query)
    nbytes = db.nbytes
    t = timings if timings is not None else {}
    scratch = torch.zeros(1, dtype=torch.float16, device=dev)
    for lo in range(0, nbytes, bs):
        chunk, j = lo // chunk_bytes, (lo % chunk_bytes) // bs
        nb = min(bs, nbytes - lo)
        n_sym = min(nb * 8 // 5, db.n_syms - (lo * 8) // 5)
        if n_sym <= 0:
            continue
        t0 = time.time()
        ascii = ops.decompress(db.packed, n_sym, lo)
        t["t_dec"] = t.get("t_dec", 0.0) + time.time() - t0
        t0 = time.time()
        D = torch.nonzero(ascii == 64).flatten()                       # msa.py:227
        starts = D + 1
        starts = starts[starts < ascii.numel()]                        # msa.py:238
        offsets = torch.cat([starts, torch.tensor([ascii.numel()], device=dev)])
        lengths = offsets - torch.cat([torch.zeros(1, dtype=offsets.dtype, device=dev),
                                       offsets[:-1]])
        lengths[:-1] -= 1
        if int(ascii[-1].item()) == 64:
            lengths[-1] -= 1
        max_seq_length = float(lengths.max().item())
        t["t_prep1"] = t.get("t_prep1", 0.0) + time.time() - t0
        if progress is not None:
            progress(lo, nb, max_seq_length)
        if abs(Lq - max_seq_length) > dlength:                         # msa.py:252
            continue
        lengths = lengths[lengths != 0]
        if lengths.numel() == 0:
            continue
        t0 = time.time()
        maxlen = int(lengths.max().item())
        n = _nextpow2(max(Lq, maxlen))
        lengths32 = lengths.to(torch.int32)
        fft_off = torch.zeros(lengths.numel() + 1, dtype=torch.int64, device=dev)
        fft_off[1:] = torch.cumsum(lengths + 1, 0)
        scores = torch.zeros(lengths.numel(), dtype=torch.float16, device=dev)
        ops.warp_corr(scores, q.qrev, q.lut_a, q.lut_b,
                      ascii, fft_off[:-1].to(torch.int32), lengths32, q.k, Lq, maxlen, n,
                      scratch)
        t["t_fft"] = t.get("t_fft", 0.0) + time.time() - t0
        t0 = time.time()
        good_idx = ops.topk_indices(scores, scores.numel() // top_fft, sorted=False)
        good_idx = good_idx[1:-1]                                       # msa.py:307
        if good_idx.numel() == 0:
            continue
        gi32 = good_idx.to(torch.int32)
        sw_scores = ops.sw_cuda_blosum62(q.codes, gi32, ascii, D.to(torch.int32),
                                         gap_penalty_param=-1)
        ktop = sw_scores.numel() // top_sw
        if ktop < top_sw_affine:   # the reference raises inside topk here
            continue
        _topk = ops.topk_indices(sw_scores, ktop, sorted=True)
        gap = ops.sw_cuda_affine(q.codes, gi32[_topk], ascii, starts.to(torch.int32), 11, 1)
        __topk = ops.topk_indices(gap, top_sw_affine, sorted=True)
        gap_sel = gap[__topk]
        pos = _topk[__topk]
        norm = gap_sel.float() / torch.tensor(float(Lq * maxlen), device=dev).sqrt()
        seg = good_idx[pos]
        hs, hl = fft_off[:-1][seg], lengths[seg]
        width = int(hl.max().item())
        if hit_width is not None:
            width = min(int(hit_width), width)
        idx = hs[:, None] + torch.arange(width, device=dev)[None, :]
        rows = ascii[idx.clamp_max(ascii.numel() - 1)]
        t["t_sw"] = t.get("t_sw", 0.0) + time.time() - t0
        rows_h = rows.cpu().numpy()
        hl_h = hl.cpu().numpy()
        norm_h = norm.cpu().numpy()
        gap_h = gap_sel.cpu().numpy()
        lin_h = sw_scores[pos].cpu().numpy()
        fft_h = scores[pos].float().cpu().numpy()
        for r in range(rows_h.shape[0]):
            L = min(int(hl_h[r]), width)
            hits.append(BatchHit(chunk, j, float(norm_h[r]), int(gap_h[r]), int(lin_h[r]),
                                 float(fft_h[r]), rows_h[r, :L].tobytes().decode("latin-1")))
    return hits


def rank_hits(hits: list[BatchHit], num_msas: int) -> list[BatchHit]:
    """Final merge (sseqs/msa.py:411-427): normalised score descending, ties in
    collection order, first num_msas."""
    if not hits:
        return []
    s = np.array([h.norm for h in hits], dtype=np.float32)
    order = np.argsort(-s.astype(np.float64), kind="stable")[:num_msas]
    return [hits[i] for i in order]
