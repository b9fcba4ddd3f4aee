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
        self.cell_off = torch.zeros(n_max, dtype=torch.int64, device=dev)
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
    ids = ids.to(torch.int64).contiguous()
    call("sseqs_db_gather", ptr(db.packed), ptr(db.seq_off), ptr(db.seq_len), ptr(ids), n, W,
         ptr(ws.rows), ptr(ws.lens), st)
    cell_off = torch.mul(ws.t_off[:n], Lq, out=ws.cell_off[:n])  # no allocation
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


# Correlation sizes of the per-batch prefilter.  The reference's size is
# nextpow2(max(Lq, batch maxlen)) (sseqs/msa.py:261); its kernels only run
# for 16..1024 and leave the scores of other batches at zero
# (sseqs/filter.cu:313-390).  Here the size is clamped to 16..2048 and
# longer segments are scored on their first 2048 residues.
BATCH_N_MIN, BATCH_N_MAX = 16, 2048


def batch_corr_size(Lq: int, maxlen: int) -> int:
    return min(max(_nextpow2(max(Lq, maxlen, 1)), BATCH_N_MIN), BATCH_N_MAX)


def batch_search(db: SeqDB, query: str, q: PreparedQuery, bs: int = 50_000_000,
                 top_fft: int = 200, top_sw: int = 10, top_sw_affine: int = 2,
                 dlength: float = 1e6, chunk_bytes: int = 1_000_000_000,
                 timings: dict | None = None, progress=None,
                 hit_width: int | None = 500, on_scores=None,
                 resident: bool = True) -> list[BatchHit]:
    """The reference cascade of sseqs/msa.py:196-370 over the resident stream.

    Hits keep their first ``hit_width`` residues (500 in the reference,
    sseqs/msa.py:183,367; None keeps them whole).  ``on_scores(b, scores)``,
    if given, receives the fp16 prefilter scores of every batch b that
    reaches the prefilter.

    With ``resident`` (and bs a multiple of 5, so batches start on whole
    8-symbol groups) the segments of every batch come from the resident
    index and all batches run together: one prefilter launch per run of
    batches with the same correlation size, then segmented selects and SW
    launches over all batches at once, and one read-back of the hits.
    Batches whose reference offsets are shifted by empty segments
    (sseqs_amd.batches) and every batch of other configurations run the
    per-batch ASCII path of the reference.
    """
    t = timings if timings is not None else {}
    Lq_sw = int(q.codes.numel())
    if (resident and bs > 0 and bs % 5 == 0 and 1 <= Lq_sw <= 2048 and top_fft > 0
            and top_sw > 0 and 1 <= top_sw_affine <= 4096):
        t0 = time.time()
        plan = db.batch_plan(bs)
        t["t_plan"] = t.get("t_plan", 0.0) + time.time() - t0
        return _resident_batches(db, query, q, plan, top_fft, top_sw, top_sw_affine, dlength,
                                 chunk_bytes, t, progress, hit_width, on_scores)
    hits: list[BatchHit] = []
    for lo in range(0, db.nbytes, bs):
        n_sym = min(min(bs, db.nbytes - lo) * 8 // 5, db.n_syms - (lo * 8) // 5)
        if n_sym <= 0:
            continue
        hits += _ascii_batch(db, query, q, lo, bs, top_fft, top_sw, top_sw_affine, dlength,
                             chunk_bytes, t, progress, hit_width, on_scores)
    return hits


def _ascii_batch(db: SeqDB, query: str, q: PreparedQuery, lo: int, bs: int, top_fft: int,
                 top_sw: int, top_sw_affine: int, dlength: float, chunk_bytes: int, t: dict,
                 progress, hit_width: int | None, on_scores) -> list[BatchHit]:
    """One batch of the reference cascade from its ASCII unpacking
    (sseqs/msa.py:211-370)."""
    dev = db.device
    Lq = len(query)
    chunk, j = lo // chunk_bytes, (lo % chunk_bytes) // bs
    nb = min(bs, db.nbytes - lo)
    n_sym = min(nb * 8 // 5, db.n_syms - (lo * 8) // 5)
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
        return []
    lengths = lengths[lengths != 0]
    if lengths.numel() == 0:
        return []
    t0 = time.time()
    maxlen = int(lengths.max().item())
    n = batch_corr_size(Lq, maxlen)
    lengths32 = lengths.to(torch.int32)
    fft_off = torch.zeros(lengths.numel() + 1, dtype=torch.int64, device=dev)
    fft_off[1:] = torch.cumsum(lengths + 1, 0)
    scores = torch.zeros(lengths.numel(), dtype=torch.float16, device=dev)
    ops.warp_corr(scores, q.qrev, q.lut_a, q.lut_b,
                  ascii, fft_off[:-1].to(torch.int32), lengths32, q.k, Lq, maxlen, n,
                  torch.zeros(1, dtype=torch.float16, device=dev))
    if on_scores is not None:
        on_scores(lo // bs, scores)
    t["t_fft"] = t.get("t_fft", 0.0) + time.time() - t0
    t0 = time.time()
    good_idx = ops.topk_indices(scores, scores.numel() // top_fft, sorted=False)
    good_idx = good_idx[1:-1]                                       # msa.py:307
    if good_idx.numel() == 0:
        return []
    gi32 = good_idx.to(torch.int32)
    sw_scores = ops.sw_cuda_blosum62(q.codes, gi32, ascii, D.to(torch.int32),
                                     gap_penalty_param=-1)
    ktop = sw_scores.numel() // top_sw
    if ktop < top_sw_affine:   # the reference raises inside topk here
        return []
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
    hits = []
    for r in range(rows_h.shape[0]):
        L = min(int(hl_h[r]), width)
        hits.append(BatchHit(chunk, j, float(norm_h[r]), int(gap_h[r]), int(lin_h[r]),
                             float(fft_h[r]), rows_h[r, :L].tobytes().decode("latin-1")))
    return hits


def _resident_batches(db: SeqDB, query: str, q: PreparedQuery, plan, top_fft: int,
                      top_sw: int, top_sw_affine: int, dlength: float, chunk_bytes: int,
                      t: dict, progress, hit_width: int | None, on_scores) -> list[BatchHit]:
    """All batches of one sweep at once from the batch plan; the same
    cascade as _ascii_batch, batch by batch, on the plan's segments."""
    dev = db.device
    st = stream_of(dev)
    Lq = len(query)
    Lq_sw = int(q.codes.numel())
    bs = plan.bs
    maxl = plan.maxlen
    if progress is not None:
        for b in range(plan.n_batches):
            progress(int(plan.lo[b]), int(min(bs, db.nbytes - plan.lo[b])), float(maxl[b]))
    active = (plan.nseg > 0) & ~(np.abs(Lq - maxl.astype(np.float64)) > dlength)
    fast = active & (plan.regular | plan.shifted)
    slow = np.nonzero(active & ~fast)[0]
    Ns = np.array([batch_corr_size(Lq, int(m)) for m in maxl], dtype=np.int64)
    fb = np.nonzero(fast)[0]

    # prefilter: one launch per run of consecutive batches with one N and one
    # kind; a shifted batch's spans may hold '@', which the generic kernels
    # score like the reference's ASCII operator
    t0 = time.time()
    sc32 = torch.empty(plan.seg_off.numel(), dtype=torch.float32, device=dev)
    # each batch's segments in length order when the plan has one (same
    # per-batch ranges), scores put back in plan order after the launches
    pf_off, pf_len = ((plan.pf_off, plan.pf_len) if plan.pf_perm is not None
                      else (plan.seg_off, plan.seg_len))
    pf_out = torch.empty_like(sc32) if plan.pf_perm is not None else sc32
    i = 0
    while i < fb.size:
        e = i
        while (e + 1 < fb.size and fb[e + 1] == fb[e] + 1 and Ns[fb[e + 1]] == Ns[fb[i]]
               and plan.shifted[fb[e + 1]] == plan.shifted[fb[i]]):
            e += 1
        b0, b1 = fb[i], fb[e]
        s_begin = int(plan.seg_base[b0])
        s_end = int(plan.seg_base[b1] + plan.nseg[b1])
        call("sseqs_db_prefilter_spans" if plan.shifted[b0] else "sseqs_db_prefilter",
             ptr(db.packed), ptr(pf_off), ptr(pf_len), s_begin,
             s_end, ptr(q.qrev), Lq, ptr(q.lut_a), ptr(q.lut_b_flat), 21, q.k, int(Ns[b0]),
             min(int(maxl[b0:b1 + 1].max()), BATCH_N_MAX), pf_out.data_ptr() + 4 * s_begin, st)
        i = e + 1
    if plan.pf_perm is not None:
        sc32[plan.pf_perm] = pf_out
        del pf_out
    sc16 = sc32.half()
    del sc32
    if on_scores is not None:
        for b in fb:
            on_scores(int(b), sc16[plan.seg_base[b]: plan.seg_base[b] + plan.nseg[b]])
    t["t_fft"] = t.get("t_fft", 0.0) + time.time() - t0

    # top 1/top_fft of each batch in index order, without its first and
    # last (sseqs/msa.py:297,307), as segment ids; linear SW -1 of those
    t0 = time.time()
    k1 = plan.nseg[fb] // top_fft
    cand1, base1 = ops.topk_segments(sc16, plan.seg_base[fb], plan.nseg[fb], k1, drop_front=1,
                                     drop_back=1, add_begin=True)
    cnt1 = np.maximum(k1 - 2, 0)
    sw1 = torch.empty(cand1.numel(), dtype=torch.int32, device=dev)
    if cand1.numel():
        # a wave scores four targets to the length of the longest: in a file
        # not in length order (the plan then has pf_perm) score them by length
        order1 = (torch.sort(plan.sw_len[cand1], stable=True).indices
                  if plan.pf_perm is not None else None)
        c_sw = cand1 if order1 is None else cand1[order1]
        out1 = sw1 if order1 is None else torch.empty_like(sw1)
        call("sseqs_db_sw_linear", ptr(db.packed), ptr(plan.sw_off), ptr(plan.sw_len),
             ptr(c_sw), c_sw.numel(), ptr(q.codes), Lq_sw, -1, ptr(out1), st)
        if order1 is not None:
            sw1[order1] = out1
            del out1, c_sw
    # top 1/top_sw by linear score (sorted), affine SW 11/1 of segment[1:]+'@'
    k2 = cnt1 // top_sw
    ok = (cnt1 > 0) & (k2 >= top_sw_affine)
    b2, base1o, k2o = fb[ok], base1[ok], k2[ok]
    if k2o.size and int(k2o.max()) > 4096:
        parts = [ops.topk_indices(sw1[int(o): int(o) + int(c)], int(k), sorted=True)
                 for o, c, k in zip(base1o, cnt1[ok], k2o)]
        pos2 = torch.cat(parts) if parts else torch.empty(0, dtype=torch.int64, device=dev)
        base2 = np.zeros(k2o.size, dtype=np.int64)
        base2[1:] = np.cumsum(k2o)[:-1]
    else:
        pos2, base2 = ops.topk_segments(sw1, base1o, cnt1[ok], k2o, sorted=True)
    rep = torch.from_numpy(np.repeat(base1o, k2o)).to(dev)
    ids2 = cand1[rep + pos2]
    n2 = int(ids2.numel())
    sw2 = torch.empty(n2, dtype=torch.int32, device=dev)
    if n2:
        aff_off = (plan.sw_off[ids2] + 1).contiguous()
        aff_len = plan.sw_len[ids2].contiguous()
        ident = torch.arange(n2, device=dev)
        call("sseqs_db_sw_affine", ptr(db.packed), ptr(aff_off), ptr(aff_len), ptr(ident), n2,
             ptr(q.codes), Lq_sw, 11, 1, ptr(sw2), st)
    # top top_sw_affine by affine score (sorted); the hits' fields
    k3 = np.full(b2.size, top_sw_affine, dtype=np.int64)
    pos3, _ = ops.topk_segments(sw2, base2, k2o, k3, sorted=True)
    g3 = torch.from_numpy(np.repeat(base2, k3)).to(dev) + pos3
    p2 = pos2[g3]                                    # index into the batch's stage-1 list
    at1 = torch.from_numpy(np.repeat(base1o, k3)).to(dev) + p2
    gseg = cand1[at1]
    gap = sw2[g3]
    lin = sw1[at1]
    fft = sc16[torch.from_numpy(np.repeat(plan.seg_base[b2], k3)).to(dev) + p2].float()  # (sic)
    den = torch.from_numpy(np.repeat(Lq * maxl[b2], k3).astype(np.float64)).to(
        device=dev, dtype=torch.float32).sqrt()
    norm = gap.float() / den
    width = int(hit_width) if hit_width is not None else (
        int(plan.seg_len[gseg].max().item()) if gseg.numel() else 1)
    rows = torch.zeros((gseg.numel(), max(width, 1)), dtype=torch.uint8, device=dev)
    hl = torch.empty(gseg.numel(), dtype=torch.int32, device=dev)
    if gseg.numel():
        call("sseqs_db_gather", ptr(db.packed), ptr(plan.seg_off), ptr(plan.seg_len), ptr(gseg),
             gseg.numel(), max(width, 1), ptr(rows), ptr(hl), st)
    rows_h, hl_h = rows.cpu().numpy(), hl.cpu().numpy()
    norm_h, gap_h = norm.cpu().numpy(), gap.cpu().numpy()
    lin_h, fft_h = lin.cpu().numpy(), fft.cpu().numpy()
    t["t_sw"] = t.get("t_sw", 0.0) + time.time() - t0

    per: dict[int, list[BatchHit]] = {}
    r = 0
    for b in b2:
        lo = int(plan.lo[b])
        chunk, j = lo // chunk_bytes, (lo % chunk_bytes) // bs
        out = []
        for _ in range(top_sw_affine):
            out.append(BatchHit(chunk, j, float(norm_h[r]), int(gap_h[r]), int(lin_h[r]),
                                float(fft_h[r]), rows_h[r, : hl_h[r]].tobytes().decode("latin-1")))
            r += 1
        per[int(b)] = out
    for b in slow:
        per[int(b)] = _ascii_batch(db, query, q, int(plan.lo[b]), bs, top_fft, top_sw,
                                   top_sw_affine, dlength, chunk_bytes, t, None, hit_width,
                                   on_scores)
    hits: list[BatchHit] = []
    for b in sorted(per):
        hits += per[b]
    return hits


def rank_hits(hits: list[BatchHit], num_msas: int) -> list[BatchHit]:
    """Final merge (sseqs/msa.py:411-427): the reference's own ordering,
    np.argsort over the float32 normalised scores reversed, so ties come out
    as numpy orders them; first num_msas."""
    if not hits:
        return []
    s = np.array([h.norm for h in hits], dtype=np.float32)
    order = np.argsort(s)[::-1][:num_msas]
    return [hits[i] for i in order]
