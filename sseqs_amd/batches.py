"""Batch plan of the per-batch cascade over the resident index.

The reference unpacks every batch of ``bs`` packed bytes to ASCII and finds
its segments by scanning for '@' (sseqs/msa.py:211-267).  The resident
index already holds every run of residues, so the segments of all batches
follow from it once per (database, bs): a batch window
[b * bs * 8/5, (b + 1) * bs * 8/5) cuts the runs it overlaps, and each
overlap is one segment (the first and last may be partial runs).

A batch is *regular* when its segments are exactly those overlaps in that
order with their reference offsets: the window starts inside a run and
consecutive runs in it are one '@' apart.  Otherwise (a window that starts
on '@', or '@@' inside it) the reference's ``lengths != 0`` filter shifts
the offsets of later segments (sseqs/msa.py:253,266-267) and the batch is
run through the ASCII path, which reproduces that shift.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class BatchPlan:
    bs: int                 # packed bytes per batch
    lo: np.ndarray          # int64 [B] first byte of each batch
    n_sym: np.ndarray       # int64 [B] symbols in each batch window
    seg_base: np.ndarray    # int64 [B] first segment of each batch in seg_off/seg_len
    nseg: np.ndarray        # int64 [B] segments (non-empty) of each batch
    maxlen: np.ndarray      # int64 [B] longest segment (0 for none)
    regular: np.ndarray     # bool [B]
    seg_off: torch.Tensor   # int64 [S] symbol offset of every segment, batch by batch
    seg_len: torch.Tensor   # int32 [S]

    @property
    def n_batches(self) -> int:
        return int(self.lo.size)


def plan_batches(seq_off: torch.Tensor, seq_len: torch.Tensor, n_syms: int, bs: int
                 ) -> BatchPlan:
    """Segments of every batch of ``bs`` bytes (a multiple of 5, so each
    batch starts on an 8-symbol group) of a stream of ``n_syms`` symbols
    whose non-empty runs are (seq_off, seq_len), in any order."""
    if bs <= 0 or bs % 5:
        raise ValueError("plan_batches: bs must be a positive multiple of 5")
    dev = seq_off.device
    nbytes = (int(n_syms) * 5 + 7) // 8
    lo = np.arange(0, nbytes, bs, dtype=np.int64)
    s0 = lo * 8 // 5
    n_sym = np.minimum(np.minimum(bs, nbytes - lo) * 8 // 5, int(n_syms) - s0)
    B = int(lo.size)
    sb = bs * 8 // 5
    off = seq_off.to(torch.int64)
    ln = seq_len.to(torch.int64)
    if off.numel() > 1 and not bool((off[1:] > off[:-1]).all().item()):
        order = torch.argsort(off)
        off, ln = off[order], ln[order]
    end = off + ln
    b_first = off // sb
    cnt = (end - 1) // sb - b_first + 1
    total = int(cnt.sum().item()) if cnt.numel() else 0
    if total == off.numel():  # no run crosses a batch boundary
        p_b, p_off, p_end = b_first, off, end
    else:
        run = torch.repeat_interleave(torch.arange(off.numel(), device=dev), cnt)
        first = torch.cumsum(cnt, 0) - cnt
        p_b = b_first[run] + (torch.arange(total, device=dev) - first[run])
        p_off = torch.maximum(off[run], p_b * sb)
        p_end = torch.minimum(end[run], (p_b + 1) * sb)
        del run, first
    p_len = p_end - p_off
    nseg_t = torch.bincount(p_b, minlength=B) if total else torch.zeros(B, dtype=torch.int64,
                                                                        device=dev)
    base_t = torch.cumsum(nseg_t, 0) - nseg_t
    maxlen_t = torch.zeros(B, dtype=torch.int64, device=dev)
    if total and dev.type == "cuda":
        # longest segment of each batch: a top-1 per batch in one launch
        from . import ops
        nseg_h = nseg_t.cpu().numpy()
        base_h = base_t.cpu().numpy()
        has_h = nseg_h > 0
        p32 = p_len.to(torch.int32)
        top, _ = ops.topk_segments(p32, base_h[has_h], nseg_h[has_h], np.ones(int(has_h.sum())),
                                   add_begin=True)
        maxlen_t[torch.from_numpy(np.nonzero(has_h)[0]).to(dev)] = p32[top].to(torch.int64)
        del p32
    elif total:
        # segmented max by a running max of (batch, length) keys: batches are
        # contiguous and in order, so the key's running max at a batch's last
        # segment holds that batch's longest segment
        run_max = torch.cummax((p_b << 32) + p_len, 0).values
        has_t = nseg_t > 0
        last = (base_t + nseg_t - 1).clamp(min=0)
        bidx = torch.arange(B, device=dev, dtype=torch.int64)
        maxlen_t = torch.where(has_t, run_max[last] - (bidx << 32), maxlen_t)
        del run_max
    # regular: the first segment starts the window, and segments of one
    # batch are one delimiter apart
    has = nseg_t > 0
    first_off = p_off[base_t.clamp(max=max(total - 1, 0))] if total else base_t
    lead_ok = first_off == torch.from_numpy(s0).to(dev)
    if total > 1:
        bad = (p_b[1:] == p_b[:-1]) & (p_off[1:] - p_end[:-1] != 1)
        nbad = torch.bincount(p_b[1:][bad], minlength=B)
    else:
        nbad = torch.zeros(B, dtype=torch.int64, device=dev)
    regular_t = has & lead_ok & (nbad == 0)
    host = torch.stack([nseg_t, base_t, maxlen_t, regular_t.to(torch.int64)]).cpu().numpy()
    return BatchPlan(bs=int(bs), lo=lo, n_sym=n_sym, seg_base=host[1], nseg=host[0],
                     maxlen=host[2], regular=host[3].astype(bool), seg_off=p_off.contiguous(),
                     seg_len=p_len.to(torch.int32).contiguous())
