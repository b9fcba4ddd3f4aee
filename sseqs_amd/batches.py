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
the offsets of later segments (sseqs/msa.py:253,266-267).  For such a
*shifted* batch the plan holds the reference's own two views of its
segments (``reference_views``): the prefilter reads segment g at the
running sum of (length + 1) of the segments before it, and the SW stages
take target g between delimiters g-1 and g (sseqs/sw.cu:67-77), which is a
different run once an empty segment was dropped.
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
    seg_off: torch.Tensor   # int64 [S] prefilter / gather offset of every segment, batch by batch
    seg_len: torch.Tensor   # int32 [S]
    shifted: np.ndarray = None     # bool [B]: not regular, reference views below
    sw_off: torch.Tensor = None    # int64 [S] SW target g of a batch (= seg_off if regular)
    sw_len: torch.Tensor = None    # int32 [S]
    # the prefilter's order when the stream is not in length order: each
    # batch's segments by length (pf_perm[i] = plan index of the i-th), so
    # that a tile of 16 does not run to the length of its longest member;
    # None when the plan order wastes little
    pf_perm: torch.Tensor = None   # int64 [S]
    pf_off: torch.Tensor = None    # int64 [S] seg_off[pf_perm]
    pf_len: torch.Tensor = None    # int32 [S] seg_len[pf_perm]

    @property
    def n_batches(self) -> int:
        return int(self.lo.size)


def reference_views(off: np.ndarray, ln: np.ndarray, W: int):
    """The reference's views of one batch window of W symbols whose runs of
    residues start at ``off`` (ascending, relative to the window) with
    lengths ``ln``; every other symbol of the window is '@'.  Returns
    (pf_off, sw_off, sw_len), all relative to the window:

    - pf_off[g]: offset the prefilter and the hit gather use for segment g,
      the running sum of (length + 1) of the non-empty segments before it
      (sseqs/msa.py:253,266-267,349-350);
    - [sw_off[g], sw_off[g] + sw_len[g]): SW target g, between delimiters
      g-1 and g (sseqs/sw.cu:67-77 with P = delimiter positions, P[-1] =
      -1); length 0 scores -1.  The affine stage reads the same span one
      symbol later (P = delimiters + 1, sseqs/msa.py:331).

    None when the runs do not reproduce the reference's segment lengths
    (the batch then stays on the ASCII path)."""
    off = np.asarray(off, dtype=np.int64)
    ln = np.asarray(ln, dtype=np.int64)
    m = off.size
    gs = np.concatenate([[0], off + ln])   # gaps between runs: the '@'s
    ge = np.concatenate([off, [W]])
    gl = ge - gs
    if m == 0 or (gl < 0).any():
        return None
    tot = int(gl.sum())
    D = np.repeat(gs - (np.cumsum(gl) - gl), gl) + np.arange(tot, dtype=np.int64)
    # the reference's segment lengths (sseqs/msa.py:227-253)
    starts = D + 1
    starts = starts[starts < W]
    offsets = np.concatenate([starts, [W]])
    lengths = offsets - np.concatenate([[0], offsets[:-1]])
    lengths[:-1] -= 1
    if tot and D[-1] == W - 1:
        lengths[-1] -= 1
    lengths = lengths[lengths != 0]
    if lengths.size != m or not np.array_equal(lengths, ln):
        return None
    pf = np.concatenate([[0], np.cumsum(ln + 1)[:-1]])
    g = np.arange(m)
    prev = np.full(m, -1, dtype=np.int64)
    if tot:
        prev[1:] = D[np.minimum(g[1:] - 1, tot - 1)]
    sw = prev + 1
    stop = np.where(g < tot, D[np.minimum(g, max(tot - 1, 0))] if tot else 0, sw)
    return pf, sw, np.maximum(stop - sw, 0)


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
    seg_off, seg_len = p_off.contiguous(), p_len.to(torch.int32).contiguous()
    regular, nseg_h, base_h = host[3].astype(bool), host[0], host[1]
    # shifted batches (a window starting on '@', or '@@' inside it): few, so
    # their reference views are computed on the host, batch by batch
    shifted = np.zeros(B, dtype=bool)
    sw_off, sw_len = seg_off, seg_len
    irr = np.nonzero((nseg_h > 0) & ~regular)[0]
    if irr.size:
        sw_off, sw_len = seg_off.clone(), seg_len.clone()
        for b in irr:
            lo_s, m = int(base_h[b]), int(nseg_h[b])
            s0b = int(s0[b])
            v = reference_views((seg_off[lo_s:lo_s + m] - s0b).cpu().numpy(),
                                seg_len[lo_s:lo_s + m].cpu().numpy(), int(n_sym[b]))
            if v is None:
                continue
            pf, so, sl = v
            seg_off[lo_s:lo_s + m] = torch.from_numpy(s0b + pf).to(dev)
            sw_off[lo_s:lo_s + m] = torch.from_numpy(s0b + so).to(dev)
            sw_len[lo_s:lo_s + m] = torch.from_numpy(sl.astype(np.int32)).to(dev)
            shifted[b] = True
    plan = BatchPlan(bs=int(bs), lo=lo, n_sym=n_sym, seg_base=base_h, nseg=nseg_h,
                     maxlen=host[2], regular=regular, seg_off=seg_off, seg_len=seg_len,
                     shifted=shifted, sw_off=sw_off, sw_len=sw_len)
    if total:
        # prefilter tiles hold 16 consecutive segments and run to the longest;
        # sort each batch by length when that wastes more than 10 %
        key = (p_b << 32) | p_len
        perm = torch.sort(key, stable=True).indices
        if _tile_cost(p_len) > 1.1 * _tile_cost(p_len[perm]):
            plan.pf_perm = perm
            plan.pf_off = seg_off[perm].contiguous()
            plan.pf_len = seg_len[perm].contiguous()
        del key, perm
    return plan


def _tile_cost(lens: torch.Tensor, tile: int = 16) -> float:
    """Sum over consecutive groups of ``tile`` of the group's longest length."""
    n = lens.numel()
    pad = (-n) % tile
    x = torch.cat([lens, lens.new_zeros(pad)]) if pad else lens
    return float(x.view(-1, tile).max(dim=1).values.sum().item())
