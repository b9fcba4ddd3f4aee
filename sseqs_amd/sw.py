"""Public Smith-Waterman entry points (mirror of sseqs/sw.py).

``sw`` scores a list of target strings with the affine kernel
(sseqs/sw.py:10-23); ``sw_affine_backtrack`` returns the local alignment of
each target as the dicts the a3m writer consumes (sseqs/sw.py:27-219), from
one HIP kernel instead of one Triton launch per anti-diagonal.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .ops import sw_cuda_affine
from ._lib import call, ptr, stream_of
from .query import AA_MAP, sw_codes

__all__ = ["sw", "sw_affine_backtrack", "AA_MAP"]


def sw(query: str, targets: list[str], gap_open: int = 11, gap_extend: int = 1) -> torch.Tensor:
    """Affine SW score (int32, on the GPU) of every target.

    Builds ``t0@t1@...@`` and the '@' positions like sseqs/sw.py:11-20.  The
    reference reads position -1 for target 0 (out of bounds); here position
    -1 is defined as -1, so target 0 is scored like every other target.
    """
    dev = torch.device("cuda")
    q = torch.from_numpy(sw_codes(query)).to(dev)
    data = "".join(t + "@" for t in targets).encode("latin-1", errors="replace")
    ascii = torch.from_numpy(np.frombuffer(data, dtype=np.uint8).copy()).to(dev)
    lens = np.array([len(t) for t in targets], dtype=np.int64)
    starts = (np.cumsum(lens + 1) - 1).astype(np.int32)
    P = torch.from_numpy(starts).to(dev)
    good_idx = torch.arange(len(targets), dtype=torch.int32, device=dev)
    return sw_cuda_affine(q, good_idx, ascii, P, gap_open=gap_open, gap_extend=gap_extend)


def sw_affine_backtrack(query: str, targets: list[str], gap_open: int = 11, gap_extend: int = 1,
                        device: str = "cuda", scores: bool = False, rows: bool = False):
    """Local affine alignment of every target (sseqs/sw.py:27-219).

    Returns ``(None, best_scores_int32, aligned)`` -- the reference's first
    element is the full int32 M matrix of the batch, which this kernel never
    materialises -- or only ``best_scores_int32`` when ``scores=True``, or
    with ``rows=True`` the host arrays (q rows, t rows, info) that the
    dicts are made from: target b's alignment is the last info[b, 0] bytes
    of its rows, info[b] = (length, q_start, q_end, t_start, t_end, score).
    Each ``aligned`` entry is a dict with q_aligned, t_aligned,
    q_start_orig_0based, q_end_orig_0based, t_start_orig_0based,
    t_end_orig_0based and score (float).
    """
    if len(targets) == 0:
        raise ValueError("sw_affine_backtrack needs at least one target")
    dev = torch.device(device)
    qb = query.encode("latin-1", errors="replace")
    tb = [t.encode("latin-1", errors="replace") for t in targets]
    Lq, B = len(qb), len(tb)
    t_len = np.array([len(t) for t in tb], dtype=np.int64)
    t_off = np.concatenate([[0], np.cumsum(t_len)[:-1]]).astype(np.int64)
    cell_off = (Lq * t_off).astype(np.int64)
    total_cells = int(Lq * t_len.sum())
    A = Lq + int(t_len.max()) + 1
    blob = b"".join(tb) or b"\0"
    d_q = torch.from_numpy(np.frombuffer(qb or b"\0", dtype=np.uint8).copy()).to(dev)
    d_t = torch.from_numpy(np.frombuffer(blob, dtype=np.uint8).copy()).to(dev)
    d_off = torch.from_numpy(t_off).to(dev)
    d_len = torch.from_numpy(t_len.astype(np.int32)).to(dev)
    d_cell = torch.from_numpy(cell_off).to(dev)
    out_q = torch.zeros((B, A), dtype=torch.uint8, device=dev)
    out_t = torch.zeros((B, A), dtype=torch.uint8, device=dev)
    info = torch.zeros((B, 6), dtype=torch.int32, device=dev)
    lib = _lib.load()
    ws = _lib.workspace(lib.sseqs_backtrack_ws_bytes(Lq, total_cells, B), dev)
    call("sseqs_backtrack", ptr(d_q), Lq, ptr(d_t), ptr(d_off), ptr(d_len), ptr(d_cell),
         total_cells, B, int(gap_open), int(gap_extend), A, ptr(out_q), ptr(out_t), ptr(info),
         ptr(ws), ws.numel(), stream_of(dev))
    best = info[:, 5].clone()
    if scores:
        return best
    if rows:
        return out_q.cpu().numpy(), out_t.cpu().numpy(), info.cpu().numpy()
    qh, th_, ih = out_q.cpu().numpy(), out_t.cpu().numpy(), info.cpu().numpy()
    aligned = []
    for b in range(B):
        n = int(ih[b, 0])
        if n:
            qa = qh[b, A - n:].tobytes().decode("latin-1")
            ta = th_[b, A - n:].tobytes().decode("latin-1")
        else:
            qa = ta = ""
        aligned.append({
            "q_aligned": qa,
            "t_aligned": ta,
            "q_start_orig_0based": int(ih[b, 1]),
            "q_end_orig_0based": int(ih[b, 2]),
            "t_start_orig_0based": int(ih[b, 3]),
            "t_end_orig_0based": int(ih[b, 4]),
            "score": float(ih[b, 5]),
        })
    return None, best, aligned
