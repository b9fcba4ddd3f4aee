"""Python mirror of the reference's three extension modules.

The reference JIT-builds ``xbit``, ``conv`` and ``sw`` (also loaded as
``sseqs_sw_ext``) from CUDA sources (sseqs/msa.py:33-35, sseqs/sw.py:9) and
binds them with pybind11 (sseqs/xbit_bind.cpp:33-36,
sseqs/filter_bind.cpp:67-82, sseqs/sw_bind.cpp:111-132).  The namespaces
below keep those module, function and argument names and defaults; each call
goes to the HIP kernels of libsseqs_amd.so on torch's current stream.

Deviations, all on inputs the reference mishandled: unsupported sizes and
dtypes raise instead of silently skipping the launch or reading out of
bounds; ``xbit`` launches on the current stream instead of the legacy default
stream; prefilter scores are exact fp32 sums rounded once to fp16 instead of
an fp16 FFT that overflows to inf on strong hits.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_of


def _cuda(t: torch.Tensor, dtype: torch.dtype, name: str) -> torch.Tensor:
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    return t.contiguous()


# ------------------------------------------------------------------ xbit --

def compress(input: torch.Tensor) -> torch.Tensor:
    """Pack ASCII symbols 5 bits each (sseqs/xbit_bind.cpp:11-18).

    Like the reference, a host or non-uint8 input is first moved to the GPU
    as uint8.  Returns ``(5n+7)//8`` bytes.
    """
    x = input.contiguous().to(device="cuda" if not input.is_cuda else input.device,
                              dtype=torch.uint8)
    n = x.numel()
    out = torch.zeros((n * 5 + 7) // 8, dtype=torch.uint8, device=x.device)
    if n:
        call("compress_launcher", ptr(x), ptr(out), n, stream_of(x.device))
    return out


def decompress(input: torch.Tensor, n: int, lo: int) -> torch.Tensor:
    """Unpack ``n`` symbols starting at byte ``lo`` (sseqs/xbit_bind.cpp:23-28)."""
    x = input.to(device="cuda" if not input.is_cuda else input.device, dtype=torch.uint8)
    x = x.contiguous().view(-1)
    n, lo = int(n), int(lo)
    if n < 0 or lo < 0:
        raise ValueError("decompress: n and lo must be >= 0")
    out = torch.empty(n, dtype=torch.uint8, device=x.device)
    if n == 0:
        return out
    need = lo + ((n + 7) // 8) * 5  # the kernel reads whole 5-byte groups
    if need > x.numel():
        if lo + (n * 5 + 7) // 8 > x.numel():
            raise ValueError("decompress: input holds fewer than n symbols after lo")
        padded = torch.zeros(need, dtype=torch.uint8, device=x.device)
        padded[: x.numel()] = x
        x = padded
    call("decompress_launcher", ptr(x), ptr(out), n, lo, stream_of(x.device))
    return out


xbit = SimpleNamespace(compress=compress, decompress=decompress)


# ------------------------------------------------------------------ conv --

def _pow2(n: int) -> bool:
    return n > 0 and (n & (n - 1)) == 0


def warp_corr(corr: torch.Tensor, query_reversed_indices: torch.Tensor,
              weight_LUTS_A: torch.Tensor, W: torch.Tensor, flat: torch.Tensor,
              offsets: torch.Tensor, lengths: torch.Tensor, k: int, query_length: int,
              maxlen: int, n: int, query_rfft_tensor: torch.Tensor) -> torch.Tensor:
    """Low-rank correlation prefilter (sseqs/filter_bind.cpp:20-64).

    Writes one fp16 score per segment into ``corr`` in place and returns the
    reference's unused ``(B, 2)`` float32 tensor.  As in the reference, a
    float32 ``corr`` is converted to a temporary fp16 copy which receives the
    scores, and the segment count, vocabulary size and query length are taken
    from ``lengths``, ``W`` and ``query_reversed_indices`` (``query_length``
    is not read).
    """
    B = int(lengths.size(0))
    V = int(W.size(1))
    q_len = int(query_reversed_indices.size(0))
    k, n, maxlen = int(k), int(n), int(maxlen)
    out = torch.empty((B, 2), dtype=torch.float32, device=corr.device)
    corr_half = corr.to(torch.float16) if corr.dtype == torch.float32 else corr
    if corr_half.dtype != torch.float16:
        raise TypeError("corr must be float16 or float32")
    if not _pow2(n) or n < 16 or n > 2048:
        raise ValueError(f"warp_corr: n must be a power of two in 16..2048, got {n}")
    if B == 0:
        return out
    qrev = _cuda(query_reversed_indices, torch.uint8, "query_reversed_indices")
    lutA = _cuda(weight_LUTS_A, torch.float32, "weight_LUTS_A")
    lutB = _cuda(W, torch.float32, "W").view(-1)
    if lutB.numel() <= k * V:  # code 21 reads one entry past rank k-1
        lutB = torch.cat([lutB, lutB.new_zeros(k * V + 1 - lutB.numel())])
    if lutA.numel() < k * V or lutA.dim() != 2:
        raise ValueError("weight_LUTS_A must hold k rows of vocab_size values")
    flat_ = _cuda(flat, torch.uint8, "flat")
    off = _cuda(offsets, torch.int32, "offsets")
    ln = _cuda(lengths, torch.int32, "lengths")
    if not corr_half.is_contiguous() or corr_half.numel() < B:
        raise ValueError("corr must be a contiguous tensor with one entry per segment")
    call("launch_warp_corr", ptr(corr_half), ptr(qrev), ptr(lutA), ptr(lutB), ptr(flat_),
         ptr(off), ptr(ln), ptr(out), B, V, k, q_len, maxlen, n, stream_of(corr.device),
         ptr(query_rfft_tensor))
    return out


conv = SimpleNamespace(warp_corr=warp_corr)


# -------------------------------------------------------------------- sw --

def sw_cuda_blosum62(query_indices_tensor: torch.Tensor, good_idx: torch.Tensor,
                     ascii_tensor: torch.Tensor, lengths_tensor: torch.Tensor,
                     gap_penalty_param: int = -10) -> torch.Tensor:
    """Linear-gap SW scores of the selected targets (sseqs/sw_bind.cpp:32-73).

    ``lengths_tensor`` holds delimiter positions; target g is
    ``ascii[P[good_idx[g]-1]+1 : P[good_idx[g]]]``.
    """
    q = _cuda(query_indices_tensor, torch.uint8, "query_indices_tensor")
    gi = _cuda(good_idx, torch.int32, "good_idx")
    a = _cuda(ascii_tensor, torch.uint8, "ascii_tensor")
    P = _cuda(lengths_tensor, torch.int32, "lengths_tensor")
    M = int(gi.size(0))
    out = torch.empty(M, dtype=torch.int32, device=q.device)
    if M:
        call("launch_sw_cuda_blosum62", ptr(q), int(q.size(0)), ptr(gi), ptr(a), ptr(P), M,
             ptr(out), int(gap_penalty_param), stream_of(q.device))
    return out


def sw_cuda_affine(query_indices_tensor: torch.Tensor, good_idx: torch.Tensor,
                   ascii_tensor: torch.Tensor, lengths_tensor: torch.Tensor,
                   gap_open: int = 11, gap_extend: int = 1) -> torch.Tensor:
    """Affine-gap SW scores of the selected targets (sseqs/sw_bind.cpp:78-108)."""
    q = _cuda(query_indices_tensor, torch.uint8, "query_indices_tensor")
    gi = _cuda(good_idx, torch.int32, "good_idx")
    a = _cuda(ascii_tensor, torch.uint8, "ascii_tensor")
    P = _cuda(lengths_tensor, torch.int32, "lengths_tensor")
    M = int(gi.size(0))
    out = torch.empty(M, dtype=torch.int32, device=q.device)
    if M:
        call("launch_sw_cuda_affine", ptr(q), int(q.size(0)), ptr(gi), ptr(a), ptr(P), M,
             ptr(out), int(gap_open), int(gap_extend), stream_of(q.device))
    return out


sw = SimpleNamespace(sw_cuda_blosum62=sw_cuda_blosum62, sw_cuda_affine=sw_cuda_affine)
sseqs_sw_ext = sw


# ----------------------------------------------------------------- top-k --

_DT = {torch.float32: _lib.F32, torch.float16: _lib.F16, torch.int32: _lib.I32}


def topk_indices(scores: torch.Tensor, k: int, sorted: bool = True) -> torch.Tensor:
    """Indices of the k largest scores, ties to the smaller index (int64).

    ``sorted=True`` orders them by (score desc, index asc); otherwise they
    come in ascending index order.
    """
    s = scores.contiguous().view(-1)
    if s.dtype not in _DT:
        raise TypeError(f"topk: unsupported dtype {s.dtype}")
    n, k = s.numel(), int(k)
    if k < 0 or k > n:
        raise ValueError(f"topk: k={k} out of range for {n} scores")
    out = torch.empty(k, dtype=torch.int64, device=s.device)
    if k == 0:
        return out
    big = sorted and k > SORTED_MAX
    ws = _lib.workspace(_lib.load().sseqs_topk_ws_bytes(n), s.device)
    call("sseqs_topk", ptr(s), _DT[s.dtype], n, k, int(bool(sorted) and not big), ptr(out),
         ptr(ws), ws.numel(), stream_of(s.device))
    if big:
        out = _order_desc(s, out)
    return out


# sseqs_topk orders its output in LDS up to this k; larger sorted selections
# are ordered after an unsorted select
SORTED_MAX = 16384


def _order_desc(s: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
    """Order selected indices ``idx`` (ascending) by (score desc, index asc),
    NaN last, as sseqs_topk's sorted output: one stable descending sort of
    the selected scores keeps ascending index order among equal scores."""
    v = s[idx]
    if v.is_floating_point():
        nan = torch.isnan(v)
        if bool(nan.any()):
            keep = idx[~nan]
            order = torch.sort(v[~nan], descending=True, stable=True).indices
            return torch.cat([keep[order], idx[nan]])
    order = torch.sort(v, descending=True, stable=True).indices
    return idx[order]


def topk_segments(scores: torch.Tensor, begin, n, k, *, sorted: bool = False,
                  drop_front: int = 0, drop_back: int = 0, add_begin: bool = False
                  ) -> tuple[torch.Tensor, np.ndarray]:
    """topk_indices of many segments of ``scores`` in one launch.

    Segment j is ``scores[begin[j] : begin[j] + n[j]]`` and selects its
    ``k[j]`` largest (host integer arrays).  Of each selection, in
    ascending index order (or (score desc, index asc) when ``sorted``,
    k <= 4096), entries ``drop_front .. k[j] - drop_back - 1`` are kept.
    Returns the kept indices of all segments concatenated (int64, within
    the segment, or within ``scores`` when ``add_begin``) and the start of
    each segment's part.
    """
    s = scores.contiguous().view(-1)
    if s.dtype not in _DT:
        raise TypeError(f"topk_segments: unsupported dtype {s.dtype}")
    begin = np.asarray(begin, dtype=np.int64)
    n = np.asarray(n, dtype=np.int64)
    k = np.minimum(np.asarray(k, dtype=np.int64), n)
    if begin.shape != n.shape or begin.shape != k.shape:
        raise ValueError("topk_segments: begin, n and k must have one entry per segment")
    if np.any(n < 0) or np.any(k < 0) or drop_front < 0 or drop_back < 0:
        raise ValueError("topk_segments: negative size")
    if n.size and (np.any(begin < 0) or np.any(begin + n > s.numel())):
        raise ValueError("topk_segments: segment outside the scores")
    if sorted and k.size and int(k.max()) > 4096:
        raise ValueError("topk_segments: sorted output limited to k <= 4096 per segment")
    cnt = np.maximum(k - drop_front - drop_back, 0)
    base = np.zeros(cnt.size, dtype=np.int64)
    if cnt.size > 1:
        base[1:] = np.cumsum(cnt)[:-1]
    total = int(cnt.sum())
    out = torch.empty(total, dtype=torch.int64, device=s.device)
    if total == 0:
        return out, base
    meta64 = torch.from_numpy(np.concatenate([begin, base])).to(s.device)
    meta32 = torch.from_numpy(np.concatenate([n, k]).astype(np.int32)).to(s.device)
    m = begin.size
    call("sseqs_topk_segments", ptr(s), _DT[s.dtype], ptr(meta64), ptr(meta32),
         ptr(meta32[m:]), ptr(meta64[m:]), m, int(k.max()), int(bool(sorted)), int(drop_front),
         int(drop_back), int(bool(add_begin)), ptr(out), stream_of(s.device))
    return out, base
