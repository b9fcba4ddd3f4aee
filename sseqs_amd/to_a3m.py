"""A3M writer (mirror of sseqs/to_a3m.py).

Aligns every hit to the query with ``sw_affine_backtrack`` in blocks of
1024 (sseqs/to_a3m.py:10-14), orders the alignments by score, and writes one
full-width row per distinct alignment: a '-' prefix up to the first
occurrence of the aligned query fragment in the query (``str.find``, as the
reference does), upper-case residues in query columns, lower-case
insertions, '-' deletions and a '-' suffix (sseqs/to_a3m.py:23-49).  Rows
with score 0, empty alignments, rows containing '@' and repeated rows are
skipped (sseqs/to_a3m.py:56-70).  Equal scores keep their input order (the
reference's numpy quicksort leaves it open).
"""
from __future__ import annotations

import numpy as np

from .sw import sw_affine_backtrack


def build_row(query: str, ares: dict, middle: str | None = None) -> str:
    """One a3m row (sseqs/to_a3m.py:23-49); ``middle`` may bring the aligned
    columns already converted (see ``_middles``)."""
    q_loc, t_loc = ares["q_aligned"], ares["t_aligned"]
    q_compact = q_loc.replace("-", "")
    start = query.find(q_compact)
    if start == -1:
        raise ValueError("Cannot locate aligned fragment in query.")
    if middle is None:
        middle = "".join("-" if t == "-" else (t.upper() if q != "-" else t.lower())
                         for q, t in zip(q_loc, t_loc))
    return "-" * start + middle + "-" * (len(query) - (start + len(q_compact)))


def _middles(aligned: list[dict]) -> list[str] | None:
    """The aligned columns of every row at once: the target letter in upper
    case under a query residue, in lower case under a query gap, '-' kept.
    For ASCII text (where str.upper/lower only move 'a'-'z' / 'A'-'Z') this
    is one vectorised pass over all rows; None when any row is not ASCII or
    has unequal aligned lengths (build_row then converts it by itself)."""
    qs = [a["q_aligned"] for a in aligned]
    ts = [a["t_aligned"] for a in aligned]
    if any(len(q) != len(t) for q, t in zip(qs, ts)):
        return None
    Q, T = "".join(qs), "".join(ts)
    if not (Q.isascii() and T.isascii()):
        return None
    if not T:
        return ["" for _ in aligned]
    qb = np.frombuffer(Q.encode("ascii"), dtype=np.uint8)
    tb = np.frombuffer(T.encode("ascii"), dtype=np.uint8)
    up = np.where((tb >= 97) & (tb <= 122), tb - 32, tb)
    low = np.where((tb >= 65) & (tb <= 90), tb + 32, tb)
    mid = np.where(qb == 45, low, up).astype(np.uint8).tobytes().decode("ascii")
    ends = np.cumsum([len(t) for t in ts])
    return [mid[e - len(t):e] for e, t in zip(ends.tolist(), ts)]


def a3m_text(query: str, aligned: list[dict]) -> str:
    # the reference's own ordering, np.argsort(scores)[::-1] over the Python
    # float scores (sseqs/to_a3m.py:16-17), so ties come out as numpy orders them
    scores = [float(a["score"]) for a in aligned]
    order = np.argsort(scores)[::-1]
    mids = _middles(aligned)
    out = [f">original_query\n{query}\n"]
    seen = set()
    for i, idx in enumerate(order):
        ares = aligned[idx]
        if ares["score"] == 0 or not ares["q_aligned"]:
            continue
        row = build_row(query, ares, None if mids is None else mids[idx])
        if "@" in row or row in seen:
            continue
        seen.add(row)
        out.append(f">target_{i} score={ares['score']}\n{row}\n")
    return "".join(out)


def _dicts(blocks) -> list[dict]:
    """Backtrack dicts (sseqs/sw.py:207-217) of row blocks."""
    out = []
    for qh, th_, ih in blocks:
        A = qh.shape[1]
        for b in range(ih.shape[0]):
            n = int(ih[b, 0])
            out.append({
                "q_aligned": qh[b, A - n:].tobytes().decode("latin-1") if n else "",
                "t_aligned": th_[b, A - n:].tobytes().decode("latin-1") if n else "",
                "q_start_orig_0based": int(ih[b, 1]), "q_end_orig_0based": int(ih[b, 2]),
                "t_start_orig_0based": int(ih[b, 3]), "t_end_orig_0based": int(ih[b, 4]),
                "score": float(ih[b, 5])})
    return out


def a3m_text_rows(query: str, blocks) -> str:
    """``a3m_text`` of backtrack rows without building dicts: ``blocks`` is a
    list of (q rows uint8 [B, A], t rows uint8 [B, A], info int32 [B, 6]),
    each alignment right-aligned in its rows (the device backtrack's
    layout).  The rows are formatted by the library's host routine
    ``sseqs_a3m_text`` in the order a3m_text uses (numpy's argsort of the
    float scores, reversed); a query or row that is not ASCII goes through
    a3m_text."""
    from . import _lib
    blocks = [(np.ascontiguousarray(q, np.uint8), np.ascontiguousarray(t, np.uint8),
               np.ascontiguousarray(i, np.int32)) for q, t, i in blocks]
    qb = query.encode("latin-1", errors="replace")
    if not blocks or not query.isascii():
        return a3m_text(query, _dicts(blocks))
    if len(blocks) == 1:
        Q, T, info = blocks[0][0].reshape(-1), blocks[0][1].reshape(-1), blocks[0][2]
    else:
        Q = np.concatenate([q.reshape(-1) for q, _, _ in blocks])
        T = np.concatenate([t.reshape(-1) for _, t, _ in blocks])
        info = np.concatenate([i for _, _, i in blocks])
    n = np.ascontiguousarray(info[:, 0])
    base, offs = 0, []
    for q, _, i in blocks:
        A = q.shape[1]
        offs.append(base + np.arange(i.shape[0], dtype=np.int64) * A + (A - i[:, 0]))
        base += q.size
    off = np.ascontiguousarray(np.concatenate(offs).astype(np.int64))
    score = np.ascontiguousarray(info[:, 5])
    order = np.ascontiguousarray(np.argsort(score.astype(np.float64))[::-1].astype(np.int64))
    lib = _lib.load()
    cap = int(lib.sseqs_a3m_bound(len(qb), n.size, int(np.maximum(n, 0).sum())))
    out = np.empty(max(cap, 1), dtype=np.uint8)
    p = lambda a: a.ctypes.data
    r = int(lib.sseqs_a3m_text(qb, len(qb), p(Q), p(T), p(off), p(n), p(score), p(order), n.size,
                               p(out), cap))
    if r == -2:
        return a3m_text(query, _dicts(blocks))
    if r == -1:
        raise ValueError("Cannot locate aligned fragment in query.")
    if r < 0:
        raise _lib.SseqsError(f"sseqs_a3m_text failed ({r})")
    return out[:r].tobytes().decode("ascii")


def to_a3m(query: str, seqs: list[str], filename: str, gap_open: int = 11,
           gap_extend: int = 1) -> str:
    """Write the a3m file for ``seqs`` aligned to ``query``; returns the text."""
    blocks = [sw_affine_backtrack(query, list(seqs[start:start + 1024]), gap_open=gap_open,
                                  gap_extend=gap_extend, rows=True)
              for start in range(0, len(seqs), 1024)]
    text = a3m_text_rows(query, blocks)
    if filename is not None:
        with open(filename, "w") as fh:
            fh.write(text)
    return text
