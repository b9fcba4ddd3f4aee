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


def to_a3m(query: str, seqs: list[str], filename: str, gap_open: int = 11,
           gap_extend: int = 1) -> str:
    """Write the a3m file for ``seqs`` aligned to ``query``; returns the text."""
    aligned: list[dict] = []
    for start in range(0, len(seqs), 1024):
        aligned += sw_affine_backtrack(query, list(seqs[start:start + 1024]),
                                       gap_open=gap_open, gap_extend=gap_extend)[-1]
    text = a3m_text(query, aligned)
    if filename is not None:
        with open(filename, "w") as fh:
            fh.write(text)
    return text
