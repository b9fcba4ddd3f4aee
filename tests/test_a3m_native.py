"""The library's host a3m formatter (sseqs_a3m_text, through
to_a3m.a3m_text_rows) against the Python a3m_text of the same alignments:
the reference's golden alignments and a3m text (tests/golden/a3m_golden.json),
and random alignments with every rule of sseqs/to_a3m.py:16-72 exercised --
tied scores, score 0, empty alignments, '@' in a row, repeated rows, lower
case letters under query gaps, fragments found at their first occurrence --
plus the Python fallback for text that is not ASCII and the error for a
fragment missing from the query.  Runs on the CPU (host code only)."""
import json
from pathlib import Path

import numpy as np
import pytest

from sseqs_amd.to_a3m import a3m_text, a3m_text_rows

GOLDEN = json.loads((Path(__file__).resolve().parent / "golden" / "a3m_golden.json").read_text())


def rows_of(aligned, A=None):
    """Right-aligned row blocks (the device backtrack's layout) of dicts."""
    n = [len(a["q_aligned"]) for a in aligned]
    A = A or max(n + [1]) + 3
    q = np.zeros((len(aligned), A), np.uint8)
    t = np.zeros((len(aligned), A), np.uint8)
    info = np.zeros((len(aligned), 6), np.int32)
    for b, a in enumerate(aligned):
        k = n[b]
        if k:
            q[b, A - k:] = np.frombuffer(a["q_aligned"].encode("latin-1"), np.uint8)
            t[b, A - k:] = np.frombuffer(a["t_aligned"].encode("latin-1"), np.uint8)
        info[b] = (k, a["q_start_orig_0based"], a["q_end_orig_0based"], a["t_start_orig_0based"],
                   a["t_end_orig_0based"], int(a["score"]))
    return q, t, info


@pytest.mark.parametrize("case", range(len(GOLDEN)))
def test_golden(case):
    g = GOLDEN[case]
    blk = rows_of(g["aligned"])
    assert a3m_text_rows(g["query"], [blk]) == g["a3m"]
    # split into two blocks of other widths
    h = len(g["aligned"]) // 2
    assert a3m_text_rows(g["query"], [rows_of(g["aligned"][:h], 90),
                                      rows_of(g["aligned"][h:], 130)]) == g["a3m"]


def _random_case(rng, Lq=60, n=300):
    AA = "ARNDCQEGHILKMFPSTWYV"
    query = "".join(rng.choice(list(AA), Lq))
    aligned = []
    for b in range(n):
        kind = rng.integers(0, 10)
        if kind == 0:
            aligned.append(dict(q_aligned="", t_aligned="", q_start_orig_0based=0,
                                q_end_orig_0based=0, t_start_orig_0based=0, t_end_orig_0based=0,
                                score=float(rng.integers(0, 3))))
            continue
        s = int(rng.integers(0, Lq - 5))
        e = int(rng.integers(s + 1, Lq + 1))
        qa, ta = [], []
        for c in query[s:e]:
            r = rng.random()
            if r < 0.1:  # query residue against a target gap
                qa.append(c)
                ta.append("-")
            elif r < 0.2:  # insertion: query gap, then the residue
                qa += ["-", c]
                ta += [rng.choice(list(AA + "@")) if rng.random() < 0.05 else rng.choice(list(AA)),
                       rng.choice(list(AA.lower() + AA))]
            else:
                qa.append(c)
                ta.append(c if rng.random() < 0.6 else rng.choice(list(AA.lower() + AA)))
        sc = float(rng.choice([0, 7, 7, 7, 12, 30, int(rng.integers(1, 400))]))
        aligned.append(dict(q_aligned="".join(qa), t_aligned="".join(ta), q_start_orig_0based=s,
                            q_end_orig_0based=e - 1, t_start_orig_0based=0, t_end_orig_0based=0,
                            score=sc))
        if rng.random() < 0.1:  # an exact repeat
            aligned.append(dict(aligned[-1]))
    return query, aligned


@pytest.mark.parametrize("seed", range(6))
def test_random_alignments(seed):
    rng = np.random.default_rng(seed)
    query, aligned = _random_case(rng)
    want = a3m_text(query, aligned)
    assert want.count("\n>") > 20
    assert a3m_text_rows(query, [rows_of(aligned)]) == want
    cut = len(aligned) // 3
    assert a3m_text_rows(query, [rows_of(aligned[:cut]), rows_of(aligned[cut:], 200)]) == want


def test_fragment_found_at_first_occurrence():
    query = "ACDACDACD"
    al = [dict(q_aligned="CD", t_aligned="cD", q_start_orig_0based=4, q_end_orig_0based=5,
               t_start_orig_0based=0, t_end_orig_0based=1, score=5.0)]
    assert a3m_text_rows(query, [rows_of(al)]) == a3m_text(query, al)
    assert a3m_text_rows(query, [rows_of(al)]).endswith("\n-CD------\n")


def test_not_ascii_uses_the_python_path():
    query = "ACDÉFG"
    al = [dict(q_aligned="CD-É", t_aligned="cDäé", q_start_orig_0based=1, q_end_orig_0based=3,
               t_start_orig_0based=0, t_end_orig_0based=3, score=9.0)]
    assert a3m_text_rows(query, [rows_of(al)]) == a3m_text(query, al)
    al2 = [dict(al[0], q_aligned="CD", t_aligned="ä-")]
    assert a3m_text_rows("ACDEFG", [rows_of(al2)]) == a3m_text("ACDEFG", al2)


def test_missing_fragment_raises():
    al = [dict(q_aligned="WW", t_aligned="WW", q_start_orig_0based=0, q_end_orig_0based=1,
               t_start_orig_0based=0, t_end_orig_0based=1, score=3.0)]
    with pytest.raises(ValueError):
        a3m_text("ACDEFG", al)
    with pytest.raises(ValueError):
        a3m_text_rows("ACDEFG", [rows_of(al)])


def test_empty():
    assert a3m_text_rows("ACD", []) == a3m_text("ACD", [])
    z = np.zeros((0, 4), np.uint8)
    assert a3m_text_rows("ACD", [(z, z, np.zeros((0, 6), np.int32))]) == a3m_text("ACD", [])
