"""The oracle's backtrack and a3m writer against vectors of the reference's
own sw_affine_backtrack / to_a3m (tests/golden/a3m_golden.json, written by
tests/golden/make_a3m_golden.py from the reference run on the CPU)."""
import json
from pathlib import Path

import pytest

GOLDEN = json.loads((Path(__file__).resolve().parent / "golden" / "a3m_golden.json").read_text())


@pytest.mark.parametrize("case", range(len(GOLDEN)))
def test_oracle_backtrack_matches_reference(orc, case):
    c = GOLDEN[case]
    for t, want in zip(c["targets"], c["aligned"]):
        got = orc.backtrack(c["query"], t, c["gap_open"], c["gap_extend"])
        assert got == want, t
    assert [int(orc.backtrack(c["query"], t, c["gap_open"], c["gap_extend"])["score"])
            for t in c["targets"]] == c["best"]


@pytest.mark.parametrize("case", range(len(GOLDEN)))
def test_oracle_a3m_matches_reference(orc, case):
    c = GOLDEN[case]
    aligned = [orc.backtrack(c["query"], t, c["gap_open"], c["gap_extend"]) for t in c["targets"]]
    assert orc.a3m(c["query"], aligned) == c["a3m"]


@pytest.mark.parametrize("case", range(len(GOLDEN)))
def test_a3m_writer_matches_reference(case):
    """The package's a3m writer on the reference's own backtrack output gives
    the reference's own text."""
    from sseqs_amd.to_a3m import a3m_text
    c = GOLDEN[case]
    assert a3m_text(c["query"], c["aligned"]) == c["a3m"]


def test_a3m_writer_rows_match_oracle(orc):
    """Many rows at once (the vectorised column conversion) against the
    oracle's per-row writer: mixed case, letters outside the 20 residues,
    gaps on both sides, repeated rows and tied scores; then the same rows
    with one non-ASCII letter and with unequal aligned lengths, which take
    the per-row path."""
    import numpy as np

    from sseqs_amd.to_a3m import a3m_text
    rng = np.random.default_rng(7)
    alph = list("ARNDCQEGHILKMFPSTWYVXBZarndcqe*")
    query = "".join(rng.choice(list("ARNDCQEGHILKMFPSTWYV"), 120))
    aligned = []
    for i in range(400):
        qs = int(rng.integers(0, 60))
        qe = int(rng.integers(qs + 1, 121))
        qa, ta = [], []
        for ch in query[qs:qe]:
            if rng.random() < 0.08:
                qa.append("-")
                ta.append(str(rng.choice(alph)))
            qa.append(ch)
            ta.append("-" if rng.random() < 0.08 else str(rng.choice(alph)))
        aligned.append({"q_aligned": "".join(qa), "t_aligned": "".join(ta),
                        "score": float(rng.integers(0, 40))})
    aligned += [dict(aligned[3]), dict(aligned[5])]
    assert a3m_text(query, aligned) == orc.a3m(query, aligned)
    odd = [dict(a) for a in aligned]
    odd[7]["t_aligned"] = odd[7]["t_aligned"][:-1] + "é"
    assert a3m_text(query, odd) == orc.a3m(query, odd)
    short = [dict(a) for a in aligned]
    short[9]["t_aligned"] = short[9]["t_aligned"][:-1]
    assert a3m_text(query, short) == orc.a3m(query, short)
