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
