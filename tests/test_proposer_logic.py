"""Host logic of the learned proposal (no GPU): the query rows a query is
embedded as (whole, then windows), the calibrated candidate count per query
length, and the per-query recall gate that decides between the learned
proposal and the plain prefilter."""
from types import SimpleNamespace

import numpy as np
import pytest

from sseqs_amd.learned import Proposer


def _prop(**kw):
    return Proposer(SimpleNamespace(n=1_000_000), **kw)


def test_rows_of_windows():
    p = _prop(window=384, stride=192)
    q_short, q_exact, q_long = "A" * 200, "C" * 384, "".join("ACDEFGHIKL"[i % 10] for i in range(1000))
    rows, count = p.rows_of([q_short, q_exact, q_long])
    assert list(count) == [1, 1, 1 + 5]
    assert rows[0] == q_short and rows[1] == q_exact and rows[2] == q_long
    wins = rows[3:]
    # windows every 192 residues, the last one ending at the query's end
    assert [len(w) for w in wins] == [384] * 5
    assert wins[0] == q_long[:384] and wins[1] == q_long[192:576] and wins[2] == q_long[384:768]
    assert wins[3] == q_long[576:960] and wins[4] == q_long[-384:]
    # no windows when disabled
    rows, count = _prop(window=0).rows_of([q_long])
    assert rows == [q_long] and list(count) == [1]


def test_group_size_checked():
    with pytest.raises(ValueError):
        _prop(group=20)
    with pytest.raises(ValueError):
        _prop(group=256)


def test_calibrated_counts_and_gate():
    p = _prop()
    assert p.meets_gate(["A" * 50], [10]).all()  # uncalibrated: always learned
    with pytest.raises(RuntimeError):
        p.top_prefilter_for(["A" * 50])
    p.k_by_len = [(0, 100, 5000), (100, 200, 300), (200, 2049, 7000)]
    p.calibration = {"top_prefilter": 100_000}
    p.use_learned = False
    qs = ["A" * 99, "A" * 100, "A" * 199, "A" * 200, "A" * 3000]
    assert list(p.top_prefilter_for(qs)) == [5000, 300, 300, 7000, 7000]
    counts = np.array([5000, 299, 100_000, 6999, 10**6])
    assert list(p.meets_gate(qs, counts)) == [True, False, True, False, True]
    # the gate held for every calibration query at 100 000: that count or more passes
    p.use_learned = True
    assert list(p.meets_gate(qs, [100_000, 1, 1, 100_000, 1])) == [True, False, False, True, False]
    p.enforce_gate = False
    assert p.meets_gate(qs, [1] * 5).all()


def test_fused_proposal_arguments_checked():
    """sseqs_propose_max rejects its bad arguments on the host, before any
    launch: query rows not a multiple of 16 or above 64, more than 32 output
    rows, offsets that do not increase or leave the G rows."""
    import ctypes as C

    from sseqs_amd._lib import SseqsError, call
    for G, off in ((80, [0, 1]), (24, [0, 1]), (64, list(range(34))), (64, [0, 3, 3]),
                   (64, [0, 5, 2]), (32, [0, 40]), (32, [-1, 4])):
        arr = np.array(off, np.int32)
        with pytest.raises(SseqsError, match="propose_max: "):
            call("sseqs_propose_max", C.c_void_p(16), 16, C.c_void_p(16), G,
                 arr.ctypes.data, len(off) - 1, C.c_void_p(16), 16, None)
