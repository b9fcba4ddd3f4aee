"""Admission of the continuous-batching engine on the host (no GPU): a
sweep takes the oldest waiting requests while their buffers and the sweep's
own fit the HBM budget, at least one, at most ``max_batch``; with
``top_prefilter="calibrated"`` each request's buffers are sized by the
proposer's operating point for its length."""
from collections import deque
from types import SimpleNamespace

import numpy as np

from sseqs_amd.serve import SearchEngine, _Request, query_bytes, sweep_bytes


def _engine(kc, budget, max_batch=64, proposer=None):
    e = SearchEngine.__new__(SearchEngine)  # no device loop
    e.db = SimpleNamespace(n=2_000_000, device="cpu")
    e.kc, e.kh, e.hit_width = kc, 200, 500
    e.max_batch, e.budget, e.proposer = max_batch, budget, proposer
    e._q = deque()
    return e


def _push(e, lens):
    for L in lens:
        e._q.append(_Request("A" * L, None, 0.0))


def test_admission_by_budget():
    lens = [300, 120, 900, 50, 700]
    e = _engine(100_000, budget=0)
    _push(e, lens)
    # a budget below one request still admits the oldest, alone
    assert [len(r.query) for r in e._admit()] == [300]
    # a budget of exactly the first three requests' buffers admits those three
    need3 = sweep_bytes(e.db, 900, 200) + sum(query_bytes(e.db, L, 100_000, 200) for L in (120, 900, 50))
    e.budget = need3
    assert [len(r.query) for r in e._admit()] == [120, 900, 50]
    assert [len(r.query) for r in e._q] == [700]
    # one byte less leaves the third waiting
    e = _engine(100_000, budget=need3 - 1)
    _push(e, [120, 900, 50])
    assert [len(r.query) for r in e._admit()] == [120, 900]
    # max_batch caps the sweep whatever the budget
    e = _engine(100_000, budget=1 << 62, max_batch=2)
    _push(e, lens)
    assert len(e._admit()) == 2 and len(e._q) == 3


def test_admission_sizes_calibrated_counts():
    prop = SimpleNamespace(top_prefilter_for=lambda qs: np.array(
        [1_000_000 if len(q) < 100 else 10_000 for q in qs]))
    e = _engine("calibrated", budget=0, proposer=prop)
    assert e._kc("A" * 50) == 1_000_000 and e._kc("A" * 500) == 10_000
    # room for a short query's million candidates, or for many long queries
    one_short = sweep_bytes(e.db, 50, 200) + query_bytes(e.db, 50, 1_000_000, 200)
    e.budget = one_short
    _push(e, [50, 500, 500])
    got = e._admit()
    assert [len(r.query) for r in got] == [50]
    assert query_bytes(e.db, 500, 10_000, 200) < query_bytes(e.db, 50, 1_000_000, 200)
