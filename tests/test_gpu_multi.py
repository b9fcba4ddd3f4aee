"""The batched sweep (many queries of mixed lengths in one sweep) against
single-query runs and the oracle.

Every query's candidates, candidate SW scores, hits and alignments from the
batched sweep must equal those of its own global_search + backtrack_hits
(same kernels, same tie rules), whatever the SW queue parameters; the SW
scores must equal the oracle's; and a HIP-graph capture of the sweep,
replayed several times with other work in between, must give the eager
results every time.
"""
import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu

LENS = [55, 64, 120, 200, 333, 512, 513, 700, 999, 1024, 1100]
TOP_PRE, TOP_HITS = 3000, 400


@pytest.fixture(scope="module")
def mworld(orc, dev):
    from sseqs_amd import synth
    from sseqs_amd.db import SeqDB
    queries = [synth.make_query(L, 40 + i, seed=9) for i, L in enumerate(LENS)]
    db = SeqDB.synthetic(30000, queries, seed=17, device=dev)
    ascii = orc.xbit_unpack(db.packed.cpu().numpy(), db.n_syms)
    return dict(db=db, queries=queries, ascii=ascii, off=db.seq_off.cpu().numpy(),
                ln=db.seq_len.cpu().numpy())


def _single(db, query):
    from sseqs_amd.query import prepare
    from sseqs_amd.search import BacktrackWorkspace, alignments_to_dicts, backtrack_hits
    from sseqs_amd.search import global_search
    q = prepare(query)
    gh = global_search(db, q, TOP_PRE, TOP_HITS)
    ws = BacktrackWorkspace(TOP_HITS, len(query), 500, db.device)
    al = alignments_to_dicts(backtrack_hits(db, q, gh.ids, ws))
    return gh, al


def _snapshot(ms):
    return [(r.candidates.clone(), r.cand_scores.clone(), r.ids.clone(), r.scores.clone(),
             r.alignment_dicts()) for r in ms.results()]


@pytest.mark.parametrize("tuning", [None, {"sw": {"1-8": {"item": 256, "blocks_per_cu": 1},
                                                   "9-16": {"item": 500, "blocks_per_cu": 3},
                                                   "17-32": {"item": 4096, "blocks_per_cu": 8}}}])
def test_multi_equals_single(orc, mworld, tuning):
    from sseqs_amd.multi import MultiSearch
    announce(f"batched sweep of {len(LENS)} queries, tuning={tuning}")
    db = mworld["db"]
    ms = MultiSearch(db, mworld["queries"], TOP_PRE, TOP_HITS, tuning=tuning or {})
    ms.run()
    res = ms.results()
    qc_all = [orc.encode_query_sw(q) for q in mworld["queries"]]
    for i, (query, r) in enumerate(zip(mworld["queries"], res)):
        gh, al = _single(db, query)
        assert torch.equal(r.candidates, gh.candidates), i
        assert torch.equal(r.cand_scores, gh.cand_scores), i
        assert torch.equal(r.ids, gh.ids), i
        assert torch.equal(r.scores, gh.scores), i
        assert r.alignment_dicts() == al, i
        cand = r.candidates.cpu().numpy()
        sw = orc.sw_affine_segments(qc_all[i], mworld["ascii"], mworld["off"][cand],
                                    mworld["ln"][cand])
        assert np.array_equal(r.cand_scores.cpu().numpy(), sw), i
        # the planted homologs of each query lead its hits
        hq = db.extra["hom_query"].cpu().numpy()
        assert np.mean(hq[r.ids[:20].cpu().numpy()] == i) > 0.8, i


def test_multi_graph_replay(mworld):
    from sseqs_amd.multi import MultiSearch
    from sseqs_amd.query import prepare
    from sseqs_amd.search import prefilter_scores
    announce("batched sweep captured in a HIP graph, replayed 3 times")
    db = mworld["db"]
    ms = MultiSearch(db, mworld["queries"][:6], TOP_PRE, TOP_HITS)
    ms.run()
    torch.cuda.synchronize()
    want = _snapshot(ms)
    ms.capture()
    other = prepare(mworld["queries"][-1])
    for rep in range(3):
        # the program's own eager work between replays, then clear the
        # outputs so that a replay that writes nothing would be caught
        prefilter_scores(db, other)
        for t in (ms.cand, ms.cs, ms.hit_ids, ms.hit_sc, ms.info):
            t.fill_(-3)
        torch.cuda.synchronize()
        ms.replay()
        torch.cuda.synchronize()
        got = _snapshot(ms)
        for a, b in zip(got, want):
            for x, y in zip(a[:4], b[:4]):
                assert torch.equal(x, y), rep
            assert a[4] == b[4], rep


def test_multi_per_query_candidate_counts(mworld):
    """One candidate count per query (lists laid out back to back): every
    query's candidates, scores and hits equal its single-query run with its
    own count, with the plain prefilter and after a graph capture."""
    from sseqs_amd.multi import MultiSearch
    from sseqs_amd.query import prepare
    from sseqs_amd.search import global_search
    announce("batched sweep with per-query candidate counts")
    db = mworld["db"]
    qs = mworld["queries"][:7]
    kc = [3000, 500, 4096, 777, 12000, 401, 2048]
    ms = MultiSearch(db, qs, kc, TOP_HITS)
    assert ms.cand.numel() == sum(kc) and ms.kh == TOP_HITS
    ms.run()
    torch.cuda.synchronize()
    want = _snapshot(ms)
    for i, (q, r) in enumerate(zip(qs, ms.results())):
        gh = global_search(db, prepare(q), kc[i], TOP_HITS)
        assert r.candidates.numel() == kc[i], i
        assert torch.equal(r.candidates, gh.candidates), i
        assert torch.equal(r.cand_scores, gh.cand_scores), i
        assert torch.equal(r.ids, gh.ids) and torch.equal(r.scores, gh.scores), i
    ms.capture()
    for t in (ms.cand, ms.cs, ms.hit_ids, ms.hit_sc):
        t.fill_(-3)
    torch.cuda.synchronize()
    ms.replay()
    torch.cuda.synchronize()
    for a, b in zip(_snapshot(ms), want):
        for x, y in zip(a[:4], b[:4]):
            assert torch.equal(x, y)
        assert a[4] == b[4]
    with pytest.raises(ValueError):
        MultiSearch(db, qs, kc[:3], TOP_HITS)
