"""Size-independent checks at the benchmark's full size (200 M sequences,
56.6 G residues, 35.5 GB packed): every launch that covers the database
must reach its end (an HSA dispatch counts work-items in 32 bits, so one
thread per 8-symbol group would stop at 2^32 groups = 34.4 G symbols).

* the delimiter scan of the generated stream finds exactly n sequences and
  reproduces the generator's index;
* sequences sampled across the whole database (including past 2^32 groups
  and the very last ones) equal the CPU generator's;
* the whole-database cascade for the 500-residue query: prefilter scores of
  a random sample within tolerance of the FP64 closed form, SW scores of
  its hits bit-exact, and the hits are the query's planted homologs;
* msa()'s per-batch cascade for the 200-residue query over all 710 batches:
  sampled prefilter scores within tolerance, and every whole hit's linear,
  affine and normalised scores recomputed from its text.
"""
import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu

N_FULL = 200_000_000


@pytest.fixture(scope="module")
def big(dev):
    from sseqs_amd import synth
    from sseqs_amd.db import SeqDB
    announce("generate 200M-sequence database")
    db, q200, q500, _ = synth.bench_db(N_FULL, device=dev)
    torch.cuda.synchronize()
    yield dict(db=db, queries=[q200, q500])
    del db
    torch.cuda.empty_cache()


def test_full_index_and_content(orc, big):
    from sseqs_amd import synth
    from sseqs_amd.db import SeqDB
    db = big["db"]
    announce("delimiter scan of the full stream")
    scan = SeqDB.from_packed(db.packed, db.n_syms)
    assert scan.n == db.n and scan.orig_id is None
    assert torch.equal(scan.seq_off, db.seq_off) and torch.equal(scan.seq_len, db.seq_len)
    del scan
    rng = np.random.default_rng(1)
    off = db.seq_off.cpu().numpy()
    past = np.nonzero(off >= (1 << 32) * 8)[0]
    ids = np.unique(np.concatenate([rng.choice(db.n, 1500, replace=False),
                                    np.arange(db.n - 50, db.n), past[:20],
                                    rng.choice(past, 300, replace=False)]))
    cls = db.extra["classes"]
    want = orc.gen_seqs(synth.gen_params(0, hom_period=synth.BENCH_HOM_PERIOD),
                        db.extra["gen_queries"], cls.start, cls.length, ids)
    got = db.strings(torch.from_numpy(ids).to(db.device), 1024)
    assert got == want


def test_full_cascade_q500(orc, big):
    from sseqs_amd.query import prepare
    from sseqs_amd.search import global_search
    db, query = big["db"], big["queries"][1]
    announce("global cascade q_len=500 over 200M sequences")
    q = prepare(query)
    gh = global_search(db, q, top_prefilter=100_000, top_hits=4096)
    ids = gh.ids.cpu().numpy()
    hq = db.extra["hom_owner"].cpu().numpy()[ids]
    assert np.mean(hq[:1000] == 1) > 0.95
    # each single query owns 5 % of the database
    own = db.extra["hom_owner"]
    assert abs(float((own == 0).float().mean()) - 0.05) < 1e-3
    assert abs(float((own == 1).float().mean()) - 0.05) < 1e-3
    # SW of the hits against the oracle
    sel = ids[:64]
    rows = db.strings(torch.from_numpy(sel).to(db.device), 1024)
    qc = orc.encode_query_sw(query)
    assert gh.scores[:64].cpu().tolist() == [orc.sw_affine_one(qc, r) for r in rows]
    # prefilter of a random sample against FP64
    rng = np.random.default_rng(2)
    samp = np.sort(rng.choice(db.n, 400, replace=False))
    rows = db.strings(torch.from_numpy(samp).to(db.device), 1024)
    A, B = orc.weighted_luts(query, 4)
    qrev = orc.encode_query_rev(query)
    args = [(r.encode(), [0], [len(r)], 10**6, orc.nextpow2(max(len(query), len(r), 16)))
            for r in rows]
    want = np.array([orc.corr(qrev, A, B, 4, *a)[0] for a in args])
    tol = np.array([orc.corr_tol(qrev, A, B, 4, *a)[0] for a in args])
    got = gh.prefilter.cpu().numpy()[samp]
    assert np.all(np.abs(got - want) <= tol)
    # the same cascade with linear-gap SW (-4): the same candidates, the
    # hits' linear scores bit-exact, and still the query's homologs
    announce("global cascade q_len=500 linear -4 over 200M sequences")
    gl = global_search(db, q, top_prefilter=100_000, top_hits=4096, linear_gap=-4)
    assert torch.equal(gl.candidates, gh.candidates)
    lids = gl.ids.cpu().numpy()
    rows = db.strings(torch.from_numpy(lids[:64]).to(db.device), 1024)
    assert gl.scores[:64].cpu().tolist() == [orc.sw_linear_one(qc, r, -4) for r in rows]
    sc = gl.scores.cpu().numpy()
    assert np.all(sc[:-1] >= sc[1:])
    assert np.mean(db.extra["hom_owner"].cpu().numpy()[lids[:1000]] == 1) > 0.95


def test_full_batch_cascade_q200(orc, big):
    """msa()'s per-batch cascade over the full database (710 batches of
    50 MB): the prefilter scores of sampled segments of three batches within
    tolerance of FP64, every hit's linear and affine scores and normalised
    score recomputed from its text (whole hits), and the hits found."""
    from sseqs_amd.db import SeqDB
    from sseqs_amd.query import prepare
    from sseqs_amd.search import batch_search
    db, query = big["db"], big["queries"][0]
    bs = 50_000_000
    announce("per-batch cascade q_len=200 over 200M sequences")
    q = prepare(query)
    keep = {0: None, 355: None, 709: None}

    def grab(j, s):
        if j in keep:
            keep[j] = s.float().cpu().numpy()
    hits = batch_search(db, query, q, bs=bs, on_scores=grab)
    plan = db.batch_plan(bs)
    assert plan.n_batches == 710 and len(hits) == 2 * 710
    # prefilter of sampled segments of three batches against FP64
    segs = SeqDB(db.packed, db.n_syms, plan.seg_off, plan.seg_len, None)
    A, B = orc.weighted_luts(query, 4)
    qrev = orc.encode_query_rev(query)
    rng = np.random.default_rng(5)
    for b, got in keep.items():
        assert got is not None and got.size == plan.nseg[b]
        loc = np.sort(rng.choice(got.size, 200, replace=False))
        rows = segs.strings(torch.from_numpy(plan.seg_base[b] + loc).to(db.device), 1024)
        n = min(max(orc.nextpow2(max(len(query), int(plan.maxlen[b]))), 16), 2048)
        want = np.array([orc.corr(qrev, A, B, 4, r.encode(), [0], [len(r)], 10**6, n)[0]
                         for r in rows])
        tol = np.array([orc.corr_tol(qrev, A, B, 4, r.encode(), [0], [len(r)], 10**6, n)[0]
                        for r in rows])
        assert np.all(np.abs(got[loc] - want) <= tol), b
    # hits: scores recomputed from their text where the text is whole
    qc = orc.encode_query_sw(query)
    checked = 0
    for i, h in enumerate(hits):
        b = i // 2
        assert (h.chunk, h.batch) == (b // 20, b % 20)
        if len(h.seq) < 500:
            assert h.linear == orc.sw_linear_one(qc, h.seq, -1)
            assert h.affine == orc.sw_affine_one(qc, h.seq[1:] + "@")
            norm = np.float32(h.affine) / np.sqrt(np.float32(len(query) * int(plan.maxlen[b])))
            assert np.float32(h.norm) == norm
            checked += 1
    assert checked > 500
    assert max(h.norm for h in hits) > 4.0
