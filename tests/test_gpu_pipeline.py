"""End-to-end parity of the two cascades and of msa()/to_a3m on a synthetic
database, stage by stage against the oracle.

Stages after the prefilter are integer and compared bit-exactly given the
same prefilter scores; the prefilter itself is compared with the FP64
closed form under the tolerance stated in test_gpu_kernels.py, and the
candidate set it induces must agree with the FP64 one except for sequences
whose FP64 score lies within that tolerance of the selection threshold.
"""
import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu

N_SEQ = 30000


@pytest.fixture(scope="module")
def world(orc, dev):
    from sseqs_amd import synth
    from sseqs_amd.db import SeqDB
    queries = [synth.make_query(200, 0, seed=1), synth.make_query(90, 1, seed=1)]
    db = SeqDB.synthetic(N_SEQ, queries, seed=3, device=dev)
    ascii = orc.xbit_unpack(db.packed.cpu().numpy(), db.n_syms)
    off = db.seq_off.cpu().numpy()
    ln = db.seq_len.cpu().numpy()
    return dict(db=db, queries=queries, ascii=ascii, off=off, ln=ln)


def _fp64_prefilter(orc, ascii, off, ln, query):
    """FP64 prefilter scores of every sequence and their tolerances."""
    A, B = orc.weighted_luts(query, 4)
    qrev = orc.encode_query_rev(query)
    out = np.zeros(ln.size)
    tol = np.zeros(ln.size)
    Lq = len(query)
    Ns = np.array([orc.nextpow2(max(Lq, int(L), 16)) for L in ln])
    for N in np.unique(Ns):
        sel = np.nonzero(Ns == N)[0]
        args = (ascii, off[sel].astype(np.int32), ln[sel], int(N), int(N))
        out[sel] = orc.corr(qrev, A, B, 4, *args)
        tol[sel] = orc.corr_tol(qrev, A, B, 4, *args, out[sel])
    return out, tol


def check_selection(orc, dev_scores, want, tol, k):
    """The device's top k (by its own scores) against the FP64 top k: the
    same set except for sequences whose FP64 score is within tolerance of
    the k-th FP64 score (an element of one set and not the other satisfies
    x64(e) + tol(e) >= x64(f) - tol(f) for some f on the other side)."""
    cand = np.sort(orc.topk_desc(np.asarray(dev_scores, np.float64), k))
    ref = np.sort(orc.topk_desc(want, k))
    if k == 0:
        return cand
    thr = np.sort(want)[::-1][k - 1]
    diff = np.setxor1d(cand, ref)
    if diff.size:
        assert np.all(np.abs(want[diff] - thr) <= tol[diff] + tol[diff].max()), diff
    return cand


@pytest.mark.parametrize("qi", [0, 1])
def test_global_cascade(orc, world, qi):
    from sseqs_amd.query import prepare
    from sseqs_amd.search import global_search
    announce(f"global cascade query {qi}")
    db, query = world["db"], world["queries"][qi]
    q = prepare(query)
    gh = global_search(db, q, top_prefilter=3000, top_hits=500)
    pre = gh.prefilter.cpu().numpy()
    want, tol = _fp64_prefilter(orc, world["ascii"], world["off"], world["ln"], query)
    assert np.all(np.abs(pre - want) <= tol)
    # candidates: exact top-k of the device scores, and the FP64 top-k up
    # to threshold ties within the tolerance
    cand = gh.candidates.cpu().numpy()
    assert np.array_equal(cand, check_selection(orc, pre, want, tol, 3000))
    # affine SW of every candidate: bit exact
    qc = orc.encode_query_sw(query)
    sw = orc.sw_affine_segments(qc, world["ascii"], world["off"][cand], world["ln"][cand])
    assert np.array_equal(gh.cand_scores.cpu().numpy(), sw)
    top = orc.topk_desc(sw, 500)
    assert np.array_equal(gh.ids.cpu().numpy(), cand[top])
    assert np.array_equal(gh.scores.cpu().numpy(), sw[top])
    # planted homologs of this query dominate the hits
    hq = world["db"].extra["hom_query"].cpu().numpy()
    assert np.mean(hq[cand[top[:50]]] == qi) > 0.9


def test_unsorted_stream(world):
    """A database file whose sequences are not in length order: the copy's
    index (sorted by length, with orig_id back to the stream) names the same
    sequences, the prefilter gives bit-identical scores per sequence (its
    tiles now read scattered stream positions), and the global cascade
    selects the same candidates and hit scores."""
    from sseqs_amd import synth
    from sseqs_amd.query import prepare
    from sseqs_amd.search import global_search, prefilter_scores
    announce("unsorted stream")
    db, query = world["db"], world["queries"][0]
    sh, perm = synth.shuffled_copy(db)
    assert sh.n == db.n and sh.orig_id is not None
    back = perm[sh.orig_id]  # index entry j of the copy is sequence back[j] of db
    assert torch.equal(sh.seq_len, db.seq_len[back])
    ids = torch.arange(0, sh.n, 97, device=sh.device)
    assert sh.strings(ids, 1200) == db.strings(back[ids], 1200)
    q = prepare(query)
    s_db, s_sh = prefilter_scores(db, q), prefilter_scores(sh, q)
    assert torch.equal(s_db[back], s_sh)
    top = torch.sort(s_db, descending=True).values
    assert float(top[2999]) != float(top[3000])  # no tie at the selection threshold
    g1 = global_search(db, q, top_prefilter=3000, top_hits=500)
    g2 = global_search(sh, q, top_prefilter=3000, top_hits=500)
    assert torch.equal(torch.sort(back[g2.candidates]).values, torch.sort(g1.candidates).values)
    assert torch.equal(torch.sort(g2.scores).values, torch.sort(g1.scores).values)


def test_global_cascade_linear_gap(orc, world):
    """The linear-gap (-4) run of the whole-database cascade: the same
    candidates, scored by linear SW bit-exactly."""
    from sseqs_amd.query import prepare
    from sseqs_amd.search import global_search
    announce("global cascade linear -4")
    db, query = world["db"], world["queries"][0]
    q = prepare(query)
    ga = global_search(db, q, top_prefilter=3000, top_hits=500)
    gl = global_search(db, q, top_prefilter=3000, top_hits=500, linear_gap=-4)
    cand = gl.candidates.cpu().numpy()
    assert np.array_equal(cand, ga.candidates.cpu().numpy())
    qc = orc.encode_query_sw(query)
    sw = orc.sw_linear_segments(qc, world["ascii"], world["off"][cand], world["ln"][cand], -4)
    assert np.array_equal(gl.cand_scores.cpu().numpy(), sw)
    top = orc.topk_desc(sw, 500)
    assert np.array_equal(gl.ids.cpu().numpy(), cand[top])


def mutate(rng, s, sub=0.3, indel=0.03):
    aa = "ARNDCQEGHILKMFPSTWYV"
    out = []
    for c in s:
        r = rng.random()
        if r < indel / 2:
            continue
        if r < indel:
            out.append("".join(rng.choice(list(aa), int(rng.integers(1, 4)))))
        out.append(str(rng.choice(list(aa))) if rng.random() < sub else c)
    return "".join(out) or "A"


def _check_batch_cascade(orc, db, ascii, query, bs, resident, **kw):
    """batch_search against the oracle cascade fed with the fp16 prefilter
    scores batch_search itself used (checked against FP64 first)."""
    from sseqs_amd.query import prepare
    from sseqs_amd.search import batch_search
    q = prepare(query)
    dev_scores = {}
    got = batch_search(db, query, q, bs=bs, resident=resident,
                       on_scores=lambda j, s: dev_scores.__setitem__(j, s.float().cpu().numpy()),
                       **kw)
    want, fft64, tol64 = orc.batch_cascade(ascii, query, bs * 8 // 5,
                                           fft_scores=lambda j: dev_scores[j], with_tol=True,
                                           **kw)
    nonempty = [j for j, w in enumerate(fft64) if w.size]
    assert sorted(dev_scores) == nonempty
    top_fft = kw.get("top_fft", 200)
    for j in nonempty:
        w, tol = fft64[j], tol64[j]
        assert np.all(np.abs(dev_scores[j] - w) <= tol), j
        # each batch's stage-1 passed set (top 1/top_fft of the device's fp16
        # scores) is the FP64-selected set up to threshold ties
        check_selection(orc, dev_scores[j], w, tol, w.size // top_fft)
    assert len(got) == len(want) > 0
    for g, w in zip(got, want):
        assert (np.float32(g.norm), g.affine, g.linear, g.seq) == w
    return got


@pytest.mark.parametrize("resident", [True, False])
def test_batch_cascade(orc, world, resident):
    announce(f"batch cascade resident={resident}")
    bs = 1_000_000  # bytes -> 1.6 M symbols per batch, several batches
    _check_batch_cascade(orc, world["db"], world["ascii"], world["queries"][0], bs, resident,
                         top_fft=50, top_sw=5, top_sw_affine=2)


@pytest.mark.parametrize("bs,qi,top_fft,top_sw", [(250_000, 0, 30, 3), (333_335, 1, 50, 4),
                                                  (99_995, 1, 20, 2)])
def test_batch_cascade_configs(orc, world, bs, qi, top_fft, top_sw):
    """Other batch sizes (batch boundaries at other places, more and smaller
    batches), the other query and other cascade fractions."""
    announce(f"batch cascade bs={bs} q={qi} top_fft={top_fft} top_sw={top_sw}")
    _check_batch_cascade(orc, world["db"], world["ascii"], world["queries"][qi], bs, True,
                         top_fft=top_fft, top_sw=top_sw, top_sw_affine=2)


def test_batch_cascade_unsorted_stream(orc, world):
    """The per-batch cascade on a database file not in length order: every
    batch mixes all lengths, so its prefilter tiles hold sequences of very
    different lengths and every batch takes the largest N."""
    from sseqs_amd import synth
    announce("batch cascade unsorted stream")
    sh, _ = synth.shuffled_copy(world["db"])
    ascii = orc.xbit_unpack(sh.packed.cpu().numpy(), sh.n_syms)
    _check_batch_cascade(orc, sh, ascii, world["queries"][0], 1_000_000, True,
                         top_fft=50, top_sw=5, top_sw_affine=2)


def test_batch_cascade_shifted_segments(orc, dev):
    """Batches that start on '@' or hold '@@': the reference's `lengths != 0`
    filter shifts the offsets of later segments; both paths reproduce it."""
    from sseqs_amd.db import SeqDB
    announce("batch cascade with shifted segments")
    rng = np.random.default_rng(11)
    aa = np.frombuffer(b"ARNDCQEGHILKMFPSTWYV", dtype=np.uint8)
    query = rng.choice(aa, 120).tobytes().decode()
    parts = []
    for i in range(5000):
        if i % 40 == 0:
            s = np.frombuffer(mutate(rng, query, sub=0.4).encode(), dtype=np.uint8)
        else:
            s = rng.choice(aa, int(rng.integers(20, 280)))
        parts.append(s.tobytes())
        parts.append(b"@@@" if rng.random() < 0.01 else b"@")
    a = np.frombuffer(b"".join(parts), dtype=np.uint8).copy()
    bs = 20_000
    sb = bs * 8 // 5
    for b in range(1, a.size // sb, 3):  # some windows start on '@', some on '@@'
        a[b * sb] = 64
        if b % 2:
            a[b * sb + 1] = 64
    packed = torch.from_numpy(orc.xbit_pack(a)).to(dev)
    db = SeqDB.from_packed(packed, a.size)
    plan = db.batch_plan(bs)
    irregular = ~plan.regular & (plan.nseg > 0)
    assert irregular.any() and plan.regular.any()
    # every such batch gets the reference's views and stays on the resident path
    assert plan.shifted[irregular].all()
    kw = dict(top_fft=10, top_sw=3, top_sw_affine=2)
    got_r = _check_batch_cascade(orc, db, a, query, bs, True, **kw)
    got_a = _check_batch_cascade(orc, db, a, query, bs, False, **kw)
    assert len(got_r) == len(got_a)


def test_batch_cascade_long_segments(orc, dev):
    """Batches whose longest segment needs a correlation size above 2048 are
    scored at 2048 on the first 2048 residues (the reference's kernels leave
    such batches at zero); both paths against the oracle."""
    from sseqs_amd.db import SeqDB
    announce("batch cascade with segments longer than 2048")
    rng = np.random.default_rng(12)
    aa = np.frombuffer(b"ARNDCQEGHILKMFPSTWYV", dtype=np.uint8)
    query = rng.choice(aa, 150).tobytes().decode()
    parts = []
    for i in range(1600):
        if i % 50 == 7:
            s = np.frombuffer(mutate(rng, query * 20, sub=0.3).encode(), dtype=np.uint8)
        elif i % 25 == 0:
            s = np.frombuffer(mutate(rng, query, sub=0.3).encode(), dtype=np.uint8)
        else:
            s = rng.choice(aa, int(rng.integers(20, 400)))
        parts.append(s.tobytes())
        parts.append(b"@")
    a = np.frombuffer(b"".join(parts), dtype=np.uint8).copy()
    db = SeqDB.from_packed(torch.from_numpy(orc.xbit_pack(a)).to(dev), a.size)
    bs = 50_000
    plan = db.batch_plan(bs)
    assert plan.maxlen.max() > 2048
    kw = dict(top_fft=10, top_sw=3, top_sw_affine=2)
    _check_batch_cascade(orc, db, a, query, bs, True, **kw)
    _check_batch_cascade(orc, db, a, query, bs, False, **kw)


def test_msa_global_a3m(orc, world, tmp_path):
    import sseqs_amd
    announce("msa global a3m")
    db, query = world["db"], world["queries"][0]
    out = tmp_path / "q.a3m"
    sseqs_amd.msa(query, str(out), num_msas=300, verbose=False, db=db, cascade="global",
                  top_prefilter=2000)
    text = out.read_text()
    from sseqs_amd.query import prepare
    from sseqs_amd.search import global_search
    gh = global_search(db, prepare(query), top_prefilter=2000, top_hits=300)
    seqs = db.strings(gh.ids, 500)
    want = orc.a3m(query, [orc.backtrack(query, s) for s in seqs])
    assert text == want
    assert text.count(">") > 100


def test_msa_batch_a3m(orc, world, tmp_path):
    import sseqs_amd
    from sseqs_amd.query import prepare
    from sseqs_amd.search import batch_search, rank_hits
    announce("msa batch a3m")
    db, query = world["db"], world["queries"][1]
    out = tmp_path / "b.a3m"
    t = sseqs_amd.msa([query], [str(out)], num_msas=64, verbose=False, db=db, bs=500_000,
                      top_fft=20, top_sw=4, return_timings=True)
    assert isinstance(t, dict)
    # the expected side is the oracle's cascade and the oracle's merge
    # (orc.rank_hits), fed the device's fp16 prefilter scores
    dev_scores = {}
    got = batch_search(db, query, prepare(query), bs=500_000, top_fft=20, top_sw=4,
                       on_scores=lambda j, s: dev_scores.__setitem__(j, s.float().cpu().numpy()))
    want_hits, _ = orc.batch_cascade(world["ascii"], query, 500_000 * 8 // 5, top_fft=20,
                                     top_sw=4, fft_scores=lambda j: dev_scores[j])
    ranked = orc.rank_hits(want_hits, 64)
    assert [h.seq for h in rank_hits(got, 64)] == [h[3] for h in ranked]
    want = orc.a3m(query, [orc.backtrack(query, h[3]) for h in ranked])
    assert out.read_text() == want


def test_server_ticket_roundtrip(orc, world, tmp_path, monkeypatch):
    """The ticket route on the real search: the tarball holds msa()'s a3m with
    the request header and '@' -> 'X', as server.py:64-106."""
    import io
    import tarfile

    pytest.importorskip("fastapi")
    from fastapi.testclient import TestClient

    import sseqs_amd
    from sseqs_amd.server import create_app
    announce("server ticket roundtrip")
    db, query = world["db"], world["queries"][1]
    for k, v in (("MSAS_NUM", "64"), ("MSAS_TOP_FFT", "20"), ("MSAS_TOP_SW", "4")):
        monkeypatch.setenv(k, v)
    client = TestClient(create_app(root=tmp_path, chunks=4,
                                   msa_kwargs=dict(db=db, verbose=False)))
    r = client.post("/ticket/msa", data={"q": f">101\n{query}"})
    assert r.status_code == 200 and r.json()["status"] == "COMPLETE"
    blob = client.get(f"/result/download/{r.json()['id']}").content
    with tarfile.open(fileobj=io.BytesIO(blob), mode="r:gz") as tar:
        got = tar.extractfile("uniref.a3m").read().decode()
    direct = tmp_path / "direct.a3m"
    sseqs_amd.msa([query], [str(direct)], num_msas=64, verbose=False, db=db, bs=100_000_000,
                  top_fft=20, top_sw=4)
    rows = direct.read_text().replace("@", "X").split("\n")
    rows[0] = ">101"
    assert got == "\n".join(rows)
    # one batch (bs = 1e8 bytes) keeps top_sw_affine = 2 hits
    assert got.count(">") == 3


def test_msa_long_query_whole_hits(orc, world, tmp_path):
    """A query past the reference's 1024-residue SW limit (prefilter N = 2048,
    stripe width 47) with whole hits instead of 500-residue prefixes."""
    import sseqs_amd
    from sseqs_amd import synth
    from sseqs_amd.query import prepare
    from sseqs_amd.search import global_search
    announce("msa long query")
    db = world["db"]
    query = synth.make_query(1500, 7, seed=4)
    out = tmp_path / "long.a3m"
    sseqs_amd.msa(query, str(out), num_msas=40, verbose=False, db=db, cascade="global",
                  top_prefilter=400, hit_width=None)
    gh = global_search(db, prepare(query), top_prefilter=400, top_hits=40)
    cand = gh.candidates.cpu().numpy()
    qc = orc.encode_query_sw(query)
    sw = orc.sw_affine_segments(qc, world["ascii"], world["off"][cand], world["ln"][cand])
    assert np.array_equal(gh.cand_scores.cpu().numpy(), sw)
    seqs = db.strings(gh.ids, 1024)
    # whole hits: every row is the full database sequence
    assert [len(x) for x in seqs] == db.seq_len[gh.ids].cpu().tolist()
    want = orc.a3m(query, [orc.backtrack(query, s) for s in seqs])
    assert out.read_text() == want


def test_batch_cascade_large_stage2(orc, dev):
    """Short sequences and top_fft = top_sw = 1: the stage-2 selects keep
    about 30 000 and 14 000 targets per batch, past the sorted limits of the
    segmented (4096) and single (16384) selects; both paths against the
    oracle."""
    from sseqs_amd.db import SeqDB
    announce("batch cascade with stage-2 selections above 16384")
    rng = np.random.default_rng(13)
    aa = np.frombuffer(b"ARNDCQEGHILKMFPSTWYV", dtype=np.uint8)
    query = rng.choice(aa, 30).tobytes().decode()
    parts = []
    for i in range(45_000):
        if i % 30 == 0:
            s = np.frombuffer(mutate(rng, query, sub=0.3).encode(), dtype=np.uint8)
        else:
            s = rng.choice(aa, int(rng.integers(8, 41)))
        parts.append(s.tobytes())
        parts.append(b"@")
    a = np.frombuffer(b"".join(parts), dtype=np.uint8).copy()
    db = SeqDB.from_packed(torch.from_numpy(orc.xbit_pack(a)).to(dev), a.size)
    bs = 500_000
    plan = db.batch_plan(bs)
    assert plan.n_batches == 2
    assert plan.nseg[0] - 2 > 16384 and 4096 < plan.nseg[1] - 2 <= 16384
    kw = dict(top_fft=1, top_sw=1, top_sw_affine=2)
    _check_batch_cascade(orc, db, a, query, bs, True, **kw)
    _check_batch_cascade(orc, db, a, query, bs, False, **kw)
