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
    A, B = orc.weighted_luts(query, 4)
    qrev = orc.encode_query_rev(query)
    out = np.zeros(ln.size)
    Lq = len(query)
    Ns = np.array([orc.nextpow2(max(Lq, int(L), 16)) for L in ln])
    for N in np.unique(Ns):
        sel = np.nonzero(Ns == N)[0]
        out[sel] = orc.corr(qrev, A, B, 4, ascii, off[sel].astype(np.int32), ln[sel], int(N),
                            int(N))
    return out


@pytest.mark.parametrize("qi", [0, 1])
def test_global_cascade(orc, world, qi):
    from sseqs_amd.query import prepare
    from sseqs_amd.search import global_search
    announce(f"global cascade query {qi}")
    db, query = world["db"], world["queries"][qi]
    q = prepare(query)
    gh = global_search(db, q, top_prefilter=3000, top_hits=500)
    pre = gh.prefilter.cpu().numpy()
    want = _fp64_prefilter(orc, world["ascii"], world["off"], world["ln"], query)
    scale = 1e-3 * np.max(np.abs(want))
    assert np.all(np.abs(pre - want) <= 1e-2 * np.abs(want) + scale)
    # candidates: exact top-k of the device scores, and the FP64 top-k up
    # to threshold ties within the tolerance
    cand = gh.candidates.cpu().numpy()
    assert np.array_equal(cand, np.sort(orc.topk_desc(pre.astype(np.float64), 3000)))
    ref = np.sort(orc.topk_desc(want, 3000))
    thr = np.sort(want)[::-1][2999]
    diff = np.setxor1d(cand, ref)
    assert np.all(np.abs(want[diff] - thr) <= 1e-2 * abs(thr) + scale)
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


def test_batch_cascade(orc, world):
    from sseqs_amd.query import prepare
    from sseqs_amd.search import batch_search
    announce("batch cascade")
    db, query = world["db"], world["queries"][0]
    q = prepare(query)
    bs = 1_000_000  # bytes -> 1.6 M symbols per batch, several batches
    got = batch_search(db, query, q, bs=bs, top_fft=50, top_sw=5, top_sw_affine=2)
    # feed the oracle the device's own fp16 prefilter scores batch by batch
    from sseqs_amd import ops
    dev_scores = []
    from sseqs_amd.search import _nextpow2

    def fft_scores(j):
        return dev_scores[j]
    for lo in range(0, db.nbytes, bs):
        n_sym = min(bs * 8 // 5, db.n_syms - lo * 8 // 5)
        a = ops.decompress(db.packed, n_sym, lo)
        an = a.cpu().numpy()
        D, starts, lengths, off = orc.segments(an)
        s = torch.zeros(lengths.size, dtype=torch.float16, device=a.device)
        maxlen = int(lengths.max())
        t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(a.device)
        ops.warp_corr(s, q.qrev, q.lut_a, q.lut_b, a, t(off[:-1].astype(np.int32)),
                      t(lengths.astype(np.int32)), q.k, len(query), maxlen,
                      _nextpow2(max(len(query), maxlen)), torch.zeros(1, dtype=torch.float16,
                                                                       device=a.device))
        dev_scores.append(s.float().cpu().numpy())
    want, fft64 = orc.batch_cascade(world["ascii"], query, bs * 8 // 5, top_fft=50, top_sw=5,
                                    top_sw_affine=2, fft_scores=fft_scores)
    for s, w in zip(dev_scores, fft64):
        scale = 1e-3 * np.max(np.abs(w))
        assert np.all(np.abs(s - w) <= 1e-2 * np.abs(w) + scale)
    assert len(got) == len(want) > 0
    for g, w in zip(got, want):
        assert (np.float32(g.norm), g.affine, g.linear, g.seq) == w


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
    hits = rank_hits(batch_search(db, query, prepare(query), bs=500_000, top_fft=20, top_sw=4),
                     64)
    want = orc.a3m(query, [orc.backtrack(query, h.seq) for h in hits])
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
