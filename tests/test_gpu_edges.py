"""Degenerate inputs through the operator mirror and the Python API, on the
GPU: empty inputs, selections of none and of every score, targets and
queries without residues, compared with the oracle where it defines them."""
import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu


def test_codec_empty(dev):
    from sseqs_amd import ops
    announce("codec empty")
    assert ops.compress(torch.empty(0, dtype=torch.uint8, device=dev)).numel() == 0
    assert ops.decompress(torch.zeros(5, dtype=torch.uint8, device=dev), 0, 0).numel() == 0


@pytest.mark.parametrize("dtype", [torch.float32, torch.float16, torch.int32])
def test_topk_none_and_all(dev, dtype):
    """k = 0 gives nothing; k = n gives every index, in (score desc, index
    asc) order when sorted and in index order otherwise, with many ties."""
    from sseqs_amd import ops
    announce(f"topk none/all {dtype}")
    x = np.random.default_rng(4).integers(-50, 50, 3001)
    s = torch.tensor(x, dtype=dtype, device=dev)
    assert ops.topk_indices(s, 0).numel() == 0
    got = ops.topk_indices(s, x.size, sorted=True).cpu().numpy()
    assert np.array_equal(got, np.lexsort((np.arange(x.size), -x)))
    got = ops.topk_indices(s, x.size, sorted=False).cpu().numpy()
    assert np.array_equal(got, np.arange(x.size))


def test_sw_operators_degenerate(orc, dev):
    """No targets; a query without residues (every target scores -1,
    sseqs/sw.cu:73-77); targets of non-residues only, an empty target
    between two delimiters and a target at index 0."""
    from sseqs_amd import ops
    announce("sw operators degenerate")
    a = np.frombuffer(b"ACDEWY@XXXX@@BZJ@MKVLLA@", dtype=np.uint8).copy()
    D = np.nonzero(a == 64)[0].astype(np.int32)
    qc = orc.encode_query_sw("MKVLLAACDE")
    at, Dt = torch.from_numpy(a).to(dev), torch.from_numpy(D).to(dev)
    qt = torch.from_numpy(qc).to(dev)
    none = torch.empty(0, dtype=torch.int32, device=dev)
    assert ops.sw_cuda_blosum62(qt, none, at, Dt).numel() == 0
    assert ops.sw_cuda_affine(qt, none, at, Dt).numel() == 0
    gi = np.array([0, 1, 2, 3, 4], dtype=np.int32)
    git = torch.from_numpy(gi).to(dev)
    for gap in (-1, -4):
        got = ops.sw_cuda_blosum62(qt, git, at, Dt, gap_penalty_param=gap).cpu().numpy()
        assert np.array_equal(got, orc.sw_linear(qc, gi, a, D, gap)), gap
    starts = (D + 1)[D + 1 < a.size].astype(np.int32)
    got = ops.sw_cuda_affine(qt, git[:4], at, torch.from_numpy(starts).to(dev)).cpu().numpy()
    assert np.array_equal(got, orc.sw_affine(qc, gi[:4], a, starts, 11, 1))
    q0 = torch.empty(0, dtype=torch.uint8, device=dev)
    assert ops.sw_cuda_blosum62(q0, git, at, Dt).tolist() == [-1] * gi.size
    assert ops.sw_cuda_affine(q0, git, at, Dt).tolist() == [-1] * gi.size


def test_batch_cascade_query_without_residues(orc, dev, tmp_path):
    """A query of letters outside the 20 residues: the prefilter scores it
    through the unknown-residue code and both SW stages score -1, so the
    hits are whatever the selections keep, as in the reference."""
    from sseqs_amd import synth
    from sseqs_amd.db import SeqDB
    from sseqs_amd.query import prepare
    from sseqs_amd.search import batch_search
    announce("batch cascade query without residues")
    db = SeqDB.synthetic(4000, [synth.make_query(80, 0, seed=7)], seed=8, device=dev)
    ascii = orc.xbit_unpack(db.packed.cpu().numpy(), db.n_syms)
    query = "XBZJ" * 12
    dev_scores = {}
    got = batch_search(db, query, prepare(query), bs=150_000, top_fft=20, top_sw=4,
                       on_scores=lambda j, s: dev_scores.__setitem__(j, s.float().cpu().numpy()))
    want, fft64, tol64 = orc.batch_cascade(ascii, query, 150_000 * 8 // 5, top_fft=20,
                                           top_sw=4, fft_scores=lambda j: dev_scores[j],
                                           with_tol=True)
    for j, w in enumerate(fft64):
        if w.size:
            assert np.all(np.abs(dev_scores[j] - w) <= tol64[j]), j
    assert len(got) == len(want) > 0
    for g, w in zip(got, want):
        assert (np.float32(g.norm), g.affine, g.linear, g.seq) == w
        assert g.affine == -1 and g.linear == -1
    # and the whole call: merge, backtrack (non-residues score -5) and a3m
    import sseqs_amd
    from sseqs_amd.search import rank_hits
    out = tmp_path / "x.a3m"
    sseqs_amd.msa(query, str(out), num_msas=64, verbose=False, db=db, bs=150_000, top_fft=20,
                  top_sw=4)
    ranked = orc.rank_hits(want, 64)
    assert [h.seq for h in rank_hits(got, 64)] == [h[3] for h in ranked]
    assert out.read_text() == orc.a3m(query, [orc.backtrack(query, h[3]) for h in ranked])


def test_msa_call_options(dev, tmp_path, monkeypatch):
    """msa()'s other call forms (sseqs/msa.py:71-115): a list of queries with
    a list of files gives each query's single-call text; blocking=False
    returns a thread whose file matches; filenames=None writes
    sha256(query)[:8] + '.a3m'; save_raw keeps the raw hits."""
    import hashlib
    import pickle
    import threading

    import sseqs_amd
    from sseqs_amd import synth
    from sseqs_amd.db import SeqDB
    announce("msa call options")
    qs = [synth.make_query(90, 0, seed=5), synth.make_query(60, 1, seed=5)]
    db = SeqDB.synthetic(6000, qs, seed=6, device=dev)
    kw = dict(num_msas=32, verbose=False, db=db, bs=200_000, top_fft=20, top_sw=4)
    single = []
    for i, q in enumerate(qs):
        fn = tmp_path / f"single{i}.a3m"
        sseqs_amd.msa(q, str(fn), **kw)
        single.append(fn.read_text())
    pair = [tmp_path / "p0.a3m", tmp_path / "p1.a3m"]
    sseqs_amd.msa(qs, [str(p) for p in pair], **kw)
    assert [p.read_text() for p in pair] == single
    t = sseqs_amd.msa(qs[0], str(tmp_path / "bg.a3m"), blocking=False, **kw)
    assert isinstance(t, threading.Thread)
    t.join()
    assert (tmp_path / "bg.a3m").read_text() == single[0]
    monkeypatch.chdir(tmp_path)
    sseqs_amd.msa(qs[1], save_raw=str(tmp_path / "raw.pkl"), **kw)
    assert (tmp_path / (hashlib.sha256(qs[1].encode()).hexdigest()[:8] + ".a3m")).read_text() \
        == single[1]
    with open(tmp_path / "raw.pkl", "rb") as fh:  # written by msa() above
        raw = pickle.load(fh)
    assert raw[0] == qs[1] and len(raw[1]) == len(raw[2]) == len(raw[4]) > 0
