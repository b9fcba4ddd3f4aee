"""The small plain-method workload at its stated size, against the oracle.

One query of length 200 against a 100 000-sequence synthetic database
(log-normal lengths, median 250, clipped to [30, 998]; BLOSUM62 background
residues; 5 % planted homologs of the query at 30-90 % identity), affine SW
11/1, the top 4096 hits backtracked to a3m.  The prefilter passes the top
20 000 to SW.  Checked: every prefilter score within tolerance of the FP64
closed form, the candidate set equal to the FP64 selection up to threshold
ties, every candidate's SW score, the 4096 hits and their order, and the a3m
text that msa() writes, character for character.
"""
import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu

N_SEQ = 100_000
TOP_PRE = 20_000
TOP_HITS = 4096


def test_config1_a3m(orc, dev, tmp_path):
    import sseqs_amd
    from sseqs_amd import synth
    from sseqs_amd.db import SeqDB
    from sseqs_amd.query import prepare
    from sseqs_amd.search import global_search
    from test_gpu_pipeline import _fp64_prefilter, check_selection

    announce(f"config 1: q_len 200 vs {N_SEQ} sequences, top {TOP_HITS} to a3m")
    query = synth.make_query(200, 0, seed=21)
    db = SeqDB.synthetic(N_SEQ, [query], seed=21, device=dev)
    assert db.n == N_SEQ
    ascii = orc.xbit_unpack(db.packed.cpu().numpy(), db.n_syms)
    off, ln = db.seq_off.cpu().numpy(), db.seq_len.cpu().numpy()
    assert ln.min() >= 30 and ln.max() <= 998
    hq = db.extra["hom_query"].cpu().numpy()
    assert abs(np.mean(hq == 0) - 0.05) < 0.001

    gh = global_search(db, prepare(query), top_prefilter=TOP_PRE, top_hits=TOP_HITS)
    pre = gh.prefilter.cpu().numpy()
    want, tol = _fp64_prefilter(orc, ascii, off, ln, query)
    assert np.all(np.abs(pre - want) <= tol)
    cand = gh.candidates.cpu().numpy()
    assert np.array_equal(cand, check_selection(orc, pre, want, tol, TOP_PRE))
    qc = orc.encode_query_sw(query)
    sw = orc.sw_affine_segments(qc, ascii, off[cand], ln[cand])
    assert np.array_equal(gh.cand_scores.cpu().numpy(), sw)
    top = orc.topk_desc(sw, TOP_HITS)
    assert np.array_equal(gh.ids.cpu().numpy(), cand[top])
    assert np.array_equal(gh.scores.cpu().numpy(), sw[top])

    # the whole call: msa() with the global cascade writes the a3m of those
    # hits, each truncated to 500 residues like the reference
    out = tmp_path / "config1.a3m"
    sseqs_amd.msa(query, str(out), num_msas=TOP_HITS, verbose=False, db=db, cascade="global",
                  top_prefilter=TOP_PRE)
    hits = [ascii[off[i]: off[i] + min(int(ln[i]), 500)].tobytes().decode()
            for i in cand[top]]
    text = out.read_text()
    assert text == orc.a3m(query, [orc.backtrack(query, h) for h in hits])
    assert text.count(">") > 1000
    del db
    torch.cuda.empty_cache()
