"""The learned prefilter's HIP kernels against a plain PyTorch fp32 reference
of the same operations (on the CPU), and the proposal against its
definition.

* forward: embeddings of sequences of mixed lengths (with non-residue
  letters) equal the reference that rounds weights and stored activations
  to bf16 where the kernels do, within 2e-3 per component of the unit
  vectors (fp32 accumulation in another order, tanh of another library);
* backward: the gradients of a pair loss with respect to every parameter
  block equal autograd's of the fp32 model within the error that bf16
  operands bring (relative L2 error below 5e-2, cosine above 0.998);
* the weighted pair loss and the soft-target InfoNCE loss (value and
  gradient) equal fp64 autograd of their formulas;
* the training step moves the loss down, with either loss;
* the FP8 table holds e4m3 of 16 * e, and rows built in chunks equal the
  same sequences embedded in one batch, bit for bit;
* proposal scores equal the fp32 dot products of the fp8 operands within
  fp16 rounding, a query's score is the exact maximum over its window rows,
  and the proposed candidates are the exact top-k of those scores; the
  learned cascade's SW scores are the oracle's;
* calibration recall equals the share of the oracle's exhaustive top k
  among the top K proposals, and a failed gate hands selection back to the
  plain prefilter.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from conftest import announce

pytestmark = pytest.mark.gpu

AA = "ARNDCQEGHILKMFPSTWYV"


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def ref_forward(params, seqs, round_bf16=True):
    """Encoder in fp32 on the CPU; round_bf16 rounds the weights, the
    embedding rows and every stored activation to bf16 like the kernels."""
    from sseqs_amd.learned import codes_of
    r = _bf if round_bf16 else (lambda t: t)
    E = params["E"]
    out = []
    for s in seqs:
        c = torch.from_numpy(codes_of(s).astype(np.int64))
        x = r(E[c]).T.unsqueeze(0)  # (1, 64, L)
        for wn, bn in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
            W = r(params[wn]).permute(2, 1, 0)  # (Cout, Cin, 9)
            u = F.conv1d(x, W, padding=4) + params[bn][None, :, None]
            x = r(F.gelu(u, approximate="tanh"))
        z = x[0].mean(dim=1)
        n = z.norm()
        out.append(z / n if n > 0 else torch.zeros_like(z))
    return torch.stack(out)


def _seqs(rng, n, lo=1, hi=300):
    out = []
    for _ in range(n):
        L = int(rng.integers(lo, hi + 1))
        s = "".join(rng.choice(list(AA + "XB"), L, p=[0.048] * 20 + [0.02, 0.02]))
        out.append(s)
    return out


@pytest.fixture(scope="module")
def enc(dev):
    from sseqs_amd.learned import Encoder
    return Encoder(dev, seed=3)


def _params_cpu(enc):
    return {k: torch.from_numpy(v.copy()) for k, v in enc.state().items()}


def test_encoder_forward_matches_reference(enc):
    announce("encoder forward vs fp32 reference")
    rng = np.random.default_rng(1)
    seqs = _seqs(rng, 40) + ["W", "XXXX", "ACDEFGHIKLMNPQRSTVWY" * 20]
    got = enc.embed_strings(seqs).cpu()
    want = ref_forward(_params_cpu(enc), seqs)
    err = (got - want).abs().max().item()
    assert err < 2e-3, err
    # unit vectors, except the sequence without residues (zero embedding)
    has = torch.tensor([any(c in AA for c in s) for s in seqs])
    assert torch.allclose(got.norm(dim=1)[has], torch.ones(int(has.sum())), atol=1e-4)
    assert got[~has].abs().max().item() == 0.0


def test_encoder_backward_matches_autograd(enc, dev):
    from sseqs_amd._lib import call, ptr, stream_of
    from sseqs_amd.learned import Encoder
    announce("encoder backward vs autograd")
    rng = np.random.default_rng(2)
    seqs = _seqs(rng, 24, 20, 250)
    nq = 6
    qi = np.repeat(np.arange(nq), 3).astype(np.int32)
    ti = (nq + np.arange(18)).astype(np.int32)
    y = rng.uniform(0.0, 0.9, 18).astype(np.float32)
    e2 = Encoder(dev, weights=enc.state())
    code, mask, rows = e2.rows_of_strings(seqs)
    e, znorm, buf = e2.forward(code, mask, rows, train=True)
    de = torch.zeros_like(e)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    # device copies held in names until the launch has been enqueued
    qi_d, ti_d, y_d = (torch.from_numpy(x).to(dev) for x in (qi, ti, y))
    call("sseqs_enc_pair_loss", ptr(e), int(e.shape[0]), ptr(qi_d), ptr(ti_d), ptr(y_d), None,
         18, ptr(de), None, ptr(loss), stream_of(dev))
    e2.backward(de, e, znorm, code, mask, rows, buf)
    torch.cuda.synchronize()
    params = {k: v.clone().requires_grad_(True) for k, v in _params_cpu(enc).items()}
    emb = ref_forward(params, seqs, round_bf16=False)
    cos = (emb[torch.from_numpy(qi).long()] * emb[torch.from_numpy(ti).long()]).sum(1)
    ref_loss = ((cos - torch.from_numpy(y)) ** 2).mean()
    ref_loss.backward()
    assert abs(loss.item() - ref_loss.item()) < 2e-3 * max(1.0, ref_loss.item())
    for name in ("E", "W1", "b1", "W2", "b2", "W3", "b3"):
        g = e2.view(e2.g, name).cpu().flatten()
        w = params[name].grad.flatten()
        if name == "E":
            g, w = g[: 20 * 64], w[: 20 * 64]
        rel = ((g - w).norm() / w.norm()).item()
        cosg = (g @ w / (g.norm() * w.norm())).item()
        assert rel < 5e-2 and cosg > 0.998, (name, rel, cosg)


def test_pair_and_infonce_losses_match_reference(dev):
    """Weighted pair loss and soft-target InfoNCE on random unit rows against
    autograd of the same formulas in fp64 on the CPU."""
    from sseqs_amd._lib import call, ptr, stream_of
    announce("weighted pair loss and InfoNCE vs fp64 autograd")
    rng = np.random.default_rng(4)
    nq, nt, tau = 7, 300, 0.07
    e64 = torch.from_numpy(rng.standard_normal((nq + nt, 256)))
    e64 = e64 / e64.norm(dim=1, keepdim=True)
    P = rng.uniform(0, 1, (nq, nt)) * (rng.uniform(0, 1, (nq, nt)) < 0.1)
    P[2] = 0.0  # a query without a distribution takes no part
    P = P / np.maximum(P.sum(1, keepdims=True), 1e-30)
    qi = rng.integers(0, nq, 50).astype(np.int32)
    ti = (nq + rng.integers(0, nt, 50)).astype(np.int32)
    y = rng.uniform(0, 1, 50).astype(np.float32)
    w = rng.uniform(0.5, 4, 50).astype(np.float32)
    active = int((P.sum(1) > 0).sum())
    scale = 0.6 / active
    # held in names until the launches have been enqueued
    e = e64.float().to(dev)
    Pd = torch.from_numpy(P.astype(np.float32)).to(dev)
    qd, td, yd, wd = (torch.from_numpy(x).to(dev) for x in (qi, ti, y, w))
    G = torch.empty((nq, nt), dtype=torch.float32, device=dev)
    de = torch.zeros_like(e)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    call("sseqs_enc_pair_loss", ptr(e), nq + nt, ptr(qd), ptr(td), ptr(yd), ptr(wd), 50, ptr(de),
         None, ptr(loss), stream_of(dev))
    call("sseqs_enc_infonce", ptr(e), nq, nt, ptr(Pd), tau, scale, ptr(G), ptr(de), ptr(loss),
         stream_of(dev))
    torch.cuda.synchronize()
    x = e.double().cpu().requires_grad_(True)
    cos = (x[torch.from_numpy(qi).long()] * x[torch.from_numpy(ti).long()]).sum(1)
    mse = (torch.from_numpy(w).double() * (cos - torch.from_numpy(y).double()) ** 2).mean()
    logp = torch.log_softmax(x[:nq] @ x[nq:].T / tau, dim=1)
    nce = -scale * (torch.from_numpy(P) * logp).sum()
    (mse + nce).backward()
    want = (mse + nce).item()
    assert abs(loss.item() - want) < 1e-4 * max(1.0, abs(want)), (loss.item(), want)
    err = (de.cpu().double() - x.grad).abs().max().item()
    assert err < 1e-4 * x.grad.abs().max().item() + 1e-6, err
    assert G[2].abs().max().item() == 0.0


def test_training_lowers_the_loss(dev):
    from sseqs_amd.learned import Encoder, make_training_set, train
    announce("encoder training on synthetic pairs")
    ts = make_training_set(60_000, 64, 64, seed=5, device=dev, lo=50, hi=300)
    assert ts.y.min() >= 0.0 and ts.y.max() <= 1.0
    hom = ts.y > 0.3
    assert hom.any() and (~hom).any()
    enc = Encoder(dev, seed=4)
    losses = train(enc, ts, steps=120, queries_per_step=16, pairs_per_query=16, lr=2e-3)
    assert np.mean(losses[-2:]) < 0.7 * losses[0], losses
    enc = Encoder(dev, seed=4)
    losses = train(enc, ts, steps=120, queries_per_step=16, pairs_per_query=16, lr=2e-3,
                   mse_weight=0.0, nce_weight=1.0)
    assert np.mean(losses[-2:]) < 0.8 * losses[0], losses


def test_table_and_proposal(orc, enc, dev):
    from sseqs_amd import synth
    from sseqs_amd._lib import call, ptr, stream_of
    from sseqs_amd.db import SeqDB
    from sseqs_amd.learned import FP8_SCALE, EmbeddingTable, Proposer
    from sseqs_amd.multi import MultiSearch
    announce("fp8 table and MFMA proposal")
    queries = [synth.make_query(L, 90 + i, seed=31) for i, L in enumerate((80, 230, 510))]
    db = SeqDB.synthetic(12000, queries, seed=32, device=dev)
    tab = EmbeddingTable(enc, db, rows=400_000)  # several chunks
    ids = np.arange(0, db.n, 37)
    seqs = db.strings(torch.from_numpy(ids).to(dev), 2048)
    e_one = enc.embed_strings(seqs)
    q8 = torch.empty((ids.size, 256), dtype=torch.uint8, device=dev)
    call("sseqs_enc_to_fp8", ptr(e_one), e_one.numel(), FP8_SCALE, ptr(q8), stream_of(dev))
    assert torch.equal(tab.table[torch.from_numpy(ids).to(dev)], q8)
    # e4m3 (OCP) of 16 * e
    dec = tab.table[:500].cpu().view(torch.float8_e4m3fn).float()
    want = (FP8_SCALE * enc.embed_strings(db.strings(torch.arange(500, device=dev), 2048))).cpu()
    assert torch.allclose(dec, want.to(torch.float8_e4m3fn).float(), atol=0, rtol=0)
    # proposal scores and candidates: every embedding row of the queries (the
    # 510-residue query adds two 384-residue windows), and a query's score
    # is the maximum over its rows
    prop16 = Proposer(tab, group=16)
    rows, count = prop16.rows_of(queries)
    assert list(count) == [1, 1, 3] and rows[3] == queries[2][:384] and rows[4] == queries[2][-384:]
    ms = MultiSearch(db, queries, top_prefilter=600, top_hits=50, proposer=prop16)
    ms.run()
    qe = enc.embed_strings(rows)
    q8r = torch.zeros((16, 256), dtype=torch.uint8, device=dev)
    call("sseqs_enc_to_fp8", ptr(qe), qe.numel(), FP8_SCALE, ptr(q8r), stream_of(dev))
    raw_rows = torch.empty((16, db.n), dtype=torch.float16, device=dev)
    call("sseqs_propose", ptr(tab.table), db.n, ptr(q8r), 16, ptr(raw_rows), db.n, stream_of(dev))
    sc_rows = raw_rows[:5].float().cpu()
    qq = (FP8_SCALE * qe).cpu().to(torch.float8_e4m3fn).float()
    tt = tab.table.cpu().view(torch.float8_e4m3fn).float()
    ref = qq @ tt.T
    assert torch.allclose(sc_rows, ref.half().float(), rtol=2e-3, atol=2e-3)
    sc = torch.stack([sc_rows[0], sc_rows[1], sc_rows[2:5].max(dim=0).values])
    assert torch.equal(prop16.scores_for(queries).float().cpu(), sc)
    for i in range(3):
        cand = ms.cand[i * 600:(i + 1) * 600].cpu().numpy()
        top = orc.topk_desc(sc[i].numpy().astype(np.float64), 600)
        assert np.array_equal(cand, np.sort(top)), i
    # exact SW of the proposed candidates
    ascii = orc.xbit_unpack(db.packed.cpu().numpy(), db.n_syms)
    off, ln = db.seq_off.cpu().numpy(), db.seq_len.cpu().numpy()
    for i, q in enumerate(queries):
        cand = ms.cand[i * 600:(i + 1) * 600].cpu().numpy()
        sw = orc.sw_affine_segments(orc.encode_query_sw(q), ascii, off[cand], ln[cand])
        assert np.array_equal(ms.cs[i * 600:(i + 1) * 600].cpu().numpy(), sw)
    # a table whose length is not a multiple of 8 (nor of the kernel's 128
    # rows), with the padded stride and with an unpadded one (scalar stores)
    db2 = SeqDB.synthetic(1003, queries, seed=33, device=dev)
    tab2 = EmbeddingTable(enc, db2)
    s_pad = Proposer(tab2).scores_for(queries).float().cpu()
    tt2 = tab2.table.cpu().view(torch.float8_e4m3fn).float()
    r2 = (qq @ tt2.T).half().float()
    want2 = torch.stack([r2[0], r2[1], r2[2:5].max(dim=0).values])
    assert torch.allclose(s_pad, want2, rtol=2e-3, atol=2e-3)
    s_whole = Proposer(tab2, window=0).scores_for(queries).float().cpu()
    qw = enc.embed_strings(queries)
    q8 = torch.zeros((16, 256), dtype=torch.uint8, device=dev)
    call("sseqs_enc_to_fp8", ptr(qw), qw.numel(), FP8_SCALE, ptr(q8), stream_of(dev))
    raw = torch.full((16, 1003), float("nan"), dtype=torch.float16, device=dev)
    call("sseqs_propose", ptr(tab2.table), 1003, ptr(q8), 16, ptr(raw), 1003, stream_of(dev))
    assert torch.equal(raw[:3].float().cpu(), s_whole)
    # the row maximum against torch's, on a stride padded to 8 with a tail
    g = torch.randn((5, 1008), device=dev).half()
    mx = torch.empty(1008, dtype=torch.float16, device=dev)
    call("sseqs_rows_max", ptr(g), 1008, 5, 1003, ptr(mx), stream_of(dev))
    assert torch.equal(mx[:1003], g[:, :1003].max(dim=0).values)
    # calibration against exhaustive SW (oracle scores of every sequence):
    # recall at K is the share of the exhaustive top k among the top K
    # proposals, and the gate decides between the learned and plain selection
    prop = ms.proposer
    k, K = 40, 600
    cal = prop.calibrate(queries, top_prefilter=K, k=k)
    for i, q in enumerate(queries):
        allsc = orc.sw_affine_segments(orc.encode_query_sw(q), ascii, off, ln)
        kth = np.sort(allsc)[-k]
        top = orc.topk_desc(sc[i].numpy().astype(np.float64), K)
        rec = min(k, int((allsc[top] >= kth).sum())) / k
        assert cal["per_query"][i]["kth_exhaustive_score"] == kth
        assert abs(cal["per_query"][i]["recall"] - round(rec, 4)) < 1e-9, i
    assert prop.use_learned == (min(p["recall"] for p in cal["per_query"]) >= 0.99)
    # the calibrated operating point gives each calibration query at least
    # the count it needed for 0.99
    Ks = prop.top_prefilter_for(queries)
    for i, pq in enumerate(cal["per_query"]):
        assert Ks[i] >= min(db.n, max(k, pq["top_prefilter_for_gate"])), i
    from sseqs_amd.serve import SearchEngine
    with pytest.raises(ValueError):
        SearchEngine(db, top_prefilter="calibrated")
    cal = prop.calibrate(queries, top_prefilter=db.n, k=k)
    assert prop.use_learned and all(p["recall"] == 1.0 for p in cal["per_query"])
    cal1 = prop.calibrate(queries, top_prefilter=1, k=k)
    assert not prop.use_learned
    # counts below each query's calibrated point fall back to the plain
    # prefilter, query by query; counts at it take the learned proposal
    Kq = prop.top_prefilter_for(queries)
    assert np.array_equal(prop.meets_gate(queries, Kq), np.ones(3, bool))
    kc = np.array([600, int(Kq[1]), 600])
    want_ok = kc >= Kq
    assert np.array_equal(prop.meets_gate(queries, kc), want_ok)
    mixed = MultiSearch(db, queries, kc, 40, proposer=prop)
    mixed.run()
    plain = MultiSearch(db, queries, kc, 40)
    plain.run()
    learned_only = MultiSearch(db, queries, kc, 40, proposer=Proposer(tab, group=16))
    learned_only.run()
    for i in range(3):
        c = slice(int(mixed.c_off_h[i]), int(mixed.c_off_h[i + 1]))
        ref = learned_only if want_ok[i] else plain
        assert torch.equal(mixed.cand[c], ref.cand[c]), i
        assert torch.equal(mixed.cs[c], ref.cs[c]), i
    assert cal1["k_by_len"] == cal["k_by_len"]


def test_fused_proposal_max_equals_rows_max(dev):
    """sseqs_propose_max against sseqs_propose + sseqs_rows_max, bit for bit:
    runs of 1 to 7 query rows with padding rows before and after them, on a
    table of 2.2 M rows (more than one table step per block) and on one of
    1003 rows (ragged tail, unpadded stride with scalar stores)."""
    from sseqs_amd._lib import call, ptr, stream_of
    announce("fused proposal maximum vs proposal + row maximum")
    st = stream_of(dev)
    g = torch.Generator(device="cpu").manual_seed(41)
    for n, G, runs in ((2_200_003, 64, [1, 6, 1, 1, 5, 2, 7, 3, 1, 1, 4, 6, 1, 2, 3, 2, 1, 1, 1,
                                         1, 3, 5]),
                       (1003, 48, [1] * 20 + [2] * 6 + [3, 4, 1, 1, 1, 1])):
        table = (torch.randn((n, 256), generator=g) * 0.08).to(torch.float8_e4m3fn)
        table = table.view(torch.uint8).to(dev)
        q = (torch.randn((G, 256), generator=g) * 0.08).to(torch.float8_e4m3fn)
        q8 = q.view(torch.uint8).to(dev)
        off = np.concatenate([[0], np.cumsum(runs)]).astype(np.int32) + 1  # row 0 is padding
        assert off[-1] < G and len(runs) <= 32
        ld = n if n < 10_000 else (n + 7) // 8 * 8
        raw = torch.empty((G, ld), dtype=torch.float16, device=dev)
        call("sseqs_propose", ptr(table), n, ptr(q8), G, ptr(raw), ld, st)
        fused = torch.full((len(runs) + 1, ld), 7.0, dtype=torch.float16, device=dev)
        call("sseqs_propose_max", ptr(table), n, ptr(q8), G, off.ctypes.data, len(runs),
             ptr(fused), ld, st)
        ld8 = (n + 7) // 8 * 8
        for o in range(len(runs)):
            src = raw[off[o]: off[o + 1]]
            src = src.contiguous() if ld == ld8 else torch.nn.functional.pad(src, (0, ld8 - ld))
            want = torch.empty(ld8, dtype=torch.float16, device=dev)
            call("sseqs_rows_max", ptr(src), ld8, int(runs[o]), n, ptr(want), st)
            assert torch.equal(fused[o, :n].view(torch.int16), want[:n].view(torch.int16)), (n, o)
        assert torch.all(fused[len(runs)] == 7.0)  # rows past the output rows untouched
        del table, raw, fused
