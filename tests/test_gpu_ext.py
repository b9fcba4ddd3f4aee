"""Calls through the compiled PyTorch extension against the ctypes
operators, and the unchanged reference call ``sseqs.msa(query, "out.a3m")``
on a database file named by DBPATH."""
import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu


def test_extension_ops_equal_ctypes_ops(orc, dev):
    from sseqs_amd import ext, ops
    from sseqs_amd.query import prepare
    announce("torch extension ops")
    rng = np.random.default_rng(3)
    aa = np.frombuffer(b"ARNDCQEGHILKMFPSTWYV", dtype=np.uint8)
    segs = [rng.choice(aa, int(rng.integers(5, 300))).tobytes() for _ in range(200)]
    flat = np.frombuffer(b"@".join(segs) + b"@", dtype=np.uint8).copy()
    a = torch.from_numpy(flat).to(dev)
    packed = ext.xbit.compress(a)
    assert torch.equal(packed, ops.xbit.compress(a))
    assert torch.equal(ext.xbit.decompress(packed, flat.size, 0), a)
    query = rng.choice(aa, 150).tobytes().decode()
    q = prepare(query)
    D = torch.nonzero(a == 64).flatten().to(torch.int32)
    gi = torch.arange(1, 150, dtype=torch.int32, device=dev)
    for gap in (-1, -4):
        assert torch.equal(ext.sw.sw_cuda_blosum62(q.codes, gi, a, D, gap_penalty_param=gap),
                           ops.sw.sw_cuda_blosum62(q.codes, gi, a, D, gap_penalty_param=gap))
    starts = (D + 1)[(D + 1) < a.numel()].to(torch.int32)
    assert torch.equal(ext.sw.sw_cuda_affine(q.codes, gi, a, starts),
                       ops.sw.sw_cuda_affine(q.codes, gi, a, starts))
    lens = torch.tensor([len(s) for s in segs], dtype=torch.int32, device=dev)
    off = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev),
                     torch.cumsum(lens.long() + 1, 0)[:-1]]).to(torch.int32)
    s1 = torch.zeros(200, dtype=torch.float16, device=dev)
    s2 = torch.zeros(200, dtype=torch.float16, device=dev)
    scratch = torch.zeros(1, dtype=torch.float16, device=dev)
    ext.conv.warp_corr(s1, q.qrev, q.lut_a, q.lut_b, a, off, lens, 4, 150, 299, 512, scratch)
    ops.conv.warp_corr(s2, q.qrev, q.lut_a, q.lut_b, a, off, lens, 4, 150, 299, 512, scratch)
    assert torch.equal(s1, s2) and s1.abs().sum().item() > 0


def test_sseqs_msa_unchanged_call(dev, tmp_path, monkeypatch):
    import sseqs
    import sseqs_amd
    from sseqs_amd import synth
    from sseqs_amd.db import SeqDB, set_default_db
    announce("sseqs.msa(query, 'out.a3m') on DBPATH")
    query = synth.make_query(140, 0, seed=41)
    db = SeqDB.synthetic(8000, [query], seed=41, device=dev)
    path = tmp_path / "db.xbit"
    path.write_bytes(db.packed[: db.nbytes].cpu().numpy().tobytes())
    monkeypatch.setenv("DBPATH", str(path))
    monkeypatch.setenv("CHUNKS", "4")
    monkeypatch.chdir(tmp_path)
    set_default_db(None)
    try:
        sseqs.msa(query, "out.a3m")  # the reference's call, unchanged
        got = (tmp_path / "out.a3m").read_text()
    finally:
        set_default_db(None)
    sseqs_amd.msa(query, str(tmp_path / "direct.a3m"), db=db, verbose=False)
    assert got == (tmp_path / "direct.a3m").read_text()
    assert got.count(">") >= 2
