"""GPU parity of every kernel against the CPU oracle, through the C ABI.

Integer work (codec, index, SW, top-k, traceback, generator) must be
bit-exact.  The prefilter is floating point: the tolerance, against the
FP64 closed form, is |x - x64| <= 1e-2 * |x64| + the error bound of fp16
operands with fp32 accumulation (orc.corr_tol; the second term matters for
scores near zero; fp16 outputs add at most 2^-11 relative).
"""
import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu

AA = "ARNDCQEGHILKMFPSTWYV"
# what a packed database can hold: residues (weighted up) and other letters
DB_ALPH = AA * 4 + "XBZUOJ"


def rand_seq(rng, n, alphabet=AA):
    return "".join(rng.choice(list(alphabet), int(n)))


def mutate(rng, s, sub=0.3, indel=0.03):
    out = []
    for c in s:
        r = rng.random()
        if r < indel / 2:
            continue
        if r < indel:
            out.append(rand_seq(rng, rng.integers(1, 4)))
        out.append(rand_seq(rng, 1) if rng.random() < sub else c)
    return "".join(out) or "A"


def corr_ok(got, want, tol):
    """|x - x64| <= tol elementwise, tol from orc.corr_tol: 1e-2 relative plus
    the error bound of fp16 operands with fp32 accumulation."""
    return np.abs(got - want) <= tol


# ----------------------------------------------------------------- codec --

@pytest.mark.parametrize("n", [1, 7, 8, 9, 31, 32, 33, 69, 1000, 123457])
def test_xbit_roundtrip(orc, dev, n):
    from sseqs_amd import ops
    announce(f"xbit n={n}")
    rng = np.random.default_rng(n)
    s = (rng.integers(0, 32, n) + 64).astype(np.uint8)
    packed = ops.compress(torch.from_numpy(s).to(dev))
    assert np.array_equal(packed.cpu().numpy(), orc.xbit_pack(s))
    assert np.array_equal(ops.decompress(packed, n, 0).cpu().numpy(), s)
    # byte offsets 5 and 15 are not 4-aligned (byte-wise kernel), 20 and 40 are
    for g in (1, 3, 4, 8):
        if 8 * g < n:
            out = ops.decompress(packed, n - 8 * g, 5 * g).cpu().numpy()
            assert np.array_equal(out, s[8 * g:])


@pytest.mark.parametrize("n", [40, 4099, 70001])
def test_xbit_any_bytes_and_alignment(orc, dev, n):
    """Bytes outside '@'..'_' keep their low 5 bits of (b - 64), as in
    sseqs/xbit.cu:36-62; inputs at odd addresses take the byte-wise kernel."""
    from sseqs_amd import ops
    announce(f"xbit any bytes n={n}")
    rng = np.random.default_rng(n + 1)
    s = rng.integers(0, 256, n + 3).astype(np.uint8)
    t = torch.from_numpy(s).to(dev)
    for shift in (0, 1, 3):
        packed = ops.compress(t[shift:shift + n])
        assert np.array_equal(packed.cpu().numpy(), orc.xbit_pack(s[shift:shift + n])), shift
    back = ops.decompress(ops.compress(t[:n]), n, 0).cpu().numpy()
    assert np.array_equal(back, ((s[:n].astype(np.int32) - 64) & 31) + 64)


def test_xbit_host_input_moves_to_gpu(dev):
    from sseqs_amd import ops
    s = torch.tensor(list(b"HELLO@WORLD@"), dtype=torch.uint8)
    packed = ops.compress(s)
    assert packed.is_cuda
    assert bytes(ops.decompress(packed, 12, 0).cpu().numpy()) == b"HELLO@WORLD@"


def _packed_db(dev, text: bytes):
    from sseqs_amd import ops
    t = torch.from_numpy(np.frombuffer(text, dtype=np.uint8).copy()).to(dev)
    return ops.compress(t)


def test_from_xbit_files(dev, tmp_path):
    """A database written as .xbit files (a main file cut at a chunk limit plus
    an antibody-style second file, sseqs/msa.py:39-40,50-62) reads back into
    the same index."""
    from sseqs_amd import synth
    from sseqs_amd.db import SeqDB
    announce("from_xbit files")
    db = SeqDB.synthetic(6000, [synth.make_query(100, 0, seed=2)], seed=5, device=dev)
    raw = db.packed[:(db.n_syms * 5 + 7) // 8].cpu().numpy()
    cb = 4000  # chunk size of this test; 3 chunks = 12000 bytes, a symbol boundary
    main, ab = tmp_path / "main.xbit", tmp_path / "ab.xbit"
    main.write_bytes(raw[:3 * cb + 123].tobytes())  # bytes past the chunk limit are ignored
    ab.write_bytes(raw[3 * cb:].tobytes())
    got = SeqDB.from_xbit(str(main), chunks=3, device=dev, ab_path=str(ab), ab_chunks=None,
                          chunk_bytes=cb)
    assert got.n == db.n
    assert torch.equal(got.seq_off, db.seq_off) and torch.equal(got.seq_len, db.seq_len)
    assert got.strings(got.seq_off.new_tensor([0, got.n - 1]).long(), 1000) == \
        db.strings(db.seq_off.new_tensor([0, db.n - 1]).long(), 1000)
    # the main file alone, whole
    alone = SeqDB.from_xbit(str(main), device=dev)
    assert alone.n > 0 and int(alone.seq_len.max()) <= 998


def test_db_index_and_gather(orc, dev):
    from sseqs_amd.db import SeqDB
    announce("db_index")
    rng = np.random.default_rng(2)
    seqs = [rand_seq(rng, L) for L in rng.integers(1, 700, 3000)]
    # leading '@', doubled '@' and no trailing '@' exercise empty runs
    text = ("@" + "@".join(seqs[:10]) + "@@" + "@".join(seqs[10:])).encode()
    db = SeqDB.from_packed(_packed_db(dev, text), len(text))
    assert db.n == len(seqs)
    lens = np.array([len(s) for s in seqs])
    order = np.argsort(lens, kind="stable")
    assert np.array_equal(db.seq_len.cpu().numpy(), lens[order])
    assert np.array_equal(db.orig_id.cpu().numpy(), order)
    ids = torch.arange(0, db.n, 7, device=dev)
    got = db.strings(ids, 500)
    want = [seqs[order[i]][:500] for i in ids.cpu().numpy()]
    assert got == want


def _index_streams(rng):
    """Residue (1) / '@' (0) patterns around the index's 16384-symbol tiles
    (64 symbols per thread): runs spanning several tiles, runs that end or
    start on a tile edge, a stream that is one run, alternation (8192 runs in
    a tile) and random densities."""
    T = 16384
    out = []
    one = np.ones(3 * T + 5, bool)
    out.append(one.copy())
    for cut in ([T - 1], [T], [T + 1], [T - 1, T], [2 * T - 1, 2 * T], [5], [3 * T + 4]):
        s = one.copy()
        s[cut] = False
        out.append(s)
    alt = np.zeros(2 * T + 3, bool)
    alt[::2] = True
    out.append(alt)
    out.append(~alt)
    for n, dens in ((T, 0.5), (T + 16, 0.02), (7 * T + 9, 0.003), (40 * T + 1, 0.0002), (1, 0.0)):
        out.append(rng.random(n) >= dens)
    for n in (1, 15, 16, 17, 63, 64, 65, T - 1, T, T + 1):
        out.append(np.ones(n, bool))
    s = one.copy()
    s[63::64] = False  # every thread's last symbol
    out.append(s)
    return out


def test_db_index_tile_edges(dev):
    from sseqs_amd import _lib
    from sseqs_amd._lib import call, ptr, stream_of
    announce("db_index tile edges")
    rng = np.random.default_rng(11)
    lib = _lib.load()
    for k, res in enumerate(_index_streams(rng)):
        n = res.size
        sym = np.where(res, rng.integers(65, 91, n), 64).astype(np.uint8)
        packed = torch.zeros((n * 5 + 7) // 8 + 64, dtype=torch.uint8, device=dev)
        c = _packed_db(dev, sym.tobytes())
        packed[: c.numel()] = c
        d = np.diff(np.concatenate([[0], res.astype(np.int8), [0]]))
        want_off = np.nonzero(d == 1)[0]
        want_len = np.nonzero(d == -1)[0] - want_off
        ws = _lib.workspace(lib.sseqs_db_index_ws_bytes(n), dev)
        cnt = torch.zeros(1, dtype=torch.int64, device=dev)
        cap = max(1, want_off.size)
        off = torch.full((cap,), -1, dtype=torch.int64, device=dev)
        ln = torch.full((cap,), -1, dtype=torch.int32, device=dev)
        call("sseqs_db_index", ptr(packed), n, ptr(off), ptr(ln), cap, ptr(cnt), ptr(ws),
             ws.numel(), stream_of(dev))
        assert int(cnt.item()) == want_off.size, k
        m = want_off.size
        assert np.array_equal(off.cpu().numpy()[:m], want_off), k
        assert np.array_equal(ln.cpu().numpy()[:m], want_len), k


# -------------------------------------------------------------------- SW --

def _batch(rng, query, n):
    """An ASCII batch of n targets joined by '@' with its delimiter positions."""
    tg = []
    for i in range(n):
        if i % 3 == 0:
            tg.append(mutate(rng, query[int(rng.integers(0, max(1, len(query) // 2))):]))
        else:
            tg.append(rand_seq(rng, rng.integers(1, 600), AA * 4 + "XBZ*-"))
    text = "@".join(tg) + "@"
    a = np.frombuffer(text.encode(), dtype=np.uint8).copy()
    D = np.nonzero(a == 64)[0].astype(np.int32)
    return a, D


@pytest.mark.parametrize("Lq", [1, 2, 5, 17, 31, 32, 33, 40, 63, 64, 65, 100, 200, 257, 500,
                                777, 1000, 1024, 1025, 2048])
def test_sw_kernels_match_oracle(orc, dev, Lq):
    from sseqs_amd import ops
    announce(f"sw Lq={Lq}")
    rng = np.random.default_rng(Lq)
    query = rand_seq(rng, Lq)
    qc = orc.encode_query_sw(query)
    a, D = _batch(rng, query, 150)
    gi = np.sort(rng.choice(len(D), size=101, replace=False)).astype(np.int32)
    gi[0] = 0  # target 0 reads P[-1]
    starts = (D + 1)[(D + 1) < a.size].astype(np.int32)
    tq, tgi, ta = (torch.from_numpy(x).to(dev) for x in (qc, gi, a))
    tD, tS = torch.from_numpy(D).to(dev), torch.from_numpy(starts).to(dev)
    for gap in (-1, -4, 0):
        got = ops.sw_cuda_blosum62(tq, tgi, ta, tD, gap_penalty_param=gap).cpu().numpy()
        assert np.array_equal(got, orc.sw_linear(qc, gi, a, D, gap)), gap
    for o, e in ((11, 1), (5, 2), (0, 0)):
        got = ops.sw_cuda_affine(tq, tgi, ta, tD, o, e).cpu().numpy()
        assert np.array_equal(got, orc.sw_affine(qc, gi, a, D, o, e)), (o, e)
    # the affine stage of the cascade passes starts = D + 1 (sseqs/msa.py:331)
    gi2 = gi[gi < starts.size]
    got = ops.sw_cuda_affine(tq, torch.from_numpy(gi2).to(dev), ta, tS).cpu().numpy()
    assert np.array_equal(got, orc.sw_affine(qc, gi2, a, starts))


@pytest.mark.parametrize("Lq", [1, 2, 15, 16, 17, 32, 33, 100, 200, 479, 500, 513, 1023, 1024,
                                1025, 2048])
def test_db_sw_matches_oracle(orc, dev, Lq, monkeypatch):
    """Resident-database SW (packed 16-bit kernel up to 1024 residues, 32-bit
    beyond and for gap parameters out of its range) equals the oracle, and
    the two kernels agree; candidates in random order, count not a multiple
    of the targets a wave holds."""
    from sseqs_amd.db import SeqDB
    from sseqs_amd.query import prepare
    from sseqs_amd.search import sw_affine_ids, sw_linear_ids
    announce(f"db sw Lq={Lq}")
    rng = np.random.default_rng(1000 + Lq)
    query = rand_seq(rng, Lq)
    seqs = [mutate(rng, query) if i % 6 == 0 else rand_seq(rng, rng.integers(1, 700), DB_ALPH)
            for i in range(203)]
    text = ("@".join(seqs) + "@").encode()
    db = SeqDB.from_packed(_packed_db(dev, text), len(text))
    q = prepare(query)
    ids = torch.from_numpy(rng.permutation(db.n)[:199]).to(dev)
    fid = ids.cpu().numpy() if db.orig_id is None else db.orig_id.cpu().numpy()[ids.cpu().numpy()]
    ascii = np.frombuffer(text, dtype=np.uint8)
    starts = np.concatenate([[0], np.cumsum([len(x) + 1 for x in seqs])[:-1]])[fid]
    lens = np.array([len(x) for x in seqs])[fid]
    qc = orc.encode_query_sw(query)
    for o, e in ((11, 1), (0, 0), (3, 2), (20000, 1)):
        want = orc.sw_affine_segments(qc, ascii, starts, lens, o, e)
        got = sw_affine_ids(db, q, ids, o, e).cpu().numpy()
        assert np.array_equal(got, want), (o, e)
        monkeypatch.setenv("SSEQS_SW_INT32", "1")
        assert np.array_equal(sw_affine_ids(db, q, ids, o, e).cpu().numpy(), want), (o, e)
        monkeypatch.delenv("SSEQS_SW_INT32")
    for gap in (-1, -4, 0, -20000):
        want = orc.sw_linear_segments(qc, ascii, starts, lens, gap)
        assert np.array_equal(sw_linear_ids(db, q, ids, gap).cpu().numpy(), want), gap


def test_sw_edge_cases(orc, dev):
    from sseqs_amd import ops
    q = orc.encode_query_sw("WCWHW")
    a = np.frombuffer(b"WCW@@HWA@W@", dtype=np.uint8).copy()
    P = np.array([3, 4, 8, 10], np.int32)
    gi = np.array([0, 1, 2, 3], np.int32)
    t = lambda x: torch.from_numpy(x).to(dev)
    got = ops.sw_cuda_affine(t(q), t(gi), t(a), t(P)).cpu().numpy()
    assert np.array_equal(got, orc.sw_affine(q, gi, a, P))
    assert got[1] == -1
    empty = ops.sw_cuda_affine(t(np.zeros(0, np.uint8)), t(gi), t(a), t(P)).cpu().numpy()
    assert empty.tolist() == [-1, -1, -1, -1]
    with pytest.raises(RuntimeError):
        ops.sw_cuda_affine(t(np.zeros(2049, np.uint8)), t(gi), t(a), t(P))
    with pytest.raises(RuntimeError):
        ops.sw_cuda_blosum62(t(q), t(gi), t(a), t(P), gap_penalty_param=2)
    with pytest.raises(TypeError):
        ops.sw_cuda_affine(t(q), t(gi.astype(np.int64)), t(a), t(P))


def test_public_sw(orc, dev):
    import sseqs_amd
    rng = np.random.default_rng(9)
    q = rand_seq(rng, 120)
    targets = [mutate(rng, q) for _ in range(5)] + [rand_seq(rng, 80) for _ in range(5)]
    got = sseqs_amd.sw(q, targets).cpu().numpy()
    qc = orc.encode_query_sw(q)
    assert got.tolist() == [orc.sw_affine_one(qc, t) for t in targets]


# ------------------------------------------------------------- prefilter --

@pytest.mark.parametrize("N,k", [(16, 1), (64, 4), (128, 2), (256, 4), (512, 4), (1024, 4),
                                 (2048, 4), (512, 8), (256, 20)])
def test_warp_corr_matches_fp64(orc, dev, N, k):
    from sseqs_amd import ops
    announce(f"warp_corr N={N} k={k}")
    rng = np.random.default_rng(N * 31 + k)
    Lq = int(rng.integers(max(2, N // 2), N + 1))
    query = rand_seq(rng, Lq, AA + "X")
    A, B = orc.weighted_luts(query, k)
    qrev = orc.encode_query_rev(query)
    segs = [mutate(rng, query) if i % 4 == 0 else rand_seq(rng, rng.integers(1, N + 1), AA + "XB")
            for i in range(97)]
    segs = [s[:N] for s in segs]
    flat = ("@".join(segs) + "@").encode()
    lens = np.array([len(s) for s in segs], np.int32)
    off = np.concatenate([[0], np.cumsum(lens + 1)[:-1]]).astype(np.int32)
    maxlen = int(lens.max()) if k != 8 else int(lens.max()) // 2  # exercise the clamp
    want = orc.corr(qrev, A, B, k, flat, off, lens, maxlen, N)
    tol = orc.corr_tol(qrev, A, B, k, flat, off, lens, maxlen, N, want)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    scores = torch.zeros(len(segs), dtype=torch.float16, device=dev)
    ops.warp_corr(scores, t(qrev), t(A), t(B), t(np.frombuffer(flat, np.uint8).copy()), t(off),
                  t(lens), k, Lq, maxlen, N, torch.zeros(1, dtype=torch.float16, device=dev))
    got = scores.float().cpu().numpy()
    ok = corr_ok(got, want, tol)
    assert ok.all(), (N, k, np.max(np.abs(got - want)), got[~ok][:5], want[~ok][:5])


def test_db_prefilter_sequences_longer_than_2048(orc, dev):
    """Sequences longer than the largest correlation size share N = 2048 and
    are scored on their first 2048 residues (whole-database prefilter)."""
    from sseqs_amd.db import SeqDB
    from sseqs_amd.query import prepare
    from sseqs_amd.search import prefilter_scores
    announce("db_prefilter with sequences longer than 2048")
    rng = np.random.default_rng(31)
    query = rand_seq(rng, 300)
    seqs = [rand_seq(rng, rng.integers(1, 600)) for _ in range(300)]
    seqs += [mutate(rng, query * 12) for _ in range(6)] + [rand_seq(rng, 5000)]
    text = ("@".join(seqs) + "@").encode()
    db = SeqDB.from_packed(_packed_db(dev, text), len(text))
    assert int(db.seq_len.max().item()) > 2048
    got = prefilter_scores(db, prepare(query)).cpu().numpy()
    order = db.orig_id.cpu().numpy() if db.orig_id is not None else np.arange(db.n)
    A, B = orc.weighted_luts(query, 4)
    qrev = orc.encode_query_rev(query)
    args = [(s.encode(), [0], [len(s)], 2048, min(2048, orc.nextpow2(max(300, len(s), 16))))
            for s in seqs]
    want = np.array([orc.corr(qrev, A, B, 4, *a)[0] for a in args])[order]
    tol = np.array([orc.corr_tol(qrev, A, B, 4, *a)[0] for a in args])[order]
    assert corr_ok(got, want, tol).all(), np.max(np.abs(got - want))


@pytest.mark.parametrize("Lq,k", [(50, 4), (200, 4), (500, 4), (200, 1), (90, 3)])
def test_db_prefilter_any_index_order(orc, dev, Lq, k):
    """Resident-database prefilter on an index that is length-sorted but not
    in stream order (a file that was not sorted), whose tiles the band
    kernel stages into per-sequence slots (or, past the stage, decodes from
    global memory), and on the same data in stream order."""
    from sseqs_amd.db import SeqDB
    from sseqs_amd.query import prepare
    from sseqs_amd.search import prefilter_scores
    announce(f"db_prefilter unsorted Lq={Lq} k={k}")
    rng = np.random.default_rng(Lq)
    # letters a real .xbit file holds besides the 20 residues take the
    # prefilter's code 20 (X) and 21 (everything else) paths
    query = rand_seq(rng, Lq, AA * 3 + "X")
    seqs = [mutate(rng, query) if i % 5 == 0 else rand_seq(rng, rng.integers(1, 1100), DB_ALPH)
            for i in range(2500)]
    text = ("@".join(seqs) + "@").encode()
    db = SeqDB.from_packed(_packed_db(dev, text), len(text))
    assert db.orig_id is not None
    if k == 4:
        # database-resident SW of some of them, affine and linear, bit-exact
        from sseqs_amd.query import prepare as prep
        from sseqs_amd.search import sw_affine_ids, sw_linear_ids
        q4 = prep(query)
        ids = torch.arange(0, db.n, 7, device=dev)
        fid = db.orig_id.cpu().numpy()[ids.cpu().numpy()]
        ascii = np.frombuffer(text, dtype=np.uint8)
        starts = np.concatenate([[0], np.cumsum([len(x) + 1 for x in seqs])[:-1]])
        lens = np.array([len(x) for x in seqs])
        qc = orc.encode_query_sw(query)
        assert np.array_equal(sw_affine_ids(db, q4, ids).cpu().numpy(),
                              orc.sw_affine_segments(qc, ascii, starts[fid], lens[fid]))
        assert np.array_equal(sw_linear_ids(db, q4, ids, -4).cpu().numpy(),
                              orc.sw_linear_segments(qc, ascii, starts[fid], lens[fid], -4))
    got = prefilter_scores(db, prepare(query, fft_rank=k)).cpu().numpy()
    order = db.orig_id.cpu().numpy()
    A, B = orc.weighted_luts(query, k)
    qrev = orc.encode_query_rev(query)
    want = np.zeros(len(seqs))
    tol = np.zeros(len(seqs))
    for i, s in enumerate(seqs):
        N = min(2048, orc.nextpow2(max(Lq, len(s), 16)))
        want[i] = orc.corr(qrev, A, B, k, s.encode(), [0], [len(s)], N, N)[0]
        tol[i] = orc.corr_tol(qrev, A, B, k, s.encode(), [0], [len(s)], N, N, want[i:i + 1])[0]
    want, tol = want[order], tol[order]
    assert corr_ok(got, want, tol).all(), np.max(np.abs(got - want))
    # the same sequences already sorted: contiguous tiles (staged path)
    srt = [seqs[i] for i in order]
    text2 = ("@".join(srt) + "@").encode()
    db2 = SeqDB.from_packed(_packed_db(dev, text2), len(text2))
    assert db2.orig_id is None
    got2 = prefilter_scores(db2, prepare(query, fft_rank=k)).cpu().numpy()
    assert np.array_equal(got2, got)


def test_warp_corr_float32_corr_is_not_updated(dev):
    """A float32 corr goes to a temporary fp16 copy (sseqs/filter_bind.cpp:34-35)."""
    from sseqs_amd import ops
    from sseqs_amd.query import prepare
    q = prepare("ACDEFGHIKLMNPQRSTVWY" * 3)
    flat = torch.tensor(list(b"ACDEF@GHIKL@"), dtype=torch.uint8, device=dev)
    c = torch.zeros(2, dtype=torch.float32, device=dev)
    ops.warp_corr(c, q.qrev, q.lut_a, q.lut_b, flat, torch.tensor([0, 6], dtype=torch.int32,
                  device=dev), torch.tensor([5, 5], dtype=torch.int32, device=dev), 4, 60, 5,
                  64, torch.zeros(1, dtype=torch.float16, device=dev))
    assert c.abs().sum().item() == 0.0


# ---------------------------------------------------------------- top-k --

@pytest.mark.parametrize("dtype", [torch.float32, torch.float16, torch.int32])
@pytest.mark.parametrize("n,k", [(1, 1), (10, 3), (5000, 100), (300000, 4096), (300000, 77777)])
def test_topk(orc, dev, dtype, n, k):
    from sseqs_amd import ops
    announce(f"topk {dtype} n={n} k={k}")
    rng = np.random.default_rng(n + k)
    if dtype == torch.int32:
        x = rng.integers(-50, 50, n).astype(np.int32)  # heavy ties
    else:
        x = np.round(rng.standard_normal(n), 2).astype(np.float32)
        x[rng.integers(0, n, max(1, n // 100))] = -np.inf
        if dtype == torch.float16:
            x = x.astype(np.float16)
    t = torch.from_numpy(x).to(dev)
    want = orc.topk_desc(x.astype(np.float64), k)
    if k <= 16384:
        got = ops.topk_indices(t, k, sorted=True).cpu().numpy()
        assert np.array_equal(got, want)
    got_u = ops.topk_indices(t, k, sorted=False).cpu().numpy()
    assert np.array_equal(got_u, np.sort(want))


@pytest.mark.parametrize("offset", [0, 3])
def test_topk_f16_ragged_and_unaligned(orc, dev, offset):
    """fp16 scores take eight per thread from 16-byte loads when the
    address is 16-byte aligned, one per thread otherwise: both equal the
    oracle, for a length that ends mid-vector and mid-tile, with heavy ties
    at the k-th score."""
    from sseqs_amd import ops
    announce(f"topk fp16 offset={offset}")
    rng = np.random.default_rng(99 + offset)
    n = 1_000_003
    x = np.round(rng.standard_normal(n + offset), 1).astype(np.float16)
    x[x == 0] = 0.0  # no -0.0: the keys order it below +0.0, the oracle ties them
    x[rng.integers(0, n, n // 50)] = -np.inf
    base = torch.from_numpy(x).to(dev)
    t = base[offset:]
    xs = x[offset:].astype(np.float64)
    for k in (1, 4096, 100_000, n):
        want = orc.topk_desc(xs, k)
        got = ops.topk_indices(t, k, sorted=False).cpu().numpy()
        assert np.array_equal(got, np.sort(want)), k
    got = ops.topk_indices(t, 5000, sorted=True).cpu().numpy()
    assert np.array_equal(got, orc.topk_desc(xs, 5000))


def test_topk_errors(dev):
    from sseqs_amd import ops
    t = torch.zeros(10, device=dev)
    with pytest.raises(ValueError):
        ops.topk_indices(t, 11)
    assert ops.topk_indices(t, 0).numel() == 0


@pytest.mark.parametrize("dtype", [torch.float16, torch.int32, torch.float32])
def test_topk_segments(orc, dev, dtype):
    """Per-segment selects in one launch equal the oracle's top-k of each
    segment (ties to the smaller index), in both orders, with the first
    and last entries dropped and with segment-relative or absolute ids."""
    from sseqs_amd import ops
    announce(f"topk_segments {dtype}")
    rng = np.random.default_rng(7)
    n = np.array([0, 1, 5, 300, 4097, 70000, 1000, 250000, 33], dtype=np.int64)
    begin = np.concatenate([[0], np.cumsum(n)[:-1]]) + 3
    total = int(begin[-1] + n[-1]) + 5
    if dtype == torch.int32:
        x = rng.integers(-40, 40, total).astype(np.int32)
    else:
        x = np.round(rng.standard_normal(total), 2).astype(np.float32)
        x[x == 0] = 0.0  # no -0.0: the keys order it below +0.0, the oracle ties them
        x[rng.integers(0, total, total // 100)] = -np.inf
        if dtype == torch.float16:
            x = x.astype(np.float16)
    t = torch.from_numpy(x).to(dev)
    for sorted_, kfun in ((False, lambda m: m // 7), (True, lambda m: min(m, 4096) // 3)):
        k = np.array([kfun(int(m)) for m in n], dtype=np.int64)
        k[2] = 5  # k == n
        for drop, add in (((0, 0), False), ((1, 1), True)):
            got, base = ops.topk_segments(t, begin, n, k, sorted=sorted_, drop_front=drop[0],
                                          drop_back=drop[1], add_begin=add)
            got = got.cpu().numpy()
            for j in range(n.size):
                seg = x[begin[j]: begin[j] + n[j]].astype(np.float64)
                want = orc.topk_desc(seg, int(k[j]))
                if not sorted_:
                    want = np.sort(want)
                want = want[drop[0]: max(int(k[j]) - drop[1], drop[0])]
                if add:
                    want = want + begin[j]
                part = got[base[j]: base[j] + want.size]
                assert np.array_equal(part, want), (j, sorted_, drop)
            assert got.size == sum(max(int(kk) - drop[0] - drop[1], 0) for kk in k)
    with pytest.raises(ValueError):
        ops.topk_segments(t, begin[:1], n[7:8], [5000], sorted=True)


# ------------------------------------------------------------- backtrack --

def test_backtrack_matches_oracle(orc, dev):
    from sseqs_amd.sw import sw_affine_backtrack
    announce("backtrack")
    rng = np.random.default_rng(21)
    for Lq in (1, 7, 64, 65, 130, 300):
        query = rand_seq(rng, Lq, AA + "X")
        targets = [mutate(rng, query) for _ in range(20)] + \
                  [rand_seq(rng, rng.integers(1, 400), AA + "B") for _ in range(20)]
        for o, e in ((11, 1), (3, 1)):
            _, best, got = sw_affine_backtrack(query, targets, o, e)
            want = [orc.backtrack(query, t, o, e) for t in targets]
            assert got == want, (Lq, o, e)
            assert best.cpu().tolist() == [int(w["score"]) for w in want]


def test_backtrack_and_a3m_match_reference_vectors(dev, tmp_path):
    """sw_affine_backtrack and to_a3m against vectors of the reference's own
    functions (tests/golden/a3m_golden.json)."""
    import json
    from pathlib import Path

    from sseqs_amd.sw import sw_affine_backtrack
    from sseqs_amd.to_a3m import to_a3m
    announce("backtrack + a3m vs reference vectors")
    golden = json.loads((Path(__file__).resolve().parent / "golden" / "a3m_golden.json").read_text())
    for i, c in enumerate(golden):
        _, best, aligned = sw_affine_backtrack(c["query"], c["targets"], c["gap_open"],
                                               c["gap_extend"])
        assert aligned == c["aligned"]
        assert best.cpu().tolist() == c["best"]
        fn = tmp_path / f"case{i}.a3m"
        to_a3m(c["query"], c["targets"], str(fn), gap_open=c["gap_open"],
               gap_extend=c["gap_extend"])
        assert fn.read_text() == c["a3m"]


# ------------------------------------------------------------- generator --

def test_generator_matches_cpu(orc, dev):
    from sseqs_amd import synth
    from sseqs_amd.db import SeqDB
    announce("generator")
    queries = [synth.make_query(200, 0), synth.make_query(57, 1)]
    n = 20000
    g = synth.generate_db(n, queries, seed=5, device=dev)
    cls = g.classes
    ascii_cpu, homq = orc.gen_db(synth.gen_params(5), queries, cls.start, cls.length, cls.sym, n)
    got = g.packed[: (g.n_syms * 5 + 7) // 8].cpu().numpy()
    assert np.array_equal(got, orc.xbit_pack(ascii_cpu))
    assert np.array_equal(g.hom_query.cpu().numpy(), homq)
    assert (homq >= 0).sum() == n // 20
    # the index the generator writes equals the delimiter scan of its stream
    db = SeqDB.from_packed(g.packed, g.n_syms)
    assert db.orig_id is None
    assert torch.equal(db.seq_off, g.seq_off) and torch.equal(db.seq_len, g.seq_len)
