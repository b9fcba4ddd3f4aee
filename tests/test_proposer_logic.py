"""Host logic of the learned proposal (no GPU): the query rows a query is
embedded as (whole, then windows), the calibrated candidate count per query
length, and the per-query recall gate that decides between the learned
proposal and the plain prefilter."""
from types import SimpleNamespace

import numpy as np
import pytest

from sseqs_amd.learned import Proposer


def _prop(**kw):
    return Proposer(SimpleNamespace(n=1_000_000), **kw)


def test_rows_of_windows():
    p = _prop(window=384, stride=192)
    q_short, q_exact, q_long = "A" * 200, "C" * 384, "".join("ACDEFGHIKL"[i % 10] for i in range(1000))
    rows, count = p.rows_of([q_short, q_exact, q_long])
    assert list(count) == [1, 1, 1 + 5]
    assert rows[0] == q_short and rows[1] == q_exact and rows[2] == q_long
    wins = rows[3:]
    # windows every 192 residues, the last one ending at the query's end
    assert [len(w) for w in wins] == [384] * 5
    assert wins[0] == q_long[:384] and wins[1] == q_long[192:576] and wins[2] == q_long[384:768]
    assert wins[3] == q_long[576:960] and wins[4] == q_long[-384:]
    # no windows when disabled
    rows, count = _prop(window=0).rows_of([q_long])
    assert rows == [q_long] and list(count) == [1]


def test_group_size_checked():
    with pytest.raises(ValueError):
        _prop(group=20)
    with pytest.raises(ValueError):
        _prop(group=256)


def test_calibrated_counts_and_gate():
    p = _prop()
    assert p.meets_gate(["A" * 50], [10]).all()  # uncalibrated: always learned
    with pytest.raises(RuntimeError):
        p.top_prefilter_for(["A" * 50])
    p.k_by_len = [(0, 100, 5000), (100, 200, 300), (200, 2049, 7000)]
    p.calibration = {"top_prefilter": 100_000}
    p.use_learned = False
    qs = ["A" * 99, "A" * 100, "A" * 199, "A" * 200, "A" * 3000]
    assert list(p.top_prefilter_for(qs)) == [5000, 300, 300, 7000, 7000]
    counts = np.array([5000, 299, 100_000, 6999, 10**6])
    assert list(p.meets_gate(qs, counts)) == [True, False, True, False, True]
    # the gate held for every calibration query at 100 000: that count or more passes
    p.use_learned = True
    assert list(p.meets_gate(qs, [100_000, 1, 1, 100_000, 1])) == [True, False, False, True, False]
    p.enforce_gate = False
    assert p.meets_gate(qs, [1] * 5).all()


def test_fused_proposal_arguments_checked():
    """sseqs_propose_max rejects its bad arguments on the host, before any
    launch: query rows not a multiple of 16 or above 64, more than 32 output
    rows, offsets that do not increase or leave the G rows; both proposal
    entry points reject a table or query rows not 16-byte aligned."""
    import ctypes as C

    from sseqs_amd._lib import SseqsError, call
    for G, off in ((80, [0, 1]), (24, [0, 1]), (64, list(range(34))), (64, [0, 3, 3]),
                   (64, [0, 5, 2]), (32, [0, 40]), (32, [-1, 4])):
        arr = np.array(off, np.int32)
        with pytest.raises(SseqsError, match="propose_max: "):
            call("sseqs_propose_max", C.c_void_p(16), 16, C.c_void_p(16), G,
                 arr.ctypes.data, len(off) - 1, C.c_void_p(16), 16, None)
    # operands the 16-byte loads cannot take
    ok = np.array([0, 3], np.int32)
    for table, q8 in ((17, 16), (16, 24)):
        with pytest.raises(SseqsError, match="16-byte aligned"):
            call("sseqs_propose_max", C.c_void_p(table), 16, C.c_void_p(q8), 16, ok.ctypes.data, 1,
                 C.c_void_p(16), 16, None)
        with pytest.raises(SseqsError, match="16-byte aligned"):
            call("sseqs_propose", C.c_void_p(table), 16, C.c_void_p(q8), 16, C.c_void_p(16), 16,
                 None)


def test_launch_groups(monkeypatch):
    """Launch groups on the host (device calls stubbed): every query in one
    group with its rows contiguous; at most `group` rows per group, first
    fit by decreasing row count; a group with a windowed query and at most 32
    queries (and at most 64 rows) gets the fused launch's row offsets, which
    are its queries' row runs, strictly increasing; other groups get none."""
    import torch

    from sseqs_amd import learned

    staged = []  # the fp32 query rows handed to the fp8 conversion, per group

    def fake_call(name, *args):
        if name == "sseqs_enc_to_fp8":
            staged.append(args[0].clone())

    monkeypatch.setattr(learned, "call", fake_call)
    monkeypatch.setattr(learned, "ptr", lambda t: t)
    monkeypatch.setattr(learned, "stream_of", lambda dev: 0)
    enc = SimpleNamespace(embed_strings=lambda rows: torch.arange(
        len(rows), dtype=torch.float32)[:, None].repeat(1, learned.DIM))
    tab = SimpleNamespace(n=1_000_000, table=torch.zeros(1, dtype=torch.uint8), enc=enc)
    rng = np.random.default_rng(3)
    lens = list(rng.integers(30, 1000, 150)) + [50] * 60
    queries = ["".join("ACDEFGHIKLMNPQRSTVWY"[(i + j) % 20] for j in range(L))
               for i, L in enumerate(lens)]
    for group in (64, 128):
        p = Proposer(tab, group=group)
        rows, count = p.rows_of(queries)
        first = np.concatenate([[0], np.cumsum(count)[:-1]])
        staged.clear()
        groups = p._groups(queries)
        assert len(staged) == len(groups)
        seen = np.concatenate([g[0] for g in groups])
        assert np.array_equal(np.sort(seen), np.arange(len(queries)))
        fused = 0
        for (qidx, q8, loc, cnt, off), q in zip(groups, staged):
            assert np.array_equal(cnt, count[qidx]) and int(cnt.sum()) <= group
            # the group's rows: each query's rows in order (the stub embeds
            # row r as the value r), then zero padding to a multiple of 16
            want = np.concatenate([first[i] + np.arange(count[i]) for i in qidx])
            assert np.array_equal(q[: want.size, 0].numpy(), want.astype(np.float32))
            assert not q[want.size:].any()
            assert q8.shape == ((int(cnt.sum()) + 15) // 16 * 16, learned.DIM)
            assert np.array_equal(loc, np.concatenate([[0], np.cumsum(cnt)[:-1]]))
            want_fused = cnt.max() > 1 and qidx.size <= 32 and q8.shape[0] <= 64
            assert (off is not None) == want_fused
            if off is not None:
                fused += 1
                assert off.dtype == np.int32 and off.size == qidx.size + 1
                assert np.array_equal(off[:-1], loc) and off[-1] == cnt.sum()
                assert np.all(np.diff(off) > 0)
        assert fused > 0 or group == 128
        assert sum(int(g[3].sum()) for g in groups) == int(count.sum()) == len(rows)
        # first fit by decreasing row count: a query did not fit any earlier
        # group when it was placed, so a later group's first query (its
        # largest) would overflow every earlier group's final fill
        fill = [int(g[3].sum()) for g in groups]
        for j in range(1, len(groups)):
            head = int(groups[j][3].max())
            assert all(fill[i] + head > group for i in range(j)), (group, j)
