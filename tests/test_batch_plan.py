"""The batch plan (segments of every batch from the resident index) against
the oracle's segmentation of each batch's ASCII window (no GPU needed)."""
import numpy as np
import pytest
import torch

from sseqs_amd.batches import plan_batches

AA = np.frombuffer(b"ARNDCQEGHILKMFPSTWYVXBZ", dtype=np.uint8)


def make_stream(rng, n_runs, lead_at, gaps_of_two):
    """ASCII stream of runs separated by '@' (sometimes '@@' or more)."""
    parts = [b"@" * lead_at]
    for i in range(n_runs):
        L = int(rng.integers(1, 60))
        parts.append(rng.choice(AA, L).tobytes())
        gap = 1
        if gaps_of_two and rng.random() < 0.1:
            gap = int(rng.integers(2, 5))
        parts.append(b"@" * gap)
    return np.frombuffer(b"".join(parts), dtype=np.uint8)


def runs_of(a):
    """The resident index: every maximal run of non-'@' symbols."""
    x = np.concatenate([[64], a, [64]])
    nz = (x != 64).astype(np.int8)
    d = np.diff(nz)
    st = np.nonzero(d == 1)[0]
    en = np.nonzero(d == -1)[0]
    return st.astype(np.int64), (en - st).astype(np.int64)


@pytest.mark.parametrize("seed,lead_at,gaps,bs,shuffle", [
    (0, 0, False, 5, False), (1, 0, False, 25, False), (2, 1, False, 10, True),
    (3, 0, True, 15, True), (4, 2, True, 100, False), (5, 0, True, 5, True),
    (6, 0, False, 1000, False),
])
def test_plan_matches_oracle_segments(orc, seed, lead_at, gaps, bs, shuffle):
    rng = np.random.default_rng(seed)
    a = make_stream(rng, 300, lead_at, gaps)
    off, ln = runs_of(a)
    if shuffle:  # the index may be in length order rather than stream order
        p = rng.permutation(off.size)
        off, ln = off[p], ln[p]
    plan = plan_batches(torch.from_numpy(off), torch.from_numpy(ln.astype(np.int32)), a.size, bs)
    sb = bs * 8 // 5
    nbytes = (a.size * 5 + 7) // 8
    assert plan.n_batches == -(-nbytes // bs)
    seg_off = plan.seg_off.numpy()
    seg_len = plan.seg_len.numpy()
    sw_off = plan.sw_off.numpy()
    sw_len = plan.sw_len.numpy()
    n_regular = n_irregular = 0
    for b in range(plan.n_batches):
        s0 = b * sb
        w = a[s0: s0 + plan.n_sym[b]]
        assert plan.n_sym[b] == min(sb, a.size - s0)
        D, starts, lengths, fft_off = orc.segments(w)
        assert plan.nseg[b] == lengths.size
        assert plan.maxlen[b] == (int(lengths.max()) if lengths.size else 0)
        # true start of every non-empty segment of the window
        x = np.concatenate([[-1], D, [w.size]])
        true_st = np.array([x[i] + 1 for i in range(x.size - 1) if x[i + 1] - x[i] > 1],
                           dtype=np.int64)
        regular = lengths.size > 0 and np.array_equal(fft_off[:-1], true_st)
        assert bool(plan.regular[b]) == regular
        sl = slice(plan.seg_base[b], plan.seg_base[b] + plan.nseg[b])
        if regular:
            n_regular += 1
            assert np.array_equal(seg_off[sl] - s0, fft_off[:-1])
            assert np.array_equal(seg_len[sl], lengths)
            assert np.array_equal(sw_off[sl], seg_off[sl]) and np.array_equal(sw_len[sl],
                                                                              seg_len[sl])
        elif lengths.size:
            n_irregular += 1
            # the shifted batch's two reference views
            assert plan.shifted[b]
            assert np.array_equal(seg_off[sl] - s0, fft_off[:-1])
            assert np.array_equal(seg_len[sl], lengths)
            # SW target g between delimiters g-1 and g (P[-1] = -1), for every
            # g the selection can reach (the first and last are dropped)
            for g in range(1, lengths.size - 1):
                st = (D[g - 1] if g > 0 else -1) + 1
                assert sw_off[sl][g] - s0 == st, (b, g)
                assert sw_len[sl][g] == D[g] - st, (b, g)
    assert n_regular > 0
    if lead_at or gaps:
        assert n_irregular > 0
    # the prefilter's order: each batch's own segments, by length, stable
    if bs >= 1000:  # ~50 whole runs of 1..59 residues per batch: sorting them pays
        assert plan.pf_perm is not None
    if plan.pf_perm is None:
        return
    perm = plan.pf_perm.numpy()
    for b in range(plan.n_batches):
        lo_s, hi_s = int(plan.seg_base[b]), int(plan.seg_base[b] + plan.nseg[b])
        pb = perm[lo_s:hi_s]
        assert np.array_equal(np.sort(pb), np.arange(lo_s, hi_s))
        assert np.array_equal(pb, lo_s + np.argsort(seg_len[lo_s:hi_s], kind="stable"))
    assert np.array_equal(plan.pf_off.numpy(), seg_off[perm])
    assert np.array_equal(plan.pf_len.numpy(), seg_len[perm])


def test_plan_keeps_length_sorted_order():
    """A stream already in length order gets no prefilter permutation."""
    ln = np.repeat(np.arange(10, 200), 40).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(ln + 1)[:-1]])
    n = int(off[-1] + ln[-1] + 1)
    plan = plan_batches(torch.from_numpy(off), torch.from_numpy(ln.astype(np.int32)), n, 500)
    assert plan.pf_perm is None


def test_plan_rejects_unaligned_batches():
    off = torch.tensor([0], dtype=torch.int64)
    ln = torch.tensor([3], dtype=torch.int32)
    with pytest.raises(ValueError):
        plan_batches(off, ln, 3, 7)


@pytest.mark.parametrize("seed", range(6))
def test_oracle_segments_are_split_lengths(orc, seed):
    """The oracle's delimiter arithmetic (sseqs/msa.py:227-267) gives the
    lengths of str.split('@') without its empty parts (a trailing '@' adds
    no part), and offsets that are the running sum of (length + 1)."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    a = rng.choice(np.frombuffer(b"@@@ARNDWY", dtype=np.uint8), n)
    if seed == 0:
        a[0] = 64
    if seed == 1:
        a[-1] = 64
    s = a.tobytes().decode()
    parts = s.split("@")
    if s.endswith("@"):
        parts = parts[:-1]
    want = [len(p) for p in parts if len(p)]
    D, starts, lengths, fft_off = orc.segments(a)
    assert lengths.tolist() == want
    assert np.array_equal(D, [i for i, c in enumerate(s) if c == "@"])
    assert fft_off.tolist() == [0] + np.cumsum(np.array(want) + 1).tolist()
