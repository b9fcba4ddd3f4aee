"""Pins the oracle's Smith-Waterman recurrences.

* Exhaustive enumeration of local alignments on tiny inputs, for the CUDA
  affine rule (gaps open from M only, so a horizontal run can not touch a
  vertical run), the CUDA linear rule, and the stripe rule (no horizontal
  move into a column j with j % tpt == 0, sseqs/sw.cu:118,250), with tpt
  forced to 1, 2 and 3 so the rule is exercised on 3-6 residue queries.
* The reference's selection rule of targets by delimiter positions
  (sseqs/sw.cu:67-77).
* Invariants at realistic sizes: stripe score <= textbook score, equality
  when horizontal gaps can not pay off, query-symmetric inputs.
"""

import numpy as np
import pytest

AA = "ARNDCQEGHILKMFPSTWYV"


def _sub(B62, t, qc):
    ti = AA.find(t) if t in AA else -1
    return -4 if ti < 0 else int(B62[ti][qc])


def brute_cuda(B62, q, t, o, e, tpt, affine=True, gap=-1):
    """max over local alignments (starting and ending with a match step)."""
    Lq, T = len(q), len(t)
    best = 0

    def dfs(i, j, score, last, run):
        nonlocal best
        if last == "D":
            best = max(best, score)
        # diagonal
        if i + 1 < T and j + 1 < Lq:
            dfs(i + 1, j + 1, score + _sub(B62, t[i + 1], q[j + 1]), "D", 0)
        # horizontal: consume query column j+1 (not at a stripe start)
        if j + 1 < Lq and (j + 1) % tpt != 0:
            if affine:
                if last == "D":
                    dfs(i, j + 1, score - o, "H", 1)
                elif last == "H":
                    dfs(i, j + 1, score - e, "H", run + 1)
            else:
                dfs(i, j + 1, score + gap, "H", 0)
        # vertical: consume target row i+1
        if i + 1 < T:
            if affine:
                if last == "D":
                    dfs(i + 1, j, score - o, "V", 1)
                elif last == "V":
                    dfs(i + 1, j, score - e, "V", run + 1)
            else:
                dfs(i + 1, j, score + gap, "V", 0)

    for i0 in range(T):
        for j0 in range(Lq):
            dfs(i0, j0, _sub(B62, t[i0], q[j0]), "D", 0)
    return best


def brute_triton(B62, q, t, o, e):
    """Local alignment with gaps opening from any state (sseqs/sw.py:511-536):
    adjacent horizontal/vertical runs each pay the open cost."""
    def s(a, b):
        ia, ib = AA.find(a) if a in AA else -1, AA.find(b) if b in AA else -1
        return -5 if ia < 0 or ib < 0 else int(B62[ia][ib])

    Lq, T = len(q), len(t)
    best = 0

    def dfs(i, j, score, last):
        nonlocal best
        if last == "D":
            best = max(best, score)
        if i + 1 < Lq and j + 1 < T:
            dfs(i + 1, j + 1, score + s(q[i + 1], t[j + 1]), "D")
        if i + 1 < Lq:
            dfs(i + 1, j, score - (e if last == "X" else o), "X")
        if j + 1 < T:
            dfs(i, j + 1, score - (e if last == "Y" else o), "Y")

    for i0 in range(Lq):
        for j0 in range(T):
            dfs(i0, j0, s(q[i0], t[j0]), "D")
    return best


def _rand(rng, n, alphabet=AA):
    return "".join(rng.choice(list(alphabet), n))


@pytest.mark.parametrize("tpt", [1, 2, 3])
def test_affine_stripe_bruteforce(orc, tpt):
    rng = np.random.default_rng(10 + tpt)
    for trial in range(60):
        Lq, T = int(rng.integers(1, 6)), int(rng.integers(1, 6))
        # a repetitive alphabet makes gapped alignments competitive
        q = _rand(rng, Lq, "WWCAHX")
        t = _rand(rng, T, "WWCAHB")
        qc = np.array([AA.index(c) for c in q if c in AA], np.uint8)
        qs = "".join(c for c in q if c in AA)
        o, e = [(11, 1), (3, 1), (1, 0), (2, 2)][trial % 4]
        want = brute_cuda(orc.BLOSUM62, qc.tolist(), t, o, e, tpt, affine=True) if qs else 0
        got = orc.sw_affine_tpt(qc, t, o, e, tpt)
        assert got == want, (q, t, o, e, tpt)


@pytest.mark.parametrize("tpt", [1, 2, 3])
def test_linear_stripe_bruteforce(orc, tpt):
    rng = np.random.default_rng(20 + tpt)
    for trial in range(60):
        Lq, T = int(rng.integers(1, 6)), int(rng.integers(1, 6))
        q = [int(x) for x in rng.integers(0, 20, Lq)]
        t = _rand(rng, T, "WWCAHRX")
        g = [-1, 0, -2, -4][trial % 4]
        want = brute_cuda(orc.BLOSUM62, q, t, 0, 0, tpt, affine=False, gap=g)
        assert orc.sw_linear_tpt(np.array(q, np.uint8), t, g, tpt) == want, (q, t, g, tpt)


def test_stripe_width_is_ceil_lq_over_32(orc):
    """orc_sw_*_one uses tpt = ceil(Lq/32) (sseqs/sw.cu:355,79)."""
    rng = np.random.default_rng(5)
    for Lq in (1, 31, 32, 33, 64, 65, 200, 500, 1000):
        q = rng.integers(0, 20, Lq).astype(np.uint8)
        t = _rand(rng, 150)
        tpt = (Lq + 31) // 32
        assert orc.sw_affine_one(q, t) == orc.sw_affine_tpt(q, t, 11, 1, tpt)
        assert orc.sw_linear_one(q, t, -1) == orc.sw_linear_tpt(q, t, -1, tpt)


def test_triton_dp_bruteforce(orc):
    rng = np.random.default_rng(7)
    for trial in range(80):
        q = _rand(rng, int(rng.integers(1, 6)), "WWCAHX")
        t = _rand(rng, int(rng.integers(1, 6)), "WWCAHB")
        o, e = [(11, 1), (3, 1), (1, 0), (2, 1)][trial % 4]
        want = brute_triton(orc.BLOSUM62, q, t, o, e)
        r = orc.backtrack(q, t, o, e)
        assert r["score"] == want, (q, t, o, e)


def test_stripe_vs_textbook(orc):
    rng = np.random.default_rng(11)
    for _ in range(40):
        Lq = int(rng.integers(33, 300))
        q = rng.integers(0, 20, Lq).astype(np.uint8)
        t = _rand(rng, int(rng.integers(20, 300)))
        s = orc.sw_affine_one(q, t)
        tb = orc.gotoh_textbook(q, t)
        assert 0 <= s <= tb
        # with prohibitive gap costs both reduce to the best ungapped segment
        assert orc.sw_affine_one(q, t, 10**4, 10**4) == orc.gotoh_textbook(q, t, 10**4, 10**4)


def test_ungapped_closed_form(orc):
    """Huge gap costs: best diagonal maximum-subarray sum (Kadane)."""
    rng = np.random.default_rng(12)
    for _ in range(20):
        q = rng.integers(0, 20, int(rng.integers(1, 80))).astype(np.uint8)
        t = _rand(rng, int(rng.integers(1, 80)))
        best = 0
        for d in range(-len(t) + 1, len(q)):
            run = 0
            for i in range(len(t)):
                j = i + d
                if 0 <= j < len(q):
                    run = max(0, run + _sub(orc.BLOSUM62, t[i], int(q[j])))
                    best = max(best, run)
        assert orc.sw_affine_one(q, t, 10**4, 10**4) == best
        assert orc.sw_linear_one(q, t, -10**4) == best


def test_self_alignment(orc):
    """Identical sequences: score >= sum of the diagonal."""
    rng = np.random.default_rng(13)
    s = _rand(rng, 120)
    q = np.array([AA.index(c) for c in s], np.uint8)
    diag = sum(int(orc.BLOSUM62[AA.index(c)][AA.index(c)]) for c in s)
    assert orc.sw_affine_one(q, s) == diag == orc.gotoh_textbook(q, s)


def test_selection_rule(orc):
    """Targets come from delimiter positions; -1 for empty targets and
    stop < start; equal positions give score 0 (sseqs/sw.cu:67-77)."""
    q = np.array([AA.index(c) for c in "WCWHW"], np.uint8)
    ascii = b"WCW@@HWA@W"
    P = np.array([3, 4, 8], np.int32)
    gi = np.array([0, 1, 2], np.int32)
    out = orc.sw_affine(q, gi, ascii, P)
    assert out[0] == orc.sw_affine_one(q, b"WCW")       # P[-1] taken as -1
    assert out[1] == -1                                  # empty segment between '@@'
    assert out[2] == orc.sw_affine_one(q, b"HWA")
    lin = orc.sw_linear(q, gi, ascii, P, -1)
    assert lin[1] == -1 and lin[2] == orc.sw_linear_one(q, b"HWA", -1)
    # stop < start -> -1 ; equal positions -> length -1, stop == start -> 0
    assert orc.sw_affine(q, np.array([1], np.int32), ascii, np.array([5, 3], np.int32))[0] == -1
    assert orc.sw_affine(q, np.array([1], np.int32), ascii, np.array([5, 5], np.int32))[0] == 0
    # empty query -> -1
    assert orc.sw_affine(np.zeros(0, np.uint8), gi, ascii, P).tolist() == [-1, -1, -1]


def test_non_residue_scores(orc):
    """CUDA kernels score non-residues -4 (sseqs/sw.cu:128-130); Triton -5
    (sseqs/sw.py:260)."""
    q = np.array([AA.index("W")], np.uint8)
    assert orc.sw_affine_one(q, b"X") == 0
    assert orc.sw_affine_one(q, b"W") == 11
    r = orc.backtrack("WXW", "WBW")
    # W/W + X/B (-5) + W/W = 17 > 11
    assert r["score"] == 17.0 and r["q_aligned"] == "WXW" and r["t_aligned"] == "WBW"
