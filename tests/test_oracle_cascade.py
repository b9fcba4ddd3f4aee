"""The oracle's per-batch cascade and final merge against hand-derived answers.

The stream below is built so that every rule of the reference cascade
(sseqs/msa.py:297-345, 354-356, 411-427) changes the expected hits if it is
broken.  The prefilter scores are supplied (fp16 values chosen by hand), so
everything after them is integer and follows from BLOSUM62 by hand:

* the query is eight W; against W, BLOSUM62 gives W 11, Y 2, F 1, P -4.
  Every candidate holds a single run of W (plus at most a Y/F tail) inside
  P filler, so its best local alignment is that run, ungapped, of at most
  eight residues: linear or affine score = 11 * (W aligned) + tail.
* "inner" runs start after "PP", so dropping the first residue (the affine
  stage scores segment[1:] + '@', sseqs/msa.py:331 with sseqs/sw.cu:77)
  leaves the score unchanged; runs at the segment start lose one W (11).

Batch 0 (20 segments, top_fft=2 -> k1 = 20 // 2 = 10 by prefilter):
  top 10 = {1, 3, 4, 6, 8, 10, 12, 14, 16, 18}; in index order without the
  first and last (sseqs/msa.py:307) -> {3, 4, 6, 8, 10, 12, 14, 16} (M = 8).
  Dropping by score instead would remove 8 (highest, 9.0) and 12 (lowest,
  4.5) and keep the strong 1 and 18.
  seg   run          linear  affine(seg[1:]+'@')
   8    inner 7W+F     78      78        (A)
   3    start 7W       77      66        (B)
   4    start 6W+YY    70      59        (C)
   6    start 6W+Y     68      57        (D)
  10    inner 6W+F     67      67        (E)
  16    start 4W       44      33
  12    inner 3W       33      33
  14    inner 2W       22      22
  k2 = M // top_sw = 8 // 2 = 4 -> {A, B, C, D}; affine top 2 -> A (78), B (66).
  - k2 from the undropped count (10 // 2 = 5) would add E (67): hits A, E.
  - linear on segment[1:] would rank A 78, E 67, B 66, C 59: hits A, E.
  - affine on the exact segment would give B 77.
  maxlen of the batch is 50 (a P-only segment that is never selected), so
  norm = affine / sqrt(8 * 50) = affine / 20 (not the candidates' lengths).
Batch 1 (12 segments, k1 = 6 -> top 6 = {0, 2, 5, 7, 9, 11}, kept
  {2, 5, 7, 9}, k2 = 4 // 2 = 2):
   2    inner 5W       55      55   -> norm 55 / sqrt(8 * 32) = 55 / 16
   5    inner 4W       44      44   -> 44 / 16
   7    inner 3W       33
   9    inner 2W       22
  (0 and 11 hold 7W runs and are dropped as first/last).
Final merge (np.argsort(norm)[::-1][:num_msas], sseqs/msa.py:426-427) with
num_msas = 3: 78/20 = 3.9, 55/16 = 3.4375, 66/20 = 3.3 (and 44/16 = 2.75
cut) -- ordering by affine score instead would put B (66) before 55.
"""
import numpy as np

Q = "W" * 8


def _inner(run: str) -> str:
    return "PP" + run + "PPP"


def _start(run: str) -> str:
    return run + "PPPP"


def _stream():
    b0 = ["P" * 6] * 20
    s0 = [0.0] * 20
    b0[1], s0[1] = _inner("W" * 7 + "Y"), 8.5     # strong, dropped as first
    b0[18], s0[18] = _inner("W" * 7 + "Y"), 8.0   # strong, dropped as last
    b0[8], s0[8] = _inner("W" * 7 + "F"), 9.0     # A
    b0[3], s0[3] = _start("W" * 7), 7.5           # B
    b0[4], s0[4] = _start("W" * 6 + "YY"), 7.0    # C
    b0[6], s0[6] = _start("W" * 6 + "Y"), 6.5     # D
    b0[10], s0[10] = _inner("W" * 6 + "F"), 6.0   # E
    b0[16], s0[16] = _start("W" * 4), 5.0
    b0[12], s0[12] = _inner("WWW"), 4.5
    b0[14], s0[14] = _inner("WW"), 5.5
    # the 11th best: would enter with k1 = 11; a strong inner run
    b0[5], s0[5] = _inner("W" * 7 + "Y"), 4.0
    b0[19] = "P" * 50                             # the batch maxlen
    for i in (0, 2, 7, 9, 11, 13, 15, 17, 19):
        s0[i] = 1.0 + 0.125 * i                   # all below 4.0
    b1 = ["P" * 5] * 12
    s1 = [0.5] * 12
    b1[0], s1[0] = _inner("W" * 7), 6.0
    b1[11], s1[11] = _inner("W" * 7), 5.0
    b1[2], s1[2] = _inner("W" * 5), 3.0
    b1[5], s1[5] = _inner("W" * 4), 4.0
    b1[7], s1[7] = _inner("WWW"), 3.5
    b1[9], s1[9] = _inner("WW"), 2.5
    b1[4] = "P" * 32                              # the batch maxlen
    for i in (1, 3, 4, 6, 8, 10):
        s1[i] = 0.25 + 0.125 * i
    t0 = "".join(s + "@" for s in b0)
    t1 = "".join(s + "@" for s in b1)
    stream = np.frombuffer((t0 + t1).encode(), dtype=np.uint8)
    scores = [np.array(s0, np.float16), np.array(s1, np.float16)]
    return stream, len(t0), scores, b0, b1


def test_hand_fixture_is_consistent(orc):
    """The fixture's premises: distinct prefilter scores, and the linear and
    affine scores claimed above for each candidate."""
    _, _, scores, b0, b1 = _stream()
    assert len(set(scores[0].tolist())) == 20 and len(set(scores[1].tolist())) == 12
    order0 = np.argsort(-scores[0].astype(np.float64))
    assert sorted(order0[:10].tolist()) == [1, 3, 4, 6, 8, 10, 12, 14, 16, 18]
    assert order0[10] == 5 and order0[0] == 8 and order0[9] == 12
    order1 = np.argsort(-scores[1].astype(np.float64))
    assert sorted(order1[:6].tolist()) == [0, 2, 5, 7, 9, 11]
    assert max(len(s) for s in b0) == 50 and max(len(s) for s in b1) == 32
    qc = orc.encode_query_sw(Q)
    claims0 = {8: (78, 78), 3: (77, 66), 4: (70, 59), 6: (68, 57), 10: (67, 67),
               16: (44, 33), 12: (33, 33), 14: (22, 22)}
    for i, (lin, aff) in claims0.items():
        assert orc.sw_linear_one(qc, b0[i], -1) == lin, i
        assert orc.sw_affine_one(qc, b0[i][1:] + "@") == aff, i
    for i, (lin, aff) in {2: (55, 55), 5: (44, 44), 7: (33, 33), 9: (22, 22)}.items():
        assert orc.sw_linear_one(qc, b1[i], -1) == lin, i
        assert orc.sw_affine_one(qc, b1[i][1:] + "@") == aff, i


def test_batch_cascade_hand_derived(orc):
    stream, bs_syms, scores, b0, b1 = _stream()
    hits, fft64 = orc.batch_cascade(stream, Q, bs_syms, top_fft=2, top_sw=2, top_sw_affine=2,
                                    fft_scores=lambda j: scores[j])
    assert [w.size for w in fft64] == [20, 12]
    want = [
        (np.float32(78) / np.float32(20), 78, 78, b0[8]),
        (np.float32(66) / np.float32(20), 66, 77, b0[3]),
        (np.float32(55) / np.float32(16), 55, 55, b1[2]),
        (np.float32(44) / np.float32(16), 44, 44, b1[5]),
    ]
    assert hits == want
    ranked = orc.rank_hits(hits, 3)
    assert [h[3] for h in ranked] == [b0[8], b1[2], b0[3]]
    assert [float(h[0]) for h in ranked] == [3.9000000953674316, 3.4375, 3.299999952316284]


def test_batch_cascade_skips_short_batches(orc):
    """A batch whose stage-2 keeps fewer than top_sw_affine candidates adds
    no hits (the reference raises inside torch.topk there); here batch 1
    keeps 4 // 3 = 1 < 2."""
    stream, bs_syms, scores, b0, _ = _stream()
    hits, _ = orc.batch_cascade(stream, Q, bs_syms, top_fft=2, top_sw=3, top_sw_affine=2,
                                fft_scores=lambda j: scores[j])
    # batch 0: k2 = 8 // 3 = 2 -> {A, B}; batch 1 skipped
    assert [h[3] for h in hits] == [b0[8], b0[3]]
