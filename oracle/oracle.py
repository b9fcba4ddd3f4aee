"""CPU oracle for the sseqs MSA-search hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker (or the timed CPU baseline);
the product package ``sseqs_amd`` never imports it.

Heavy recurrences live in oracle.c (built to oracle/_oracle.so by
oracle/Makefile); the host-side steps of the reference (LUTs, encodings,
segmentation, the per-batch cascade, the a3m writer) are restated here in
numpy.  Each function cites the reference file:line it follows.  Tie orders
that the reference leaves to torch.topk are fixed here the same way the
product fixes them: score descending, then index ascending; where the
reference sorts with numpy (final merge, a3m rows) the same numpy call is
made.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "_oracle.so"

AA = "ARNDCQEGHILKMFPSTWYV"
AA_TO_POS = {a: i for i, a in enumerate(AA)}

# BLOSUM62 restated independently of the product (sseqs/msa.py:110).
BLOSUM62 = np.array([
    4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0,
    -1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3,
    -2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3,
    -2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3,
    0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1,
    -1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2,
    -1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2,
    0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3,
    -2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3,
    -1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3,
    -1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1,
    -1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2,
    -1, -1, -2, -3, -1, 0, -2, -3, -2, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1,
    -2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1,
    -1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2,
    1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2,
    0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0,
    -3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3,
    -2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -1,
    0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4,
]).reshape(20, 20)

_lib = None


def build() -> Path:
    """Compile oracle.c with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return SO


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not SO.exists():
            build()
        L = C.CDLL(str(SO))
        vp, i32, i64 = C.c_void_p, C.c_int, C.c_int64
        L.orc_xbit_pack.argtypes = [vp, i64, vp]
        L.orc_xbit_unpack.argtypes = [vp, i64, i64, vp]
        L.orc_sw_linear_one.argtypes = [vp, i32, vp, i32, i32]
        L.orc_sw_affine_one.argtypes = [vp, i32, vp, i32, i32, i32]
        L.orc_gotoh_textbook.argtypes = [vp, i32, vp, i32, i32, i32]
        L.orc_sw_linear_tpt.argtypes = [vp, i32, vp, i32, i32, i32]
        L.orc_sw_affine_tpt.argtypes = [vp, i32, vp, i32, i32, i32, i32]
        L.orc_sw_linear.argtypes = [vp, i32, vp, i32, vp, vp, i32, vp]
        L.orc_sw_affine.argtypes = [vp, i32, vp, i32, vp, vp, i32, i32, vp]
        L.orc_sw_affine_segments.argtypes = [vp, i32, vp, vp, vp, i64, i32, i32, vp]
        L.orc_sw_linear_segments.argtypes = [vp, i32, vp, vp, vp, i64, i32, vp]
        L.orc_corr.argtypes = [vp, i32, vp, vp, i32, i32, i32, vp, vp, vp, i32, i32, i32, vp]
        L.orc_backtrack.argtypes = [vp, i32, vp, i32, i32, i32, vp, vp, vp]
        L.orc_gen_db.argtypes = [vp, vp, vp, i32, vp, vp, vp, i32, i64, vp, vp]
        L.orc_gen_query.argtypes = [vp, i64, i32, vp]
        for f in ("orc_sw_linear_one", "orc_sw_affine_one", "orc_gotoh_textbook", "orc_backtrack",
                  "orc_sw_linear_tpt", "orc_sw_affine_tpt"):
            getattr(L, f).restype = i32
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _u8(x) -> np.ndarray:
    if isinstance(x, str):
        x = x.encode()
    return np.ascontiguousarray(np.frombuffer(bytes(x), dtype=np.uint8) if isinstance(x, (bytes, bytearray)) else np.asarray(x, dtype=np.uint8))


# ------------------------------------------------------------- codec ----

def xbit_pack(ascii_bytes) -> np.ndarray:
    """sseqs/xbit.cu:36-62 (compress)."""
    a = _u8(ascii_bytes)
    out = np.zeros((a.size * 5 + 7) // 8, dtype=np.uint8)
    lib().orc_xbit_pack(_p(a), a.size, _p(out))
    return out


def xbit_unpack(packed, n: int, lo: int = 0) -> np.ndarray:
    """sseqs/xbit.cu:8-34 (decompress), reading from byte lo."""
    p = _u8(packed)
    need = lo + ((n + 7) // 8) * 5
    if p.size < need:
        p = np.concatenate([p, np.zeros(need - p.size, np.uint8)])
    out = np.zeros(n, dtype=np.uint8)
    lib().orc_xbit_unpack(_p(p), n, lo, _p(out))
    return out


# ---------------------------------------------------- encodings, LUTs ----

def encode_query_sw(query: str) -> np.ndarray:
    """SW query codes: residues only, others dropped (sseqs/msa.py:156)."""
    return np.array([AA_TO_POS[c] for c in query if c in AA_TO_POS], dtype=np.uint8)


def encode_query_rev(query: str) -> np.ndarray:
    """Reversed prefilter codes, 20 for anything else (sseqs/msa.py:119,157)."""
    return np.array([AA_TO_POS.get(c, 20) for c in query[::-1]], dtype=np.uint8)


def weighted_luts(query: str, fft_rank: int = 4):
    """LUT_A, LUT_B (20 x 21 float32) of sseqs/msa.py:125-145."""
    B = BLOSUM62.copy()
    if fft_rank == 1:
        B[B < 0] = 0
    w = np.sqrt(np.array([query.count(a) for a in AA]) / len(query))
    U, S, Vt = np.linalg.svd(w[:, None] * B)
    A = np.stack([np.append(np.sqrt(s) * u, 0.0) for s, u in zip(S, U.T)]).astype(np.float32)
    Bt = np.stack([np.append(np.sqrt(s) * v, 0.0) for s, v in zip(S, Vt)]).astype(np.float32)
    return A, Bt


# ------------------------------------------------------------ segments ----

def segments(ascii_batch: np.ndarray):
    """Delimiter positions, segment lengths (after the `!= 0` filter) and FFT
    offsets of one decompressed batch (sseqs/msa.py:227-267)."""
    a = np.asarray(ascii_batch, dtype=np.uint8)
    D = np.nonzero(a == 64)[0].astype(np.int64)
    starts = D + 1
    starts = starts[starts < a.size]
    offsets = np.concatenate([starts, [a.size]]).astype(np.int64)
    lengths = offsets - np.concatenate([[0], offsets[:-1]])
    lengths[:-1] -= 1
    if a.size and a[-1] == 64:
        lengths[-1] -= 1
    lengths = lengths[lengths != 0]
    fft_off = np.zeros(lengths.size + 1, dtype=np.int64)
    fft_off[1:] = np.cumsum(lengths + 1)
    return D, starts, lengths, fft_off


# --------------------------------------------------------------- SW -------

def sw_linear(q_codes, good_idx, ascii, P, gap: int) -> np.ndarray:
    """sseqs/sw.cu:55-179 via the reference selection rule."""
    q, g, a, p = _u8(q_codes), np.ascontiguousarray(good_idx, np.int32), _u8(ascii), np.ascontiguousarray(P, np.int32)
    out = np.zeros(g.size, dtype=np.int32)
    lib().orc_sw_linear(_p(q), q.size, _p(g), g.size, _p(a), _p(p), int(gap), _p(out))
    return out


def sw_affine(q_codes, good_idx, ascii, P, o: int = 11, e: int = 1) -> np.ndarray:
    """sseqs/sw.cu:182-325 via the reference selection rule."""
    q, g, a, p = _u8(q_codes), np.ascontiguousarray(good_idx, np.int32), _u8(ascii), np.ascontiguousarray(P, np.int32)
    out = np.zeros(g.size, dtype=np.int32)
    lib().orc_sw_affine(_p(q), q.size, _p(g), g.size, _p(a), _p(p), int(o), int(e), _p(out))
    return out


def sw_affine_one(q_codes, target: bytes | str, o=11, e=1) -> int:
    q, t = _u8(q_codes), _u8(target)
    return lib().orc_sw_affine_one(_p(q), q.size, _p(t), t.size, o, e)


def sw_linear_one(q_codes, target: bytes | str, gap=-1) -> int:
    q, t = _u8(q_codes), _u8(target)
    return lib().orc_sw_linear_one(_p(q), q.size, _p(t), t.size, gap)


def sw_affine_tpt(q_codes, target, o, e, tpt) -> int:
    q, t = _u8(q_codes), _u8(target)
    return lib().orc_sw_affine_tpt(_p(q), q.size, _p(t), t.size, o, e, tpt)


def sw_linear_tpt(q_codes, target, gap, tpt) -> int:
    q, t = _u8(q_codes), _u8(target)
    return lib().orc_sw_linear_tpt(_p(q), q.size, _p(t), t.size, gap, tpt)


def gotoh_textbook(q_codes, target, o=11, e=1) -> int:
    q, t = _u8(q_codes), _u8(target)
    return lib().orc_gotoh_textbook(_p(q), q.size, _p(t), t.size, o, e)


def sw_affine_segments(q_codes, ascii, off, ln, o=11, e=1) -> np.ndarray:
    q, a = _u8(q_codes), _u8(ascii)
    off = np.ascontiguousarray(off, np.int64)
    ln = np.ascontiguousarray(ln, np.int32)
    out = np.zeros(off.size, dtype=np.int32)
    lib().orc_sw_affine_segments(_p(q), q.size, _p(a), _p(off), _p(ln), off.size, o, e, _p(out))
    return out


def sw_linear_segments(q_codes, ascii, off, ln, gap=-4) -> np.ndarray:
    """sseqs/sw.cu:55-179 over exact segments (offset, length)."""
    q, a = _u8(q_codes), _u8(ascii)
    off = np.ascontiguousarray(off, np.int64)
    ln = np.ascontiguousarray(ln, np.int32)
    out = np.zeros(off.size, dtype=np.int32)
    lib().orc_sw_linear_segments(_p(q), q.size, _p(a), _p(off), _p(ln), off.size, gap, _p(out))
    return out


# ------------------------------------------------------------ prefilter --

def corr(qrev, lutA: np.ndarray, lutB: np.ndarray, k: int, flat, offsets, lengths,
         maxlen: int, N: int) -> np.ndarray:
    """FP64 closed form of the FFT prefilter (sseqs/filter.cu:92-159,
    sseqs/fft.cu:17-415): max over even m of 0.5 * circular correlation."""
    q = _u8(qrev)
    A = np.ascontiguousarray(lutA, np.float32)
    Bm = np.ascontiguousarray(lutB, np.float32)
    V = A.shape[1]
    f = _u8(flat)
    off = np.ascontiguousarray(offsets, np.int32)
    ln = np.ascontiguousarray(lengths, np.int32)
    out = np.zeros(ln.size, dtype=np.float64)
    lib().orc_corr(_p(q), q.size, _p(A), _p(Bm), Bm.shape[0], V, k, _p(f), _p(off), _p(ln),
                   ln.size, int(maxlen), int(N), _p(out))
    return out


def corr_tol(qrev, lutA: np.ndarray, lutB: np.ndarray, k: int, flat, offsets, lengths,
             maxlen: int, N: int, x64: np.ndarray | None = None) -> np.ndarray:
    """Per-sequence tolerance of a device prefilter score against the FP64
    closed form: 1e-2 * |x64| plus the error that rounding the two operands
    a = 0.5 * LUT_A and b = LUT_B to fp16 and summing n = k * Lq products in
    fp32 can cause.  An operand rounded to fp16 is off by at most 2^-11 of
    itself plus 2^-25 (half the smallest subnormal, for values that
    underflow), so a product is off by at most (2^-10 + 2^-22) |a||b| +
    2^-25 (1 + 2^-11) (|a| + |b|) + 2^-50, and the fp32 sum adds at most
    n * 2^-24 of the absolute sum.  The sum of |a_r| * |b_r| over the terms
    of a shift is at most the closed form evaluated with |LUT_A| and
    |LUT_B| (the maximum over shifts of those absolute sums)."""
    if x64 is None:
        x64 = corr(qrev, lutA, lutB, k, flat, offsets, lengths, maxlen, N)
    absum = corr(qrev, np.abs(lutA), np.abs(lutB), k, flat, offsets, lengths, maxlen, N)
    n_terms = int(k) * len(qrev)
    amax = 0.5 * float(np.max(np.abs(np.asarray(lutA, np.float64)[:k])))
    bmax = float(np.max(np.abs(np.asarray(lutB, np.float64))))
    under = n_terms * (2.0**-25 * (1 + 2.0**-11) * (amax + bmax) + 2.0**-50)
    return 1e-2 * np.abs(x64) + (2.0**-10 + 2.0**-22 + n_terms * 2.0**-24) * absum + under


def corr_fft_emulation(qrev, lutA, lutB, k: int, seg_codes_fn, L: int, maxlen: int, N: int) -> float:
    """The reference's own transform pipeline (packed real FFT, frequency
    product, packed inverse, 1/N, even samples) evaluated in FP64 with
    numpy, for one segment; used to pin the closed form."""
    H = N // 2
    Bflat = np.asarray(lutB, np.float64).reshape(-1)
    V = np.asarray(lutA).shape[1]
    acc = np.zeros(H, dtype=np.complex128)
    for r in range(k):
        a = np.zeros(N)
        for t in range(min(len(qrev), N)):
            a[t] = lutA[r][qrev[t]]
        b = np.zeros(N)
        for t in range(min(L, maxlen, N)):
            b[t] = Bflat[r * V + seg_codes_fn(t)]
        # forward: H-point FFT of packed pairs, then the real split
        def rfft_packed(x):
            z = x[0::2] + 1j * x[1::2]
            Z = np.fft.fft(z)
            X = np.zeros(H + 1, dtype=np.complex128)
            for kk in range(H + 1):
                zk = Z[kk % H]
                zc = np.conj(Z[(H - kk) % H])
                w = np.exp(-2j * np.pi * kk / N)
                X[kk] = 0.5 * (zk + zc) - 0.5j * w * (zk - zc)
            return X
        P = rfft_packed(a) * rfft_packed(b)
        # inverse: pack the half spectrum, unnormalised H-point inverse
        Zi = np.zeros(H, dtype=np.complex128)
        for kk in range(H):
            pk, pc = P[kk], np.conj(P[H - kk])
            w = np.exp(2j * np.pi * kk / N)
            Zi[kk] = 0.5 * (pk + pc) + 0.5j * w * (pk - pc)
        acc += np.fft.ifft(Zi) * H
    return float(np.max((acc / N).real))


# ------------------------------------------------------------ backtrack --

def backtrack(query: str, target: str, o=11, e=1) -> dict:
    """DP + stateless traceback of sseqs/sw.py:408-768 for one target, as the
    dict sw_affine_backtrack builds (sseqs/sw.py:196-217)."""
    q, t = _u8(query), _u8(target)
    qa = np.zeros(q.size + t.size + 2, np.uint8)
    ta = np.zeros_like(qa)
    info = np.zeros(6, np.int32)
    lib().orc_backtrack(_p(q), q.size, _p(t), t.size, o, e, _p(qa), _p(ta), _p(info))
    n = int(info[0])
    return {
        "q_aligned": qa[:n].tobytes().decode("latin-1"),
        "t_aligned": ta[:n].tobytes().decode("latin-1"),
        "q_start_orig_0based": int(info[1]),
        "q_end_orig_0based": int(info[2]),
        "t_start_orig_0based": int(info[3]),
        "t_end_orig_0based": int(info[4]),
        "score": float(info[5]),
    }


def a3m(query: str, aligned: list[dict]) -> str:
    """a3m text of sseqs/to_a3m.py:16-72 from backtrack dicts, ordered by
    np.argsort of the Python float scores, reversed (sseqs/to_a3m.py:16-17)."""
    scores = [float(a["score"]) for a in aligned]
    order = np.argsort(scores)[::-1]
    aligned = [aligned[i] for i in order]
    out = f">original_query\n{query}\n"
    seen = set()
    for i, ares in enumerate(aligned):
        if ares["score"] == 0 or not ares["q_aligned"]:
            continue
        q_loc, t_loc = ares["q_aligned"], ares["t_aligned"]
        qc = q_loc.replace("-", "")
        s = query.find(qc)
        if s == -1:
            raise ValueError("Cannot locate aligned fragment in query.")
        mid = "".join("-" if tc == "-" else (tc.upper() if qch != "-" else tc.lower())
                      for qch, tc in zip(q_loc, t_loc))
        row = "-" * s + mid + "-" * (len(query) - (s + len(qc)))
        if "@" in row or row in seen:
            continue
        seen.add(row)
        out += f">target_{i} score={ares['score']}\n{row}\n"
    return out


# ------------------------------------------------------------ selection --

def topk_desc(scores: np.ndarray, k: int) -> np.ndarray:
    """Indices of the k largest, (score desc, index asc)."""
    s = np.asarray(scores)
    order = np.lexsort((np.arange(s.size), -s.astype(np.float64)))
    return order[:k]


# ------------------------------------------------------------ generator --

def gen_db(params, queries: list[str], cls_start, cls_len, cls_sym, n: int):
    """CPU restatement of the synthetic generator: ASCII symbols and the
    homolog query id per sequence."""
    codes = [np.array([AA_TO_POS[c] for c in q], np.uint8) for q in queries]
    q_off = np.concatenate([[0], np.cumsum([c.size for c in codes])]).astype(np.int64)
    qc = np.concatenate(codes) if codes else np.zeros(1, np.uint8)
    cs = np.ascontiguousarray(cls_start, np.int64)
    cl = np.ascontiguousarray(cls_len, np.int32)
    cy = np.ascontiguousarray(cls_sym, np.int64)
    n_syms = int(cy[-1]) if cy.size == cl.size + 1 else int(cy[-1] + cl[-1] + 1)
    out = np.zeros(n_syms, dtype=np.uint8)
    homq = np.zeros(n, dtype=np.int32)
    lib().orc_gen_db(C.addressof(params), _p(qc), _p(q_off), len(queries), _p(cs), _p(cl), _p(cy),
                     cl.size, n, _p(out), _p(homq))
    return out, homq


# ------------------------------------------------------------- cascades --

def nextpow2(x: int) -> int:
    """sseqs/msa.py:261."""
    return 1 << (x - 1).bit_length()


def batch_cascade(db_ascii: np.ndarray, query: str, bs_syms: int, top_fft=200, top_sw=10,
                  top_sw_affine=2, fft_rank=4, fft_scores=None, with_tol=False):
    """Per-batch cascade of sseqs/msa.py:217-370 on a decompressed stream cut
    into batches of bs_syms symbols.  fft_scores(j) may supply the batch's
    prefilter scores (the stages after it are integer and exact); otherwise
    the FP64 closed form rounded to fp16 is used.  Returns the hits as
    (norm float32, affine, linear, seq) in collection order, plus the
    batch's FP64 prefilter scores (and, with_tol, their corr_tol
    tolerances)."""
    A, Bt = weighted_luts(query, fft_rank)
    k = max(1, min(fft_rank, A.shape[0]))
    qrev, qc = encode_query_rev(query), encode_query_sw(query)
    Lq = len(query)
    hits, fft64, tol64 = [], [], []
    for j, lo in enumerate(range(0, db_ascii.size, bs_syms)):
        a = db_ascii[lo: lo + bs_syms]
        D, starts, lengths, off = segments(a)
        if lengths.size == 0:
            fft64.append(np.zeros(0))
            tol64.append(np.zeros(0))
            continue
        maxlen = int(lengths.max())
        # sseqs/msa.py:261; sizes outside 16..2048 (scored 0 by the
        # reference's kernels, sseqs/filter.cu:313-390) are clamped, longer
        # segments scored on their first 2048 residues
        n = min(max(nextpow2(max(Lq, maxlen)), 16), 2048)
        s64 = corr(qrev, A, Bt, k, a, off[:-1], lengths, maxlen, n)
        fft64.append(s64)
        if with_tol:
            tol64.append(corr_tol(qrev, A, Bt, k, a, off[:-1], lengths, maxlen, n, s64))
        s = fft_scores(j) if fft_scores is not None else s64.astype(np.float16)
        good = np.sort(topk_desc(np.asarray(s, np.float64), s.size // top_fft))[1:-1]
        if good.size == 0:
            continue
        lin = sw_linear(qc, good.astype(np.int32), a, D.astype(np.int32), -1)
        ktop = lin.size // top_sw
        if ktop < top_sw_affine:
            continue
        t1 = topk_desc(lin, ktop)
        gap = sw_affine(qc, good[t1].astype(np.int32), a, starts.astype(np.int32), 11, 1)
        t2 = topk_desc(gap, top_sw_affine)
        pos = t1[t2]
        norm = gap[t2].astype(np.float32) / np.sqrt(np.float32(Lq * maxlen))
        for r, p in enumerate(pos):
            seg = int(good[p])
            st, L = int(off[seg]), int(lengths[seg])
            hits.append((np.float32(norm[r]), int(gap[t2][r]), int(lin[p]),
                         a[st: st + min(L, 500)].tobytes().decode("latin-1")))
    if with_tol:
        return hits, fft64, tol64
    return hits, fft64


def rank_hits(hits, num_msas):
    """sseqs/msa.py:426-427: np.argsort of the float32 normalised scores,
    reversed (ties as numpy orders them)."""
    s = np.array([h[0] for h in hits], dtype=np.float32)
    return [hits[i] for i in np.argsort(s)[::-1][:num_msas]]


def gen_seqs(params, queries: list[str], cls_start, cls_len, ids) -> list[str]:
    """Selected sequences of the synthetic database (same algorithm as
    gen_db), for spot checks of databases too large to restate whole."""
    codes = [np.array([AA_TO_POS[c] for c in q], np.uint8) for q in queries]
    q_off = np.concatenate([[0], np.cumsum([c.size for c in codes])]).astype(np.int64)
    qc = np.concatenate(codes) if codes else np.zeros(1, np.uint8)
    cs = np.ascontiguousarray(cls_start, np.int64)
    cl = np.ascontiguousarray(cls_len, np.int32)
    ids = np.ascontiguousarray(ids, np.int64)
    out = np.zeros((ids.size, 1024), np.uint8)
    ln = np.zeros(ids.size, np.int32)
    L = lib()
    L.orc_gen_seqs.argtypes = [C.c_void_p] * 4 + [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                                  C.c_int64, C.c_void_p, C.c_void_p]
    L.orc_gen_seqs.argtypes[3] = C.c_int
    L.orc_gen_seqs(C.addressof(params), _p(qc), _p(q_off), len(queries), _p(cs), _p(cl), cl.size,
                   _p(ids), ids.size, _p(out), _p(ln))
    return [out[r, : ln[r]].tobytes().decode("latin-1") for r in range(ids.size)]
