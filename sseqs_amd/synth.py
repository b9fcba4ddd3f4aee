"""Synthetic protein database and queries for the benchmark workloads.

The reference searches a downloaded ColabFold ``.xbit`` file
(sseqs/msa.py:18-25), which is not available offline.  This module defines
the synthetic stand-in used by the benchmark and the tests:

* residues i.i.d. from the BLOSUM62 background frequencies
  (Henikoff & Henikoff 1992) over ARNDCQEGHILKMFPSTWYV;
* lengths log-normal (median 250, sigma 0.5), rounded, clipped to
  [30, 998], laid out in ascending length order (the reference's database
  appears length-sorted), with the exact quantiles of that distribution;
* one sequence in every ``hom_period`` (default 20, i.e. 5 %) is a planted
  homolog of one of the queries: a copy at 30-90 % identity with
  substitutions drawn from BLOSUM62 target frequencies
  P(b|a) ~ p_b * 2^(s_ab/2), b != a, and geometric indels (rate 0.02 per
  position, mean length 3), embedded in background residues.

Everything is integer arithmetic on a counter-based hash, so the GPU
generator (csrc/gen.hip) and the CPU restatement give identical bytes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import GenParams, call, ptr, stream_of

AA = "ARNDCQEGHILKMFPSTWYV"

# BLOSUM62 background residue frequencies (Henikoff & Henikoff 1992),
# order ARNDCQEGHILKMFPSTWYV.
BG_FREQ = np.array([0.074, 0.052, 0.045, 0.054, 0.025, 0.034, 0.054, 0.074, 0.026, 0.068,
                    0.099, 0.058, 0.025, 0.047, 0.039, 0.057, 0.051, 0.013, 0.032, 0.073])

# BLOSUM62, same order (sseqs/msa.py:110).
BLOSUM62 = np.array([
    [4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0],
    [-1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3],
    [-2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3],
    [-2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3],
    [0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1],
    [-1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2],
    [-1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2],
    [0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3],
    [-2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3],
    [-1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3],
    [-1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1],
    [-1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2],
    [-1, -1, -2, -3, -1, 0, -2, -3, -2, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1],
    [-2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1],
    [-1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2],
    [1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2],
    [0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0],
    [-3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3],
    [-2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -1],
    [0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4],
], dtype=np.int64)

ONE = 1 << 24  # probabilities are 24-bit integers


def _cdf24(p: np.ndarray) -> np.ndarray:
    c = np.cumsum(p / p.sum())
    out = np.floor(c * ONE + 0.5).astype(np.int64)
    out[-1] = ONE
    return out.astype(np.int32)


def gen_params(seed: int = 0, hom_period: int = 20, p_indel: float = 0.02,
               mean_indel: float = 3.0, id_lo: float = 0.30, id_hi: float = 0.90) -> GenParams:
    """Generator parameters as the C struct shared with the kernels."""
    gp = GenParams()
    gp.seed = int(seed) & ((1 << 64) - 1)
    for a, v in enumerate(_cdf24(BG_FREQ)):
        gp.bg_cdf[a] = int(v)
    for a in range(20):
        w = BG_FREQ * np.power(2.0, BLOSUM62[a] / 2.0)
        w[a] = 0.0
        for b, v in enumerate(_cdf24(w)):
            gp.sub_cdf[20 * a + b] = int(v)
    gp.p_indel = int(round(p_indel * ONE))
    gp.p_ins = ONE // 2
    gp.p_ext = int(round((1.0 - 1.0 / mean_indel) * ONE))  # geometric, mean `mean_indel`
    gp.id_lo = int(round(id_lo * ONE))
    gp.id_span = int(round((id_hi - id_lo) * ONE))
    gp.hom_period = int(hom_period)
    return gp


@dataclass
class LengthClasses:
    """Sorted length layout of n sequences: class c has count[c] sequences
    of length length[c]; start/sym are cumulative sequence/symbol offsets."""

    length: np.ndarray  # int32 [K]
    count: np.ndarray   # int64 [K]
    start: np.ndarray   # int64 [K+1]
    sym: np.ndarray     # int64 [K+1] (sym[K] = total symbols incl. delimiters)

    @property
    def n(self) -> int:
        return int(self.start[-1])

    @property
    def n_syms(self) -> int:
        return int(self.sym[-1])


def length_classes(n: int, median: float = 250.0, sigma: float = 0.5, lo: int = 30,
                   hi: int = 998) -> LengthClasses:
    """Exact-quantile counts of round(LogNormal(ln median, sigma)) clipped to
    [lo, hi] for n sequences, in ascending length order."""
    ls = np.arange(lo, hi + 1)
    z = (np.log(ls + 0.5) - math.log(median)) / sigma
    F = np.array([0.5 * math.erfc(-x / math.sqrt(2.0)) for x in z])
    F[-1] = 1.0
    cum = np.floor(F * n + 0.5).astype(np.int64)
    cum[-1] = n
    cnt = np.diff(np.concatenate([[0], cum]))
    keep = cnt > 0
    length = ls[keep].astype(np.int32)
    count = cnt[keep].astype(np.int64)
    start = np.concatenate([[0], np.cumsum(count)]).astype(np.int64)
    sym = np.concatenate([[0], np.cumsum(count * (length.astype(np.int64) + 1))]).astype(np.int64)
    return LengthClasses(length, count, start, sym)


def mix64(z: int) -> int:
    m = (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def rnd24(seed: int, stream: int, ctr: int) -> int:
    m = (1 << 64) - 1
    h = mix64(((seed * 0x9E3779B97F4A7C15) & m) ^ mix64((stream * 0xD1B54A32D192ED03 + ctr) & m))
    return h >> 40


def make_query(length: int, qid: int = 0, seed: int = 0) -> str:
    """Background-distributed query of the given length (stream 2^60 | qid)."""
    cdf = _cdf24(BG_FREQ)
    out = []
    for p in range(length):
        r = rnd24(seed, (1 << 60) | qid, p)
        out.append(AA[int(np.searchsorted(cdf, r, side="right"))])
    return "".join(out)


def encode_codes(seq: str) -> np.ndarray:
    idx = {a: i for i, a in enumerate(AA)}
    return np.array([idx[c] for c in seq], dtype=np.uint8)


@dataclass
class SyntheticDB:
    packed: torch.Tensor     # uint8, ceil(5*n_syms/8) + db.PAD padding bytes
    seq_off: torch.Tensor    # int64 [n]
    seq_len: torch.Tensor    # int32 [n]
    hom_query: torch.Tensor  # int32 [n], query id or -1
    classes: LengthClasses
    n_syms: int


def generate_db(n: int, queries: list[str], seed: int = 0, device="cuda",
                hom_period: int = 20, **length_kw) -> SyntheticDB:
    """Generate the length-sorted synthetic database on the GPU."""
    dev = torch.device(device)
    cls = length_classes(n, **length_kw)
    if int(cls.length.max()) > 1024:
        raise ValueError("synthetic sequences are limited to 1024 residues")
    gp = gen_params(seed, hom_period=hom_period)
    codes = [encode_codes(q) for q in queries]
    q_off = np.concatenate([[0], np.cumsum([len(c) for c in codes])]).astype(np.int64)
    qc = np.concatenate(codes) if codes else np.zeros(1, np.uint8)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d_q, d_qoff = t(qc), t(q_off)
    d_start, d_len, d_sym = t(cls.start), t(cls.length), t(cls.sym[:-1].copy())
    n_syms = cls.n_syms
    from .db import PAD
    packed = torch.zeros((n_syms * 5 + 7) // 8 + PAD, dtype=torch.uint8, device=dev)
    seq_off = torch.empty(n, dtype=torch.int64, device=dev)
    seq_len = torch.empty(n, dtype=torch.int32, device=dev)
    homq = torch.empty(n, dtype=torch.int32, device=dev)
    lib = _lib.load()
    ws = _lib.workspace(lib.sseqs_gen_ws_bytes(n, hom_period), dev)
    call("sseqs_gen_db", gp, ptr(d_q), ptr(d_qoff), len(queries), ptr(d_start), ptr(d_len),
         ptr(d_sym), len(cls.length), n, n_syms, ptr(packed), ptr(seq_off), ptr(seq_len),
         ptr(homq), ptr(ws), ws.numel(), stream_of(dev))
    return SyntheticDB(packed, seq_off, seq_len, homq, cls, n_syms)


def shuffled_copy(db, seed: int = 3):
    """The sequences of ``db`` written to a new packed stream in random order
    and indexed by ``SeqDB.from_packed``, which sorts the index by length:
    the layout of a database file that is not in length order.  Returns
    (database, perm): stream sequence i of the copy is sequence perm[i] of
    ``db``, so index entry j of the copy is sequence perm[orig_id[j]]."""
    from . import ops
    from .db import SeqDB
    dev = db.device
    asc = ops.decompress(db.packed, db.n_syms, 0)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    perm = torch.randperm(db.n, device=dev, generator=g)
    ln = db.seq_len[perm].to(torch.int64)
    out_off = torch.cumsum(ln + 1, 0) - (ln + 1)
    src_off = db.seq_off[perm]
    total = int((ln + 1).sum().item())
    out = torch.full((total,), 64, dtype=torch.uint8, device=dev)
    # a million sequences per step: each index array stays far below 2^31
    # elements, which whole-database gathers and scatters exceed
    for c0 in range(0, db.n, 1 << 20):
        lc = ln[c0:c0 + (1 << 20)]
        seg = torch.repeat_interleave(torch.arange(lc.numel(), device=dev), lc)
        pos = torch.arange(seg.numel(), device=dev) - (torch.cumsum(lc, 0) - lc)[seg]
        out[out_off[c0:c0 + (1 << 20)][seg] + pos] = asc[src_off[c0:c0 + (1 << 20)][seg] + pos]
        del seg, pos
    del asc
    return SeqDB.from_packed(ops.compress(out), total), perm


# ------------------------------------------------------- benchmark set-up --

# The benchmark database holds planted homologs of three query sets, one
# homolog slot per block of BENCH_HOM_PERIOD sequences (12.5 %), the slot's
# query drawn uniformly from a list in which the 200- and 500-residue
# queries appear BENCH_REPEAT times each and every batched query once:
# 512 / 1280 of 12.5 % = 5 % of the database for each single query (the
# share the single-query workloads state), and 12.5 % / 1280 (about
# 19 500 sequences of 200 M) for each of the 256 batched queries.
BENCH_HOM_PERIOD = 8
BENCH_REPEAT = 512
BENCH_BATCH = 256


def batch_query_lengths(n: int = BENCH_BATCH, seed: int = 1, lo: int = 50,
                        hi: int = 1000) -> np.ndarray:
    """Lengths of the batched workload's queries, uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=n)


def bench_queries(seed: int = 1, n_batch: int = BENCH_BATCH) -> tuple[str, str, list[str]]:
    """(200-residue query, 500-residue query, batched queries)."""
    q200 = make_query(200, 0, seed=seed)
    q500 = make_query(500, 1, seed=seed)
    lens = batch_query_lengths(n_batch, seed)
    batch = [make_query(int(L), 2 + i, seed=seed) for i, L in enumerate(lens)]
    return q200, q500, batch


def bench_query_list(q200: str, q500: str, batch: list[str]) -> tuple[list[str], np.ndarray]:
    """The generator's query list and, per list entry, the workload query it
    stands for (0: q200, 1: q500, 2 + i: batched query i)."""
    qs = [q200] * BENCH_REPEAT + [q500] * BENCH_REPEAT + list(batch)
    owner = np.concatenate([np.zeros(BENCH_REPEAT, np.int64), np.ones(BENCH_REPEAT, np.int64),
                            2 + np.arange(len(batch), dtype=np.int64)])
    return qs, owner


def bench_db(n: int, device="cuda", seed: int = 0, query_seed: int = 1):
    """The benchmark database of n sequences and its queries: returns
    (SeqDB, q200, q500, batch).  db.extra["hom_owner"] holds, per sequence,
    the workload query it is a homolog of (-1 for background)."""
    from .db import SeqDB
    q200, q500, batch = bench_queries(query_seed)
    qs, owner = bench_query_list(q200, q500, batch)
    db = SeqDB.synthetic(n, qs, seed=seed, device=device, hom_period=BENCH_HOM_PERIOD)
    hq = db.extra["hom_query"]
    own = torch.from_numpy(owner.astype(np.int32)).to(hq.device)
    db.extra["hom_owner"] = torch.where(hq >= 0, own[hq.clamp_min(0).long()], -1)
    db.extra["gen_queries"] = qs
    return db, q200, q500, batch
