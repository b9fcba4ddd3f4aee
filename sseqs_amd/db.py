"""The HBM-resident sequence database.

The reference keeps its ``.xbit`` database in pinned host RAM and streams it
over PCIe, 1 GB chunk by chunk, for every search (sseqs/msa.py:13-64,
196-208), re-finding the delimiters of every batch on the GPU
(sseqs/msa.py:227-267).  Here the packed bytes are copied to HBM once and an
index of every sequence (symbol offset, length) is built once by the
delimiter-scan kernel.  The index is kept in ascending length order (a
permutation of the file order when the file is not already sorted) so that
the prefilter can run one correlation size per contiguous range and the
MFMA tiles hold sequences of similar length.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_of

# Tail bytes after the stream: symbol reads touch 2 bytes and the prefilter
# copies 16-byte chunks up to 40 bytes past the last symbol's byte.
PAD = 64
# The reference reads its database in chunks of 1 GB (sseqs/msa.py:50-55).
CHUNK_BYTES = 1_000_000_000


@dataclass
class XbitLayout:
    starts: list[int]   # byte offset of each file's data in the stream
    sizes: list[int]    # bytes read from each file
    total: int          # stream bytes (a multiple of 5: whole 8-symbol groups)


def xbit_layout(sizes: list[int], chunks: list[int | None], chunk_bytes: int = CHUNK_BYTES
                ) -> XbitLayout:
    """Placement of several .xbit files in one packed stream.  File i
    contributes its first chunks[i] * chunk_bytes bytes (all of it for None)
    and is padded with zero bits to a multiple of 5 bytes, so the next file
    starts on a symbol boundary; the zero symbols decode as '@' (empty runs,
    which the index skips)."""
    starts, used, pos = [], [], 0
    for size, c in zip(sizes, chunks):
        m = size if c is None else min(size, int(c) * int(chunk_bytes))
        starts.append(pos)
        used.append(m)
        pos += (m + 4) // 5 * 5
    return XbitLayout(starts, used, pos)


@dataclass
class SeqDB:
    packed: torch.Tensor            # uint8, 5-bit stream + PAD bytes
    n_syms: int                     # symbols in the stream
    seq_off: torch.Tensor           # int64 [n], symbol offset, length order
    seq_len: torch.Tensor           # int32 [n], length order
    orig_id: torch.Tensor | None    # int64 [n] file-order id, None when identity
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.seq_len.numel())

    @property
    def device(self) -> torch.device:
        return self.packed.device

    @property
    def nbytes(self) -> int:
        """Packed bytes of the database (what one sweep streams)."""
        return (self.n_syms * 5 + 7) // 8

    @property
    def residues(self) -> int:
        return int(self.extra.get("residues", 0)) or int(self.seq_len.sum().item())

    # ------------------------------------------------------------ ranges --
    def length_ranges(self, Lq: int, nmax: int = 2048) -> list[tuple[int, int, int]]:
        """[(N, begin, end)] with one correlation size N = nextpow2(max(Lq, L))
        per contiguous range of the length-sorted index (sequences longer
        than nmax share N = nmax and are truncated to it)."""
        key = ("ranges", Lq, nmax)
        if key in self.extra:
            return self.extra[key]
        bounds = []
        N = 16
        while N < Lq:
            N <<= 1
        lens = self.seq_len
        begin = 0
        while begin < self.n:
            Ncur = min(N, nmax)
            if Ncur < nmax:
                end = int(torch.searchsorted(lens, torch.tensor(Ncur, dtype=lens.dtype,
                                                                 device=lens.device),
                                             right=True).item())
            else:
                end = self.n
            if end > begin:
                bounds.append((Ncur, begin, end))
                begin = end
            N <<= 1
        self.extra[key] = bounds
        return bounds

    def batch_plan(self, bs: int):
        """Segments of every batch of ``bs`` packed bytes (cached per bs);
        see sseqs_amd.batches."""
        key = ("batch_plan", int(bs))
        if key not in self.extra:
            from .batches import plan_batches
            self.extra[key] = plan_batches(self.seq_off, self.seq_len, self.n_syms, int(bs))
        return self.extra[key]

    # ----------------------------------------------------------- gathers --
    def gather(self, ids: torch.Tensor, width: int = 500) -> tuple[torch.Tensor, torch.Tensor]:
        """ASCII rows (n, width) and lengths of sequences ``ids`` (index order),
        truncated to ``width`` (the reference keeps at most 500 residues per
        hit, sseqs/msa.py:183,367)."""
        ids = ids.to(torch.int64).contiguous()
        n = ids.numel()
        out = torch.zeros((n, width), dtype=torch.uint8, device=self.device)
        ln = torch.empty(n, dtype=torch.int32, device=self.device)
        if n:
            call("sseqs_db_gather", ptr(self.packed), ptr(self.seq_off), ptr(self.seq_len),
                 ptr(ids), n, int(width), ptr(out), ptr(ln), stream_of(self.device))
        return out, ln

    def strings(self, ids: torch.Tensor, width: int = 500) -> list[str]:
        rows, ln = self.gather(ids, width)
        rows, ln = rows.cpu().numpy(), ln.cpu().numpy()
        return [rows[i, : ln[i]].tobytes().decode("latin-1") for i in range(rows.shape[0])]

    def file_ids(self, ids: torch.Tensor) -> torch.Tensor:
        return ids if self.orig_id is None else self.orig_id[ids]

    # ------------------------------------------------------ constructors --
    @staticmethod
    def from_packed(packed: torch.Tensor, n_syms: int | None = None) -> "SeqDB":
        """Index a packed stream already in HBM (every non-'@' run is one
        sequence; empty runs are skipped like sseqs/msa.py:253)."""
        dev = packed.device
        nbytes = packed.numel()
        if n_syms is None:
            n_syms = nbytes * 8 // 5
        if (n_syms * 5 + 7) // 8 + PAD > nbytes:
            buf = torch.zeros((n_syms * 5 + 7) // 8 + PAD, dtype=torch.uint8, device=dev)
            buf[: min(nbytes, buf.numel())] = packed[: buf.numel()]
            packed = buf
        lib = _lib.load()
        ws = _lib.workspace(lib.sseqs_db_index_ws_bytes(n_syms), dev)
        cnt = torch.zeros(1, dtype=torch.int64, device=dev)
        # pass 1 with capacity 0 only counts; pass 2 writes
        dummy = torch.empty(1, dtype=torch.int64, device=dev)
        call("sseqs_db_index", ptr(packed), n_syms, ptr(dummy), ptr(dummy), 0, ptr(cnt), ptr(ws),
             ws.numel(), stream_of(dev))
        n = int(cnt.item())
        off = torch.empty(n, dtype=torch.int64, device=dev)
        ln = torch.empty(n, dtype=torch.int32, device=dev)
        call("sseqs_db_index", ptr(packed), n_syms, ptr(off), ptr(ln), n, ptr(cnt), ptr(ws),
             ws.numel(), stream_of(dev))
        orig = None
        if n > 1 and bool((ln[1:] < ln[:-1]).any().item()):
            orig = torch.sort(ln, stable=True).indices
            off, ln = off[orig].contiguous(), ln[orig].contiguous()
        return SeqDB(packed, int(n_syms), off, ln, orig)

    @staticmethod
    def from_xbit(path: str | os.PathLike, chunks: int | None = None, device="cuda",
                  ab_path: str | os.PathLike | None = None, ab_chunks: int | None = 23,
                  chunk_bytes: int = CHUNK_BYTES) -> "SeqDB":
        """Load an .xbit file (its first ``chunks`` chunks of ``chunk_bytes``,
        like the reference's CHUNKS, sseqs/msa.py:14,50-55) into HBM and
        index it.  ``ab_path``: a second database appended after it (the
        antibody set of sseqs/msa.py:39-40,57-62, first ``ab_chunks`` chunks).
        A file shorter than its chunk count is read to its end (the
        reference left the rest of its pinned chunks uninitialised)."""
        dev = torch.device(device)
        files = [(path, chunks)] + ([(ab_path, ab_chunks)] if ab_path is not None else [])
        sizes = [os.path.getsize(f) for f, _ in files]
        lay = xbit_layout(sizes, [c for _, c in files], chunk_bytes)
        packed = torch.zeros(lay.total + PAD, dtype=torch.uint8, device=dev)
        step = 1 << 28
        host = torch.empty(step, dtype=torch.uint8, pin_memory=True)
        for (f, _), start, size in zip(files, lay.starts, lay.sizes):
            with open(f, "rb") as fh:
                pos = 0
                while pos < size:
                    m = min(step, size - pos)
                    got = fh.readinto(host[:m].numpy())
                    if got != m:
                        raise IOError(f"short read from {f}")
                    packed[start + pos: start + pos + m].copy_(host[:m], non_blocking=False)
                    pos += m
        return SeqDB.from_packed(packed, lay.total * 8 // 5)

    @staticmethod
    def synthetic(n: int, queries: list[str], seed: int = 0, device="cuda",
                  hom_period: int = 20) -> "SeqDB":
        """The benchmark's synthetic database (sseqs_amd.synth), generated
        directly in HBM, already in length order."""
        from .synth import generate_db
        g = generate_db(n, queries, seed=seed, device=device, hom_period=hom_period)
        cls = g.classes
        residues = int((cls.count * cls.length.astype(np.int64)).sum())
        return SeqDB(g.packed, g.n_syms, g.seq_off, g.seq_len, None,
                     extra={"hom_query": g.hom_query, "classes": cls, "residues": residues})


_default_db: SeqDB | None = None


def env_chunks() -> int | None:
    """CHUNKS as the reference reads it (sseqs/msa.py:14): the number of 1 GB
    chunks of the database to load, 4 when unset and at least 1.
    ``CHUNKS=all`` (an addition) loads the whole file and returns None."""
    v = os.environ.get("CHUNKS", "4").strip()
    if v.lower() == "all":
        return None
    c = int(v)
    if c <= 0:
        raise ValueError(f"CHUNKS must be a positive number of 1 GB chunks or 'all', got {v!r}")
    return c


def chunks_tag() -> str:
    """The database-size part of the server's cache folder names
    (server.py:25,72): CHUNKS as loaded by default_db."""
    c = env_chunks()
    return "all" if c is None else str(c)


def default_db(device="cuda") -> SeqDB:
    """Database named by DBPATH (default uniref_bfd_mgy_cf.xbit), first CHUNKS
    GB (sseqs/msa.py:14,18; 4 when unset), followed by the antibody database
    named by AB when set, loaded on first use."""
    global _default_db
    if _default_db is None:
        path = os.environ.get("DBPATH", "uniref_bfd_mgy_cf.xbit")
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"Didn't find MSA database at `{path}` (set DBPATH, or pass db=...)")
        chunks = env_chunks()
        ab = os.environ.get("AB") or None  # sseqs/msa.py:39-40,57-62
        _default_db = SeqDB.from_xbit(path, chunks, device=device, ab_path=ab)
        _default_db.extra["chunks_tag"] = chunks_tag()
    return _default_db


def set_default_db(db: SeqDB | None) -> None:
    global _default_db
    _default_db = db
