""".xbit ingestion at GB scale: a 4.5 GB synthetic database written to a file
and a second (antibody-style) file, loaded through SeqDB.from_xbit with the
reference's CHUNKS and AB handling (sseqs/msa.py:14,39-62): the first
CHUNKS GB of the main file plus the first chunk of the second, each padded
to a whole 5-byte group.  The loaded index must equal the delimiter scan of
the same bytes, the complete runs between the delimiters of sampled windows
(unpacked by the oracle) must be exactly the index entries inside them, and
sampled sequences must read back as written."""
import time

import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu


def test_from_xbit_multi_gb(orc, dev, tmp_path):
    from sseqs_amd import synth
    from sseqs_amd.db import CHUNK_BYTES, SeqDB, xbit_layout
    announce("xbit ingestion: 4.5 GB main file + 0.25 GB second file")
    main = SeqDB.synthetic(25_000_000, [synth.make_query(300, 0, seed=5)], seed=51, device=dev)
    second = SeqDB.synthetic(1_400_000, [synth.make_query(120, 1, seed=5)], seed=52, device=dev)
    p_main, p_ab = tmp_path / "main.xbit", tmp_path / "ab.xbit"
    for db, p in ((main, p_main), (second, p_ab)):
        with open(p, "wb") as fh:
            step = 1 << 28
            for b in range(0, db.nbytes, step):
                fh.write(db.packed[b: min(db.nbytes, b + step)].cpu().numpy().tobytes())
    sizes = [p_main.stat().st_size, p_ab.stat().st_size]
    assert sizes[0] > 4 * CHUNK_BYTES
    chunks = 3
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    db = SeqDB.from_xbit(p_main, chunks=chunks, device=dev, ab_path=p_ab, ab_chunks=1)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    lay = xbit_layout(sizes, [chunks, 1])
    print(f"[ingest] {lay.sizes[0] + lay.sizes[1]} bytes from 2 files, {db.n} sequences indexed "
          f"in {sec:.2f} s ({(lay.sizes[0] + lay.sizes[1]) / sec / 1e9:.2f} GB/s)", flush=True)
    assert lay.sizes == [chunks * CHUNK_BYTES, sizes[1]]
    # the device holds exactly those bytes, each file padded to a 5-byte group
    want = np.zeros(lay.total, dtype=np.uint8)
    with open(p_main, "rb") as fh:
        want[: lay.sizes[0]] = np.frombuffer(fh.read(lay.sizes[0]), dtype=np.uint8)
    with open(p_ab, "rb") as fh:
        want[lay.starts[1]: lay.starts[1] + lay.sizes[1]] = np.frombuffer(fh.read(), np.uint8)
    got = db.packed[: lay.total].cpu().numpy()
    assert np.array_equal(got, want)
    # the index equals a fresh scan of the same bytes
    scan = SeqDB.from_packed(db.packed, db.n_syms)
    assert torch.equal(scan.seq_off, db.seq_off) and torch.equal(scan.seq_len, db.seq_len)
    # sampled windows: the complete runs between delimiters that the oracle
    # finds are exactly the index entries inside the window
    rng = np.random.default_rng(3)
    off = db.seq_off.cpu().numpy()
    ln = db.seq_len.cpu().numpy()
    order = np.argsort(off)
    off_s, ln_s = off[order], ln[order]
    W = 200_000
    for w0 in list(rng.integers(0, db.n_syms - W, 6)) + [lay.starts[1] * 8 // 5 - W // 2]:
        w0 = int(w0) // 8 * 8
        a = orc.xbit_unpack(want, W, w0 * 5 // 8)
        D = np.nonzero(a == 64)[0]
        pieces = [(int(D[i]) + 1, int(D[i + 1] - D[i] - 1)) for i in range(D.size - 1)
                  if D[i + 1] - D[i] > 1]
        i0, i1 = np.searchsorted(off_s, w0 + D[0] + 1), np.searchsorted(off_s, w0 + D[-1])
        runs = [(int(o - w0), int(L)) for o, L in zip(off_s[i0:i1], ln_s[i0:i1])]
        assert pieces == runs and len(runs) > 100, w0
    # sampled sequences of the main file read back as the generator wrote them
    cut = lay.sizes[0] * 8 // 5 - 2000
    sel = np.sort(rng.choice(np.nonzero(off < cut)[0], 300, replace=False))
    m_off, m_len = main.seq_off.cpu().numpy(), main.seq_len.cpu().numpy()
    j = np.searchsorted(m_off, off[sel])
    assert np.array_equal(m_off[j], off[sel]) and np.array_equal(m_len[j], ln[sel])
    assert db.strings(torch.from_numpy(sel).to(dev), 1100) == \
        main.strings(torch.from_numpy(j).to(dev), 1100)
