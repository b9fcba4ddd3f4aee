"""Placement of .xbit files in the HBM stream (no GPU needed)."""
from sseqs_amd.db import CHUNK_BYTES, xbit_layout


def test_single_file_whole_and_chunked():
    lay = xbit_layout([12], [None])
    assert (lay.starts, lay.sizes, lay.total) == ([0], [12], 15)
    lay = xbit_layout([3 * CHUNK_BYTES + 7], [2])
    assert (lay.sizes, lay.total) == ([2 * CHUNK_BYTES], 2 * CHUNK_BYTES)
    # fewer bytes than the chunk count asks for: read to the end
    lay = xbit_layout([11], [4], chunk_bytes=5)
    assert (lay.sizes, lay.total) == ([11], 15)


def test_second_file_starts_on_symbol_boundary():
    lay = xbit_layout([13, 8], [None, None])
    assert lay.starts == [0, 15] and lay.sizes == [13, 8] and lay.total == 25
    assert all(st % 5 == 0 for st in lay.starts)
    # 1 GB chunks are whole 8-symbol groups already (1e9 = 5 * 2e8)
    lay = xbit_layout([5 * CHUNK_BYTES, 40], [3, 23])
    assert lay.starts == [0, 3 * CHUNK_BYTES] and lay.total == 3 * CHUNK_BYTES + 40
