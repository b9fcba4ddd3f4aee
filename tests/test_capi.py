"""The C-ABI library loads without a GPU and exports exactly what
include/sseqs_amd.h declares (no compute calls here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "sseqs_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "return", "sizeof")))


def test_header_parses():
    names = header_functions()
    assert "launch_sw_cuda_affine" in names and "sseqs_topk" in names
    assert len(names) >= 15


def test_library_exports_every_symbol():
    from sseqs_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding table covers the same set
    assert sorted(_lib.SIGNATURES) == header_functions()
    assert lib.sseqs_abi_version() == 2


def test_argument_errors_need_no_gpu():
    """Host-side argument checks return before any launch."""
    from sseqs_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load()
    assert lib.launch_sw_cuda_affine(None, 3000, None, None, None, 5, None, 11, 1, None) == -1
    assert b"query length" in lib.sseqs_last_error()
    assert lib.launch_sw_cuda_blosum62(None, 10, None, None, None, 5, None, 3, None) == -1
    assert lib.sseqs_topk(None, 0, 10, 11, 1, None, None, 0, None) == -1
    assert lib.launch_warp_corr(None, None, None, None, None, None, None, None, 4, 21, 4, 10, 10,
                                100, None, None) == -1
    assert b"power of two" in lib.sseqs_last_error()
    # segmented selects: bad sizes, dtype and the sorted limit
    assert lib.sseqs_topk_segments(None, 0, None, None, None, None, -1, 1, 0, 0, 0, 0, None,
                                   None) == -1
    assert lib.sseqs_topk_segments(None, 7, None, None, None, None, 3, 1, 0, 0, 0, 0, None,
                                   None) == -1
    assert lib.sseqs_topk_segments(None, 0, None, None, None, None, 3, 5000, 1, 0, 0, 0, None,
                                   None) == -1
    assert b"4096" in lib.sseqs_last_error()
    assert lib.sseqs_topk_segments(None, 0, None, None, None, None, 3, 5, 0, 0, 0, 0, None,
                                   None) == -1
    assert b"null" in lib.sseqs_last_error()
    # nothing to do is not an error
    assert lib.sseqs_topk_segments(None, 0, None, None, None, None, 0, 5, 0, 0, 0, 0, None,
                                   None) == 0
    # resident-database SW: query length and gap sign
    assert lib.sseqs_db_sw_linear(None, None, None, None, 5, None, 0, -1, None, None) == -1
    assert lib.sseqs_db_sw_linear(None, None, None, None, 5, None, 10, 2, None, None) == -1
    assert lib.sseqs_db_sw_affine(None, None, None, None, 5, None, 2049, 11, 1, None, None) == -1
    assert lib.sseqs_db_sw_affine(None, None, None, None, -1, None, 20, 11, 1, None, None) == -1
    assert lib.sseqs_db_index(None, 0, None, None, 0, None, None, 0, None) == -1
    assert lib.sseqs_db_index(None, 100, None, None, 0, None, None, 0, None) == -1
    assert b"null" in lib.sseqs_last_error()
    # a misaligned stream is refused before anything is launched
    assert lib.sseqs_db_index(3, 100, None, None, 0, 8, 8, 1 << 20, None) == -1
    assert b"aligned" in lib.sseqs_last_error()
    assert lib.sseqs_db_index(64, 100, None, None, 0, 8, 8, 8, None) == -2
    assert lib.sseqs_db_index_ws_bytes(1 << 40) > lib.sseqs_db_index_ws_bytes(1 << 20)


def test_prefilter_kernel_plan_needs_no_gpu():
    """The prefilter's kernel choice per length range (host-side plan)."""
    from sseqs_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load()
    buf = ctypes.create_string_buffer(64)

    def name(lq, k, n):
        assert lib.sseqs_db_prefilter_kernel(lq, k, n, n, buf, 64) == 0
        return buf.value.decode()

    # q_len 200: every one of the 8 diagonals is nonzero at N = 256, 8 of 16 at 512
    assert name(200, 4, 256) == "k_corr_band<8, true, true, 1>"
    assert name(200, 4, 512) == "k_corr_band<8, false, true, 1>"
    assert name(200, 4, 1024) == "k_corr_band<8, false, false, 1>"
    assert name(500, 4, 512) == "k_corr_band<16, true, true, 1>"
    assert name(500, 4, 1024).startswith("k_corr_band<18, false")
    assert name(200, 3, 256) == "k_corr_db<4>"
    assert name(50, 4, 64) == "k_corr_db4<1>"
    assert lib.sseqs_db_prefilter_kernel(200, 4, 300, 300, buf, 64) == -1
