"""Loader for libsseqs_amd.so (the C ABI declared in include/sseqs_amd.h).

The library is built in-tree by ``python -m sseqs_amd.build``.  There is no
fallback: if the library is missing or a call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libsseqs_amd.so"

F32, F16, I32 = 0, 1, 2  # dtype codes of sseqs_topk

_vp, _i32, _i64, _sz = C.c_void_p, C.c_int, C.c_int64, C.c_size_t


class GenParams(C.Structure):
    """Mirror of sseqs_gen_params (1712 bytes)."""

    _fields_ = [
        ("seed", C.c_uint64),
        ("bg_cdf", C.c_int32 * 20),
        ("sub_cdf", C.c_int32 * 400),
        ("p_indel", C.c_int32),
        ("p_ins", C.c_int32),
        ("p_ext", C.c_int32),
        ("id_lo", C.c_int32),
        ("id_span", C.c_int32),
        ("hom_period", C.c_int32),
    ]


# name -> (restype, argtypes); must match include/sseqs_amd.h
SIGNATURES = {
    "sseqs_last_error": (C.c_char_p, []),
    "sseqs_abi_version": (_i32, []),
    "sseqs_a3m_bound": (_i64, [_i64, _i64, _i64]),
    "sseqs_a3m_text": (_i64, [C.c_char_p, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64]),
    "decompress_launcher": (_i32, [_vp, _vp, _i64, _i64, _vp]),
    "compress_launcher": (_i32, [_vp, _vp, _i64, _vp]),
    "launch_warp_corr": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32,
                                _i32, _i32, _vp, _vp]),
    "launch_sw_cuda_blosum62": (_i32, [_vp, _i32, _vp, _vp, _vp, _i32, _vp, _i32, _vp]),
    "launch_sw_cuda_affine": (_i32, [_vp, _i32, _vp, _vp, _vp, _i32, _vp, _i32, _i32, _vp]),
    "sseqs_db_index_ws_bytes": (_sz, [_i64]),
    "sseqs_db_index": (_i32, [_vp, _i64, _vp, _vp, _i64, _vp, _vp, _sz, _vp]),
    "sseqs_db_prefilter": (_i32, [_vp, _vp, _vp, _i64, _i64, _vp, _i32, _vp, _vp, _i32, _i32, _i32,
                                  _i32, _vp, _vp]),
    "sseqs_db_prefilter_spans": (_i32, [_vp, _vp, _vp, _i64, _i64, _vp, _i32, _vp, _vp, _i32, _i32,
                                        _i32, _i32, _vp, _vp]),
    "sseqs_db_gather": (_i32, [_vp, _vp, _vp, _vp, _i64, _i32, _vp, _vp, _vp]),
    "sseqs_db_prefilter_kernel": (_i32, [_i32, _i32, _i32, _i32, C.c_char_p, _i32]),
    "sseqs_db_sw_affine": (_i32, [_vp, _vp, _vp, _vp, _i64, _vp, _i32, _i32, _i32, _vp, _vp]),
    "sseqs_db_sw_linear": (_i32, [_vp, _vp, _vp, _vp, _i64, _vp, _i32, _i32, _vp, _vp]),
    "sseqs_db_sw_batch": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32,
                                 _i32, _i32, _i32, _vp, _vp, _vp]),
    "sseqs_enc_conv": (_i32, [_vp, _i32, _vp, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _i32,
                              _vp]),
    "sseqs_enc_wgrad": (_i32, [_vp, _i32, _vp, _i32, _i64, _vp, _vp]),
    "sseqs_enc_bias_grad": (_i32, [_vp, _i32, _i64, _vp, _vp]),
    "sseqs_enc_embed": (_i32, [_vp, _i64, _vp, _vp, _vp]),
    "sseqs_enc_embed_grad": (_i32, [_vp, _i64, _vp, _vp, _vp]),
    "sseqs_enc_pool": (_i32, [_vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "sseqs_enc_pool_grad": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "sseqs_enc_pair_loss": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "sseqs_enc_infonce": (_i32, [_vp, _i32, _i32, _vp, C.c_float, C.c_float, _vp, _vp, _vp,
                                 _vp]),
    "sseqs_enc_adam": (_i32, [_vp, _vp, _vp, _vp, _i64, C.c_float, C.c_float, C.c_float, C.c_float,
                              C.c_float, _i32, _vp]),
    "sseqs_enc_pack_weights": (_i32, [_vp, _i32, _i32, _vp, _vp, _vp]),
    "sseqs_enc_db_rows": (_i32, [_vp, _vp, _vp, _i64, _i64, _vp, _i32, _vp, _vp, _vp]),
    "sseqs_enc_to_fp8": (_i32, [_vp, _i64, C.c_float, _vp, _vp]),
    "sseqs_propose": (_i32, [_vp, _i64, _vp, _i32, _vp, _i64, _vp]),
    "sseqs_rows_max": (_i32, [_vp, _i64, _i32, _i64, _vp, _vp]),
    "sseqs_propose_max": (_i32, [_vp, _i64, _vp, _i32, _vp, _i32, _vp, _i64, _vp]),
    "sseqs_topk_ws_bytes": (_sz, [_i64]),
    "sseqs_topk": (_i32, [_vp, _i32, _i64, _i64, _i32, _vp, _vp, _sz, _vp]),
    "sseqs_topk_noalloc_max_n": (_i64, []),
    "sseqs_topk_noalloc": (_i32, [_vp, _i32, _i64, _i64, _i32, _vp, _vp]),
    "sseqs_topk_segments": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32,
                                   _i32, _vp, _vp]),
    "sseqs_backtrack_ws_bytes": (_sz, [_i32, _i64, _i32]),
    "sseqs_backtrack": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32, _vp,
                               _vp, _vp, _vp, _sz, _vp]),
    "sseqs_gen_ws_bytes": (_sz, [_i64, _i32]),
    "sseqs_gen_db": (_i32, [C.POINTER(GenParams), _vp, _vp, _i32, _vp, _vp, _vp, _i32, _i64, _i64,
                            _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
}

_lib = None


class SseqsError(RuntimeError):
    """A C-ABI call returned an error."""


def load() -> C.CDLL:
    """Load the library once (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise SseqsError(
            f"{LIB_PATH} is missing: build it with `python -m sseqs_amd.build` "
            "(the HIP kernels are the only implementation)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().sseqs_last_error().decode(errors="replace")
        raise SseqsError(f"{what} failed ({rc}): {msg}")


def call(name: str, *args) -> None:
    """Call an int-returning entry point and raise on a non-zero return."""
    check(getattr(load(), name)(*args), name)


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def stream_of(device: torch.device | None = None) -> int:
    """Raw hipStream_t of torch's current stream on `device`."""
    return torch.cuda.current_stream(device).cuda_stream


def require_cuda(*tensors: torch.Tensor) -> None:
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise ValueError("sseqs_amd operators take device tensors")


def workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def exported_symbols() -> list[str]:
    return [n for n in SIGNATURES]


if os.environ.get("SSEQS_EAGER_LOAD"):
    load()
