"""The reference's extension modules ``xbit``, ``conv`` and ``sw`` (also
``sseqs_sw_ext``) as a compiled PyTorch extension (csrc/torch_ext.cpp) over
the C ABI of libsseqs_amd.so: the form in which the reference JIT-builds
and imports them (sseqs/msa.py:33-35, sseqs/sw.py:9), with the same
function names, argument names and defaults (sseqs/xbit_bind.cpp:33-36,
sseqs/filter_bind.cpp:67-82, sseqs/sw_bind.cpp:111-132).  Built in-tree by
``python -m sseqs_amd.build``; importing it without the build raises.
"""
import torch  # noqa: F401  (loads the torch libraries the extension links)

try:
    from . import _torch_ext
except ImportError as e:  # no fallback: the extension is the implementation
    raise ImportError("sseqs_amd/_torch_ext.so is missing or stale: build it with "
                      "`python -m sseqs_amd.build`") from e

xbit = _torch_ext.xbit
conv = _torch_ext.conv
sw = _torch_ext.sw
sseqs_sw_ext = sw

__all__ = ["xbit", "conv", "sw", "sseqs_sw_ext"]
