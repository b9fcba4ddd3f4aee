"""sseqs/sw.py's public names (sseqs/sw.py:9-27)."""
from sseqs_amd.ext import sseqs_sw_ext
from sseqs_amd.sw import sw, sw_affine_backtrack

__all__ = ["sw", "sw_affine_backtrack", "sseqs_sw_ext"]
