"""sseqs/msa.py's public names: ``msa`` and the extension modules it
builds at import (sseqs/msa.py:33-35), here the compiled HIP extension."""
from sseqs_amd.ext import conv, sw, xbit
from sseqs_amd.msa import msa

__all__ = ["msa", "xbit", "conv", "sw"]
