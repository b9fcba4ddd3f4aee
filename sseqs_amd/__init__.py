"""MI355X-native MSA search with the API of sseqs (sseqs/__init__.py:1-2).

``msa`` and ``sw`` keep the reference's signatures; the extension-operator
mirrors are in ``sseqs_amd.ops`` (``xbit``, ``conv``, ``sw``).  The HIP
library is loaded on the first operator call; importing the package needs
no GPU.
"""
from .msa import msa
from .sw import sw

__all__ = ["msa", "sw"]
