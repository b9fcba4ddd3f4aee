"""``import sseqs`` for code written against the reference package
(sseqs/__init__.py:1-2): ``sseqs.msa`` and ``sseqs.sw`` are the MI355X
implementations in ``sseqs_amd``; the submodules ``sseqs.msa``,
``sseqs.sw`` and ``sseqs.to_a3m`` name the same functions as the
reference's modules (and the extension modules as compiled HIP)."""
from .msa import msa
from .sw import sw

__all__ = ["msa", "sw"]
