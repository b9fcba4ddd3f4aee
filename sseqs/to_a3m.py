"""sseqs/to_a3m.py's public name (sseqs/to_a3m.py:5)."""
from sseqs_amd.to_a3m import to_a3m

__all__ = ["to_a3m"]
