"""Per-query host preparation: encodings and the weighted-SVD prefilter LUTs.

Restates sseqs/msa.py:109-160: the prefilter approximates
w_a * BLOSUM62[a][b] (w = sqrt of the query's residue frequencies) by its
leading singular triplets, LUT_A[r] = sqrt(s_r) u_r for the query side and
LUT_B[r] = sqrt(s_r) v_r for the database side, each padded with a 0 column
for unknown residues.  For rank 1 the negative BLOSUM62 entries are zeroed
first (sseqs/msa.py:125).  The SVD is numpy's, as in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .synth import AA, BLOSUM62

AA_MAP = {a: i for i, a in enumerate(AA)}
AA_TO_POS = dict(AA_MAP)
AA_TO_POS.update({"X": 20, "-": 20, "*": 20})


def sw_codes(query: str) -> np.ndarray:
    """Residue codes for Smith-Waterman; non-residues are dropped
    (sseqs/msa.py:156, sseqs/sw.py:11)."""
    return np.array([AA_MAP[c] for c in query if c in AA_MAP], dtype=np.uint8)


def reversed_codes(query: str) -> np.ndarray:
    """Prefilter codes of the reversed query, 20 for anything else
    (sseqs/msa.py:157)."""
    return np.array([AA_TO_POS.get(c, 20) for c in query[::-1]], dtype=np.uint8)


def weighted_luts(query: str, fft_rank: int = 4) -> tuple[np.ndarray, np.ndarray]:
    """(LUT_A, LUT_B), each 20 x 21 float32 (sseqs/msa.py:125-145)."""
    B = BLOSUM62.copy()
    if fft_rank == 1:
        B[B < 0] = 0
    w = np.sqrt(np.array([query.count(a) for a in AA]) / len(query))
    U, S, Vt = np.linalg.svd(w[:, None] * B)
    A = np.stack([np.append(np.sqrt(s) * u, 0.0) for s, u in zip(S, U.T)]).astype(np.float32)
    Bt = np.stack([np.append(np.sqrt(s) * v, 0.0) for s, v in zip(S, Vt)]).astype(np.float32)
    return A, Bt


@dataclass
class PreparedQuery:
    """Device-resident per-query inputs of the search kernels."""

    text: str
    codes: torch.Tensor      # uint8 SW codes (non-residues dropped)
    qrev: torch.Tensor       # uint8 reversed prefilter codes
    lut_a: torch.Tensor      # float32 [20, 21]
    lut_b: torch.Tensor      # float32 [20, 21]
    lut_b_flat: torch.Tensor  # float32 [20*21 + 1]: flat LUT_B plus a 0 that code 21
                              # of the last row reads (see launch_warp_corr)
    k: int                   # prefilter rank actually used
    qbytes: torch.Tensor     # uint8 raw query text (the backtrack scores non-residues -5)

    @property
    def length(self) -> int:
        return len(self.text)


def prepare(query: str, fft_rank: int = 4, device="cuda") -> PreparedQuery:
    if not query:
        raise ValueError("empty query")
    A, Bt = weighted_luts(query, fft_rank)
    k = max(1, min(int(fft_rank), A.shape[0]))  # sseqs/msa.py:262
    dev = torch.device(device)
    lut_b = np.concatenate([Bt.reshape(-1), np.zeros(1, np.float32)])
    return PreparedQuery(
        text=query,
        codes=torch.from_numpy(sw_codes(query)).to(dev),
        qrev=torch.from_numpy(reversed_codes(query)).to(dev),
        lut_a=torch.from_numpy(A).to(dev),
        lut_b=torch.from_numpy(Bt).to(dev),
        lut_b_flat=torch.from_numpy(lut_b).to(dev),
        k=k,
        qbytes=torch.from_numpy(np.frombuffer(query.encode("latin-1", errors="replace"),
                                              dtype=np.uint8).copy()).to(dev),
    )
