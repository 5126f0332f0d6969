"""ORACLE — genome artefact encoder, restated per base (test infrastructure,
not product code).

Follows beguidedoff/genome_converter.py:44-100 step by step for small inputs:
each contig upper-cased (:46), every base mapped through NUC_TO_2BIT with 0
for anything else (:57, encoding.py:6), the N mask 1 where the upper-cased
base is 'N' (:62), the reverse complement as COMPLEMENT_2BIT of the encoded
bases in reverse order (:67-70, encoding.py:11), contigs concatenated, then
packed 16 bases per uint32 with base i at bits 2i and zero padding
(encoding.py:36-46), and the N mask through np.packbits (:96).
It does not import beguidedoff_amd.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np

NUC_TO_2BIT = {"A": 0, "C": 1, "G": 2, "T": 3, "N": 0}
COMPLEMENT_2BIT = {0: 3, 1: 2, 2: 1, 3: 0}


def _pack(codes: List[int]) -> np.ndarray:
    n_words = (len(codes) + 15) // 16
    out = np.zeros(n_words, dtype=np.uint32)
    for i, c in enumerate(codes):
        out[i // 16] |= np.uint32(c << (2 * (i % 16)))
    return out


def _packbits(bits: List[int]) -> np.ndarray:
    out = np.zeros((len(bits) + 7) // 8, dtype=np.uint8)
    for i, b in enumerate(bits):
        if b:
            out[i // 8] |= np.uint8(1 << (7 - i % 8))
    return out


def encode(contigs: List[Tuple[str, str]]) -> Dict[str, np.ndarray]:
    fwd: List[int] = []
    rc: List[int] = []
    nmask: List[int] = []
    offsets, lengths = [], []
    for _, seq in contigs:
        seq = seq.upper()
        offsets.append(len(fwd))
        lengths.append(len(seq))
        enc = [NUC_TO_2BIT.get(c, 0) for c in seq]
        fwd += enc
        nmask += [1 if c == "N" else 0 for c in seq]
        rc += [COMPLEMENT_2BIT[b] for b in reversed(enc)]
    return {
        "genome_forward": _pack(fwd),
        "genome_rc": _pack(rc),
        "n_mask": _packbits(nmask),
        "chrom_offsets": np.array(offsets, dtype=np.int64),
        "chrom_lengths": np.array(lengths, dtype=np.int64),
        "total_bp": np.int64(len(fwd)),
    }
