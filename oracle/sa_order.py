"""ORACLE — test infrastructure, not product code.

The reference fast mode's anchor-occurrence cap, restated by brute force:
the FM-index returns an anchor's occurrences as its suffix-array interval and
keeps the first `max_hits` entries (csrc/fmindex_binding.cpp:141-176,
fast/finder.py:149-159); the indexed text is the contigs' bases as symbols
A C G T = 1 2 3 4, every other byte (N included) = 1, a separator 6 between
contigs (fast/fmindex.py:63-128, fast/backend.py:28-49), and the terminal 0
the suffix-array construction appends.  Here the suffixes are sorted
literally (Python byte-string order), so only small genomes are practical.

Parity: pinned by the definition of a suffix array (sorted suffixes) on
tiny inputs and by hand-derived known answers in tests/test_sa_order.py.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

_SYM = {"A": 1, "C": 2, "G": 3, "T": 4}


def index_text(contigs: Sequence[Tuple[str, str]]) -> Tuple[bytes, List[int]]:
    """(text, text offset of each contig) of fast/fmindex.py:63-128."""
    parts, offs, t = [], [], 0
    for k, (_, seq) in enumerate(contigs):
        if k:
            parts.append(bytes([6]))
            t += 1
        offs.append(t)
        parts.append(bytes(_SYM.get(c, 1) for c in seq.upper()))
        t += len(seq)
    return b"".join(parts) + bytes([0]), offs


def capped_occurrences(contigs: Sequence[Tuple[str, str]], anchor: str,
                       max_hits: int) -> Tuple[List[int], int]:
    """(the first `max_hits` occurrences of `anchor` in suffix-array order, as
    0-based genome positions of the separator-free concatenation; the total
    occurrence count).  max_hits <= 0 keeps every occurrence."""
    text, offs = index_text(contigs)
    pat = bytes(_SYM.get(c, 1) for c in anchor.upper())
    occ = [i for i in range(len(text) - len(pat) + 1) if text[i:i + len(pat)] == pat]
    order = sorted(occ, key=lambda i: text[i:])
    kept = order if max_hits <= 0 else order[:max_hits]
    # text position -> genome position: subtract the separators before it
    out = []
    for t in kept:
        k = max(j for j, o in enumerate(offs) if o <= t)
        out.append(t - k)
    return out, len(occ)
