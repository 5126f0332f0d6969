"""CCTOP penalty (beguidedoff/scoring.py:5-24): sum of 1.2**p over 1-based
mismatch positions p, summed in the order given (ascending in practice)."""
from typing import List, Sequence

CCTOP_BASE = 1.2


def calculate_cctop_score(mismatch_positions: Sequence[int]) -> float:
    if not mismatch_positions:
        return 0.0
    return sum(CCTOP_BASE ** p for p in mismatch_positions)


def cctop_table() -> List[float]:
    """table[p] = 1.2 ** p for p = 0..20, computed by this host's float pow.

    The GPU epilogue sums table entries in ascending position order, which is
    the same sequence of IEEE double additions as ``calculate_cctop_score``.
    """
    return [CCTOP_BASE ** p for p in range(21)]
