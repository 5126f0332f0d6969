"""IUPAC PAM masks (beguidedoff/pam.py:7-45).

A PAM position is a 4-bit set of allowed bases: bit 0 = A, 1 = C, 2 = G, 3 = T.
"""
from collections import defaultdict
from itertools import product
from typing import Dict, List, Tuple

import numpy as np

IUPAC_MASK = {
    "A": 0b0001, "C": 0b0010, "G": 0b0100, "T": 0b1000,
    "R": 0b0101, "Y": 0b1010, "M": 0b0011, "K": 0b1100,
    "S": 0b0110, "W": 0b1001, "H": 0b1011, "B": 0b1110,
    "V": 0b0111, "D": 0b1101, "N": 0b1111,
}
VALID_IUPAC = set(IUPAC_MASK)
MAX_PAM_LEN = 6


def validate_pam(pam: str) -> None:
    pam = pam.upper()
    if not 1 <= len(pam) <= MAX_PAM_LEN:
        raise ValueError(f"PAM must be 1-6 characters, got {len(pam)}")
    bad = set(pam) - VALID_IUPAC
    if bad:
        raise ValueError(f"Invalid IUPAC codes in PAM: {bad}")


def encode_pam_mask(pam: str) -> np.ndarray:
    """PAM string -> uint8[6] allowed-base nibbles, zero padded."""
    pam = pam.upper()
    validate_pam(pam)
    mask = np.zeros(MAX_PAM_LEN, dtype=np.uint8)
    mask[: len(pam)] = [IUPAC_MASK[c] for c in pam]
    return mask


def pam_matches_at_position(genome_2bit, position: int, pam_mask, pam_len: int) -> bool:
    for i in range(pam_len):
        g = position + i
        if g >= len(genome_2bit) or not ((1 << int(genome_2bit[g])) & int(pam_mask[i])):
            return False
    return True


def group_guides_by_pam(guides: List) -> Dict[Tuple[str, str], List]:
    groups: Dict[Tuple[str, str], List] = defaultdict(list)
    for g in guides:
        groups[(g.pam.upper(), g.pam_position.lower())].append(g)
    return dict(groups)


def all_possible_pams(pam: str) -> List[str]:
    pam = pam.upper()
    validate_pam(pam)
    choices = [[b for b, bit in zip("ACGT", (1, 2, 4, 8)) if IUPAC_MASK[c] & bit] for c in pam]
    return ["".join(p) for p in product(*choices)]
