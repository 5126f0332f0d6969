"""encode_pam_packed (beguidedoff/cuda_search.py:101-160) is a re-expression
of the IUPAC allowed-set test: pinned by exhaustive enumeration."""
from itertools import product

import numpy as np

from beguidedoff_amd.cuda_search import _pam_nibble_fields, encode_pam_packed
from beguidedoff_amd.pam import IUPAC_MASK, encode_pam_mask


def packed_ok(pat_info, cmp, codes):
    pam = 0
    for i, c in enumerate(codes):
        pam |= c << (2 * i)
    return ((pam ^ pat_info) & cmp) == 0


def test_nibble_fields_classes():
    assert _pam_nibble_fields(0b0001) == (0, 0b11)
    assert _pam_nibble_fields(0b1111) == (0, 0b00)
    assert _pam_nibble_fields(IUPAC_MASK["R"]) == (0, 0b01)
    assert _pam_nibble_fields(IUPAC_MASK["Y"]) == (1, 0b01)
    assert _pam_nibble_fields(IUPAC_MASK["M"]) == (0, 0b10)
    assert _pam_nibble_fields(IUPAC_MASK["K"]) == (2, 0b10)
    for c in "SWHBVD":
        assert _pam_nibble_fields(IUPAC_MASK[c]) is None


def test_every_nibble_every_base():
    for nib in range(16):
        f = _pam_nibble_fields(nib)
        for b in range(4):
            allowed = bool(nib & (1 << b))
            if f is None:
                continue
            assert (((b ^ f[0]) & f[1]) == 0) == allowed


def test_packed_equals_iupac_for_pams():
    rng = np.random.default_rng(0)
    codes = list(IUPAC_MASK)
    pams = ["NGG", "NRG", "NNGRRT", "NNNRRT", "NG", "TTTV", "A", "KMRY"]
    pams += ["".join(rng.choice(codes, size=int(rng.integers(1, 7)))) for _ in range(40)]
    masks = np.stack([encode_pam_mask(p) for p in pams])
    pats, cmps = encode_pam_packed(masks, [len(p) for p in pams], [1] * len(pams))
    for p, info, cmp in zip(pams, pats.tolist(), cmps.tolist()):
        assert (info >> 12) & 7 == len(p) and (info >> 16) & 1 == 1
        exotic = any(_pam_nibble_fields(IUPAC_MASK[c]) is None for c in p)
        assert bool((info >> 15) & 1) == exotic
        if exotic:
            assert cmp == 0
            continue
        windows = product(range(4), repeat=len(p)) if len(p) <= 4 else (
            tuple(rng.integers(0, 4, size=len(p))) for _ in range(300))
        for w in windows:
            want = all(IUPAC_MASK[c] & (1 << b) for c, b in zip(p, w))
            assert packed_ok(info & 0xFFF, cmp, w) == want


def test_disabled_lengths():
    pats, cmps = encode_pam_packed(np.zeros((2, 6), np.uint8), [0, 7], [1, 1])
    assert pats.tolist() == [0, 0]
