"""2-bit nucleotide codec.

Contract of the reference (beguidedoff/encoding.py:6-102): A=0, C=1, G=2,
T=3, any other character (N included) encodes as 0; packed genomes hold 16 nt
per uint32 with base i at bits 2*(i % 16); a guide packs into one integer with
base i at bits 2i.  Implemented with numpy lookup tables.
"""
from typing import List

import numpy as np

NUC_TO_2BIT = {"A": 0, "C": 1, "G": 2, "T": 3, "N": 0}
BIT_TO_NUC = {0: "A", 1: "C", 2: "G", 3: "T"}
COMPLEMENT_2BIT = {0: 3, 1: 2, 2: 1, 3: 0}

_ENC = np.zeros(256, dtype=np.uint8)
for _c, _v in (("A", 0), ("C", 1), ("G", 2), ("T", 3)):
    _ENC[ord(_c)] = _v
_DEC = np.frombuffer(b"ACGT", dtype=np.uint8)
_COMP = {"A": "T", "T": "A", "G": "C", "C": "G", "N": "N"}
_LOW40 = 0x5555555555


def encode_dna(seq: str) -> np.ndarray:
    """DNA string -> uint8 codes (upper-cased first; unknown -> 0)."""
    raw = np.frombuffer(seq.upper().encode("latin-1", "replace"), dtype=np.uint8)
    return _ENC[raw]


_RC_ENC = np.zeros(256, dtype=np.uint8)  # code of a character's complement (others -> 0)
for _c, _v in (("A", 3), ("C", 2), ("G", 1), ("T", 0)):
    _RC_ENC[ord(_c)] = _v


def _char_rows(seqs: List[str], length: int) -> np.ndarray:
    raw = np.frombuffer("".join(seqs).upper().encode("latin-1", "replace"), dtype=np.uint8)
    return raw.reshape(len(seqs), length)


def encode_dna_rows(seqs: List[str], length: int) -> np.ndarray:
    """Equal-length DNA strings -> uint8[n, length] codes, as encode_dna row by row."""
    return _ENC[_char_rows(seqs, length)]


def reverse_complement_rows(seqs: List[str], length: int) -> np.ndarray:
    """uint8[n, length] codes of reverse_complement_seq(s) for equal-length strings."""
    return _RC_ENC[_char_rows(seqs, length)[:, ::-1]]


def decode_dna(encoded: np.ndarray) -> str:
    """uint8 codes (0..3) -> DNA string."""
    codes = np.asarray(encoded, dtype=np.uint8)
    if codes.size and int(codes.max()) > 3:
        return "".join(BIT_TO_NUC.get(int(b), "N") for b in codes)
    return _DEC[codes].tobytes().decode("ascii")


def reverse_complement_2bit(encoded: np.ndarray) -> np.ndarray:
    return (3 - np.asarray(encoded, dtype=np.uint8)[::-1]).astype(np.uint8)


def reverse_complement_seq(seq: str) -> str:
    """Reverse complement; characters other than ACGTN are kept as they are."""
    return "".join(_COMP.get(b, b) for b in reversed(seq.upper()))


def pack_2bit_to_uint32(encoded: np.ndarray) -> np.ndarray:
    """uint8 codes -> uint32 words, 16 nt per word, base i at bits 2(i%16)."""
    codes = np.asarray(encoded, dtype=np.uint32)
    n_words = (codes.size + 15) // 16
    padded = np.zeros(n_words * 16, dtype=np.uint32)
    padded[: codes.size] = codes
    shifts = (2 * np.arange(16, dtype=np.uint32))[None, :]
    return np.bitwise_or.reduce(padded.reshape(n_words, 16) << shifts, axis=1).astype(np.uint32)


def unpack_uint32_to_2bit(packed: np.ndarray, total_bp: int) -> np.ndarray:
    words = np.asarray(packed, dtype=np.uint32)
    shifts = (2 * np.arange(16, dtype=np.uint32))[None, :]
    codes = ((words[:, None] >> shifts) & 3).astype(np.uint8).reshape(-1)
    return codes[:total_bp]


def extract_window(packed: np.ndarray, start_bp: int, length: int) -> np.ndarray:
    """Codes of bases [start_bp, start_bp + length) of a packed array."""
    idx = np.arange(start_bp, start_bp + length, dtype=np.int64)
    words = np.asarray(packed)[idx >> 4].astype(np.uint32)
    return ((words >> ((idx & 15) * 2).astype(np.uint32)) & 3).astype(np.uint8)


def encode_guide_64bit(encoded: np.ndarray) -> int:
    """Codes -> Python int, base i at bits 2i (all bases kept)."""
    val = 0
    for i, b in enumerate(np.asarray(encoded).tolist()):
        val |= int(b) << (2 * i)
    return val


def decode_guide_64bit(val: int, length: int = 20) -> np.ndarray:
    return np.array([(val >> (2 * i)) & 3 for i in range(length)], dtype=np.uint8)


def _mismatch_bits(val1: int, val2: int) -> int:
    x = val1 ^ val2
    return (x | (x >> 1)) & _LOW40


def hamming_distance_2bit(val1: int, val2: int, length: int = 20) -> int:
    """Number of differing bases among the first 20 (the mask is 40 bits)."""
    return bin(_mismatch_bits(val1, val2)).count("1")


def mismatch_positions_2bit(val1: int, val2: int, length: int = 20) -> List[int]:
    """1-based positions i <= length where the 2-bit codes differ."""
    bits = _mismatch_bits(val1, val2)
    return [i + 1 for i in range(length) if (bits >> (2 * i)) & 1]
