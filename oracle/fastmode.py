"""CPU BASELINE — the reference's fast mode (FM-index anchors at a CCTOP
cutoff), restated in C (oracle/bgo_fast.c) over this directory's own
exact-match index.  Test / measurement infrastructure: used by tests and by
bench.py's CPU baseline leg, never by the product.

Restates beguidedoff/fast/finder.py:163-277 (candidates), :304-441 (per-guide
filters) and regions.py:80-123 (cold region).  20-nt guides only: for 21-nt
guides the reference's '+' and '-' anchors cover different guide bases
(fast/finder.py:186-189).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import scan as _scan
from .finder_oracle import editable

_CODE = {"A": 0, "C": 1, "G": 2, "T": 3}
_IUPAC = {"A": 1, "C": 2, "G": 4, "T": 8, "R": 5, "Y": 10, "S": 6, "W": 9, "K": 12, "M": 3,
          "B": 14, "D": 13, "H": 11, "V": 7, "N": 15}


def _lib():
    lib = _scan._load()
    if not getattr(lib, "_fast_sigs", False):
        lib.bgo_fast_build.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int,
                                       C.c_int, C.POINTER(C.c_void_p)]
        lib.bgo_fast_build.restype = C.c_int
        lib.bgo_fast_free.argtypes = [C.c_void_p]
        lib.bgo_fast_free.restype = None
        lib.bgo_fast_bytes.argtypes = [C.c_void_p]
        lib.bgo_fast_bytes.restype = C.c_int64
        lib.bgo_fast_search.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p,
                                        C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.bgo_fast_search.restype = C.c_int
        lib._fast_sigs = True
    return lib


def _p(a: Optional[np.ndarray]):
    return C.c_void_p(0) if a is None else a.ctypes.data_as(C.c_void_p)


def cold_region(cutoff: float, guide_length: int = 20) -> Tuple[int, int]:
    """(first 0-based position, length) of the exact 'cold' block: positions p
    (1-based) with 1.2**p >= cutoff (regions.py:80-123)."""
    cold = [p for p in range(1, guide_length + 1) if 1.2 ** p >= cutoff]
    if not cold:
        raise ValueError(f"search cutoff {cutoff} leaves no exact 'cold' region")
    if cold != list(range(cold[0], cold[-1] + 1)):
        raise ValueError("fast mode requires a contiguous cold region")
    return cold[0] - 1, len(cold)


@dataclass
class FastHits:
    """Candidates of one guide (global forward leftmost, strand 0 '+' / 1 '-',
    2-bit site codes (base i at bits 2i), mismatch mask (bit p-1 = position p),
    score) and the anchor-cap flag."""

    leftmost: np.ndarray
    strand: np.ndarray
    site: np.ndarray
    mm: np.ndarray
    score: np.ndarray
    overflow: bool


class FastIndex:
    """Exact-match anchor index of a genome artefact's forward strand."""

    def __init__(self, genome: dict, threads: int = 0):
        self.genome = genome
        self._words = np.ascontiguousarray(genome["genome_forward"], dtype=np.uint32)
        self._offsets = np.ascontiguousarray(genome["chrom_offsets"], dtype=np.int64)
        self._lengths = np.ascontiguousarray(genome["chrom_lengths"], dtype=np.int64)
        nm = genome.get("n_mask_le_fwd")
        self._nmask = None if nm is None else np.ascontiguousarray(nm, dtype=np.uint8)
        h = C.c_void_p()
        t0 = time.perf_counter()
        rc = _lib().bgo_fast_build(_p(self._words), int(genome["total_bp"]), _p(self._offsets),
                                   _p(self._lengths), len(self._offsets), int(threads), C.byref(h))
        if rc != 0:
            raise MemoryError("fast index build failed")
        self.build_s = time.perf_counter() - t0
        self._h = h

    @property
    def nbytes(self) -> int:
        return int(_lib().bgo_fast_bytes(self._h))

    def close(self) -> None:
        if self._h is not None:
            _lib().bgo_fast_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def search(self, guides: Sequence[Tuple[str, str]], cutoff: float = 5.0,
               max_hits: int = 10000, threads: int = 0) -> List[FastHits]:
        """Candidates of every (20-nt guide, PAM) at the cutoff."""
        n = len(guides)
        g = np.zeros((n, 20), np.uint8)
        pm = np.zeros((n, 6), np.uint8)
        pl = np.zeros(n, np.uint8)
        for i, (seq, pam) in enumerate(guides):
            if len(seq) != 20:
                raise ValueError("the fast-mode baseline takes 20-nt guides")
            g[i] = [_CODE.get(c, 0) for c in seq.upper()]
            pl[i] = len(pam)
            pm[i, : len(pam)] = [_IUPAC[c] for c in pam.upper()]
        c0, k = cold_region(cutoff)
        tab = np.array([1.2 ** p for p in range(21)], dtype=np.float64)
        counts = np.zeros(n, np.int64)
        ovf = np.zeros(n, np.uint8)
        lib = _lib()
        args = (self._h, _p(self._nmask), _p(g), _p(pm), _p(pl), n, c0, k, float(cutoff), _p(tab),
                int(max_hits), int(threads), _p(counts))
        if lib.bgo_fast_search(*args, None, None, None, None, None, _p(ovf)) != 0:
            raise ValueError("cold region shorter than the index key")
        tot = int(counts.sum())
        left = np.empty(tot, np.int64)
        site = np.empty(tot, np.uint64)
        mm = np.empty(tot, np.uint32)
        strand = np.empty(tot, np.int8)
        score = np.empty(tot, np.float64)
        lib.bgo_fast_search(*args, _p(left), _p(site), _p(mm), _p(strand), _p(score), _p(ovf))
        out, b = [], 0
        for i in range(n):
            e = b + int(counts[i])
            out.append(FastHits(left[b:e], strand[b:e], site[b:e], mm[b:e], score[b:e],
                                bool(ovf[i])))
            b = e
        return out


_BASES = "ACGT"


def site_string(code: int) -> str:
    return "".join(_BASES[(int(code) >> (2 * i)) & 3] for i in range(20))


def any_editable_row(hits: FastHits, editor_type: str, start: int, end: int,
                     on_target: Optional[Tuple[int, str]] = None) -> bool:
    """The fast search's "this guide has an output row" with return_all off:
    an editable, non-on-target candidate (fast/finder.py:379-417).
    on_target = (global forward leftmost, '+' / '-') or None."""
    for lm, st, sv in zip(hits.leftmost.tolist(), hits.strand.tolist(), hits.site.tolist()):
        if on_target is not None and (lm, "+" if st == 0 else "-") == on_target:
            continue
        if editable(site_string(sv), editor_type, start, end):
            return True
    return False
