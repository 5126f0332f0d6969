"""ORACLE — ctypes wrapper of oracle/libbgo_oracle.so (test infrastructure).

``scan_strand`` has the contract of the reference driver
``search_genome_cuda`` (beguidedoff/cuda_search.py:384-488) without result
buffers: every site is returned, ordered by (position, guide).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path
from typing import Optional, Tuple

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libbgo_oracle.so"
_lib = None


SOURCES = ("bgo_oracle.c", "bgo_fast.c")


def build(force: bool = False) -> Path:
    newest = max((HERE / s).stat().st_mtime for s in SOURCES)
    if force or not LIB.exists() or LIB.stat().st_mtime < newest:
        subprocess.run(["make", "-s", "-C", str(HERE), "-B" if force else "libbgo_oracle.so"],
                       check=True)
    return LIB


def _load():
    global _lib
    if _lib is None:
        build()
        lib = C.CDLL(str(LIB))
        P = C.c_void_p
        lib.bgo_oracle_scan.restype = C.c_int64
        lib.bgo_oracle_scan.argtypes = [P, C.c_int64, C.c_int64, P, C.c_int64, P, P, P, P,
                                        C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_int64,
                                        C.c_int32, C.c_int64, P, P, P]
        _lib = lib
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else C.c_void_p(0)


def scan_strand(words: np.ndarray, total_bp: int, guides: np.ndarray, pam_masks: np.ndarray,
                pam_lens: np.ndarray, pam_3prime: np.ndarray, max_mm: int,
                n_mask: Optional[np.ndarray] = None, n_encoded: int = 0,
                pos_begin: int = 0, pos_end: Optional[int] = None,
                threads: Optional[int] = None) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(guide index int32[], position int64[], n_mm int8[]) of every site."""
    lib = _load()
    words = np.ascontiguousarray(words, dtype=np.uint32)
    guides = np.ascontiguousarray(guides, dtype=np.uint64)
    n = guides.size
    masks = np.zeros((n, 6), dtype=np.uint8)
    pm = np.asarray(pam_masks, dtype=np.uint8).reshape(n, -1)
    masks[:, : min(6, pm.shape[1])] = pm[:, :6]
    lens = np.ascontiguousarray(pam_lens, dtype=np.uint8)
    tp = np.ascontiguousarray(pam_3prime, dtype=np.uint8)
    nm = None if n_mask is None else np.ascontiguousarray(n_mask, dtype=np.uint8)
    if nm is not None and nm.size == 0:
        nm = None
    end = total_bp if pos_end is None else pos_end
    T = threads or min(16, os.cpu_count() or 1)
    cap = 1 << 16
    while True:
        og = np.empty(cap, np.int32)
        op = np.empty(cap, np.int64)
        om = np.empty(cap, np.int8)
        got = lib.bgo_oracle_scan(_p(words), words.size, int(total_bp), _p(nm),
                                  0 if nm is None else nm.size, _p(guides), _p(masks), _p(lens),
                                  _p(tp), n, int(max_mm), int(n_encoded), int(pos_begin), int(end),
                                  T, cap, _p(og), _p(op), _p(om))
        if got < 0:
            raise RuntimeError("oracle scan failed")
        if got <= cap:
            return og[:got], op[:got], om[:got]
        cap = int(got)
