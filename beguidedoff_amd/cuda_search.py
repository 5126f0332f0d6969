"""Drop-in replacement of the reference's GPU scan driver.

Same module-level names and the same call contract as
``beguidedoff/cuda_search.py`` (``MAX_GUIDES``, ``MAX_PAM_LEN``,
``encode_pam_packed``, ``search_genome_cuda``), so a caller written for the
reference (``finder.py:92-98,179-191``) works unchanged.  The scan itself runs
in hand-written gfx950 kernels of ``libbgo_hip.so`` (C ABI ``bgo_scan_plain``);
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import threading
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib

MAX_GUIDES = 2048
MAX_PAM_LEN = 6
_DEFAULT_RESULTS_PER_GUIDE = 10240


def _pam_nibble_fields(nib: int) -> Optional[Tuple[int, int]]:
    """(pattern2, compare2) for a 4-bit allowed-base set, or None when the set
    is not expressible as fixed bits of the 2-bit code
    (beguidedoff/cuda_search.py:101-121).

    One base -> both bits fixed; N -> none; R/Y (same bit 0) -> bit 0 fixed;
    M/K (same bit 1) -> bit 1 fixed; S, W and the 3-base codes -> None.
    """
    bases = [b for b in range(4) if nib & (1 << b)]
    if len(bases) == 1:
        return bases[0], 0b11
    if len(bases) == 4:
        return 0, 0b00
    if len(bases) == 2:
        lo, hi = bases
        if (lo & 1) == (hi & 1) and (lo >> 1) != (hi >> 1):
            return lo & 1, 0b01
        if (lo >> 1) == (hi >> 1) and (lo & 1) != (hi & 1):
            return lo & 2, 0b10
    return None


def encode_pam_packed(pam_masks: np.ndarray, pam_lens, pam_3prime) -> Tuple[np.ndarray, np.ndarray]:
    """Per-guide packed PAM test fields (beguidedoff/cuda_search.py:124-160).

    pats[g] = pattern (bits 0-11) | plen << 12 | exotic << 15 | 3prime << 16
    cmps[g] = compared bits (0 for exotic guides).  plen 0 or > 6 leaves the
    guide disabled (pats = 0).
    """
    masks = np.ascontiguousarray(pam_masks, dtype=np.uint8)
    n = int(masks.shape[0])
    pats = np.zeros(n, dtype=np.uint32)
    cmps = np.zeros(n, dtype=np.uint32)
    for g in range(n):
        plen = int(pam_lens[g])
        if plen == 0 or plen > MAX_PAM_LEN:
            continue
        pat = cmp = 0
        exotic = False
        for i in range(plen):
            fields = _pam_nibble_fields(int(masks[g, i]) & 0xF)
            if fields is None:
                exotic = True
                break
            pat |= fields[0] << (2 * i)
            cmp |= fields[1] << (2 * i)
        info = (pat & 0xFFF) | (plen << 12)
        if exotic:
            info |= 1 << 15
            cmp = 0
        if pam_3prime[g]:
            info |= 1 << 16
        pats[g] = info
        cmps[g] = cmp
    return pats, cmps


class DeviceStrand:
    """One strand array (+ N mask) resident on the GPU (bgo_strand)."""

    def __init__(self, genome_packed: np.ndarray, total_bp: int,
                 n_mask_packed: Optional[np.ndarray] = None):
        lib = _lib.load()
        words = np.ascontiguousarray(genome_packed, dtype=np.uint32)
        mask = None if n_mask_packed is None else np.ascontiguousarray(n_mask_packed, dtype=np.uint8)
        h = C.c_void_p()
        _lib.check(lib.bgo_strand_create(
            _lib.ptr(words), words.size, _lib.ptr(mask), 0 if mask is None else mask.size,
            int(total_bp), C.byref(h)))
        self._h = h
        self.total_bp = int(total_bp)
        self.has_n_mask = mask is not None and mask.size > 0

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise RuntimeError("strand released")
        return self._h

    def close(self) -> None:
        if self._h is not None:
            _lib.load().bgo_strand_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Upload cache keyed by the identity of the host arrays, like the
# reference's _GENOME_DEVICE_CACHE (cuda_search.py:321-362), so overflow
# re-scans do not re-upload the genome.  Eviction only drops the cache's
# references, as the reference's does: a strand that a running call still
# holds stays alive, and its device memory is released when the last
# reference goes (DeviceStrand.__del__ -> bgo_strand_destroy, which also
# waits for a scan in progress on the handle).
_CACHE: Dict[Tuple[int, int], Tuple[np.ndarray, Optional[np.ndarray], DeviceStrand]] = {}
_CACHE_LIMIT = 4
_CACHE_LOCK = threading.Lock()


def _device_strand(genome_packed, total_bp, n_mask_packed) -> DeviceStrand:
    key = (id(genome_packed), id(n_mask_packed) if n_mask_packed is not None else 0)
    with _CACHE_LOCK:
        hit = _CACHE.get(key)
        if (hit is not None and hit[0] is genome_packed and hit[1] is n_mask_packed
                and hit[2].total_bp == int(total_bp)):
            return hit[2]
        if len(_CACHE) >= _CACHE_LIMIT:
            _CACHE.clear()
        strand = DeviceStrand(genome_packed, total_bp, n_mask_packed)
        _CACHE[key] = (genome_packed, n_mask_packed, strand)
        return strand


def clear_device_cache() -> None:
    """Drop the cached uploads; strands still held by a running call are
    released when that call returns."""
    with _CACHE_LOCK:
        _CACHE.clear()


def last_kernel_ms() -> float:
    """Kernel time of the last search_genome_cuda call on this thread."""
    ms = C.c_float(0.0)
    _lib.check(_lib.load().bgo_scan_plain_last_ms(C.byref(ms)))
    return float(ms.value)


def search_genome_cuda(genome_packed, total_bp, guide_encoded, pam_masks, pam_lens, pam_3prime,
                       max_mismatches, chunk_size=200_000_000, return_overflow=False,
                       results_per_guide=_DEFAULT_RESULTS_PER_GUIDE, n_mask_packed=None,
                       n_encoded=0):
    """Scan one strand array for every guide (beguidedoff/cuda_search.py:384-488).

    Returns ``{g: [(pos, n_mm), ...]}``; with ``return_overflow`` also
    ``{g: count > results_per_guide}`` and ``{g: exact raw count}``.  Raises
    RuntimeError for more than MAX_GUIDES guides.  An empty N mask (None or
    size 0) scans ungated.
    """
    guides = np.ascontiguousarray(guide_encoded, dtype=np.uint64)
    ng = int(guides.shape[0])
    if ng > MAX_GUIDES:
        raise RuntimeError(
            f"Too many guides ({ng}) for GPU search. Maximum supported is {MAX_GUIDES}. "
            "Consider splitting the input into smaller batches, or increase MAX_GUIDES in "
            "cuda_search.py.")
    lib = _lib.load()
    masks = np.ascontiguousarray(pam_masks, dtype=np.uint8).reshape(ng, -1)
    if masks.shape[1] < MAX_PAM_LEN:
        masks = np.pad(masks, ((0, 0), (0, MAX_PAM_LEN - masks.shape[1])))
    masks = np.ascontiguousarray(masks[:, :MAX_PAM_LEN])
    pats, cmps = encode_pam_packed(masks, pam_lens, pam_3prime)
    nmask = None
    if n_mask_packed is not None:
        nmask = np.asarray(n_mask_packed, dtype=np.uint8)
        if nmask.size == 0:
            nmask = None
    strand = _device_strand(genome_packed, total_bp, n_mask_packed if nmask is not None else None)
    rpg = int(results_per_guide)
    counts = np.zeros(ng, dtype=np.int32)
    positions = np.empty((ng, rpg), dtype=np.int64)
    mms = np.empty((ng, rpg), dtype=np.int8)
    if ng:
        _lib.check(lib.bgo_scan_plain(
            strand.handle, _lib.ptr(guides), _lib.ptr(masks), _lib.ptr(pats), _lib.ptr(cmps), ng,
            int(max_mismatches), int(chunk_size), rpg, int(n_encoded), _lib.ptr(counts),
            _lib.ptr(positions), _lib.ptr(mms)))
    results: Dict[int, List[Tuple[int, int]]] = {}
    overflow: Dict[int, bool] = {}
    hit_counts: Dict[int, int] = {}
    for g in range(ng):
        cnt = int(counts[g])
        k = min(cnt, rpg)
        overflow[g] = cnt > rpg
        hit_counts[g] = cnt
        results[g] = list(zip(positions[g, :k].tolist(), mms[g, :k].tolist()))
    if return_overflow:
        return results, overflow, hit_counts
    return results
