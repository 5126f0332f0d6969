"""Batched persistent scan: Python side of ``bgo_genome`` / ``bgo_plan``.

One launch scans both strands of the whole genome (or a position range of
it) for any mix of guides, PAMs, editors and mismatch budgets, and returns
fully scored site records.  Sites are those of the reference extensive kernel
(beguidedoff/cuda_search.py:163-314) for each guide; the records carry the
finder's per-hit epilogue (beguidedoff/finder.py:435-470).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .base_editing import EDITABLE_BASE, edit_window_mask
from .encoding import encode_dna_rows
from .pam import encode_pam_mask
from .scoring import cctop_table

HIT_DTYPE = _lib.HIT_DTYPE
_EDITOR_CODE = {"A": 0, "C": 1}


@dataclass(frozen=True)
class GuideSpec:
    """One guide of a batched scan.  Only the first 20 nt are compared and the
    PAM is read right after them (the reference's 20-nt layout,
    cuda_search.py:239-240)."""

    seq: str
    pam: str
    max_mm: int
    editor_type: str = "ABE"
    window: Tuple[int, int] = (4, 8)


def guide_descs(specs: Sequence[GuideSpec]) -> np.ndarray:
    """bgo_guide_desc records: the first 20 bases packed 2 bits each (base i
    at bits 2i), PAM nibbles and length, budget, editor and edit window.
    Sequences are encoded per length group and PAMs, editors and windows once
    per distinct value."""
    n = len(specs)
    d = np.zeros(n, dtype=_lib.GUIDE_DESC_DTYPE)
    if n == 0:
        return d
    seqs = [s.seq for s in specs]
    lens = np.array([len(x) for x in seqs], dtype=np.int64)
    packed = np.zeros(n, dtype=np.uint64)
    for L in np.unique(lens).tolist():
        idx = np.flatnonzero(lens == L)
        k = min(L, 20)
        if k == 0:
            continue
        codes = encode_dna_rows([seqs[i] for i in idx], L)[:, :k].astype(np.uint64)
        packed[idx] = (codes << (2 * np.arange(k, dtype=np.uint64))).sum(axis=1, dtype=np.uint64)
    d["seq2bit"] = packed
    mm = np.fromiter((s.max_mm for s in specs), dtype=np.int64, count=n)
    bad = np.flatnonzero((mm < 0) | (mm > 20))
    if bad.size:
        raise ValueError(f"max_mm must be in 0..20, got {specs[int(bad[0])].max_mm}")
    d["max_mm"] = mm
    # PAMs, editors and windows: a table per distinct value, indexed per guide
    pam_ix: Dict[str, int] = {}
    pi = np.fromiter((pam_ix.setdefault(s.pam, len(pam_ix)) for s in specs), dtype=np.int64,
                     count=n)
    d["pam_mask"] = np.stack([encode_pam_mask(p) for p in pam_ix])[pi]
    d["pam_len"] = np.array([len(p) for p in pam_ix], dtype=np.int64)[pi]
    ed_ix: Dict[str, int] = {}
    ei = np.fromiter((ed_ix.setdefault(s.editor_type, len(ed_ix)) for s in specs),
                     dtype=np.int64, count=n)
    codes = []
    for e in ed_ix:
        target = EDITABLE_BASE.get(e)
        if target is None:
            raise ValueError(f"Unknown editor_type: {e}")
        codes.append(_EDITOR_CODE[target])
    d["editor"] = np.array(codes, dtype=np.int64)[ei]
    win_ix: Dict[tuple, int] = {}
    wi = np.fromiter((win_ix.setdefault(tuple(s.window), len(win_ix)) for s in specs),
                     dtype=np.int64, count=n)
    d["edit_mask"] = np.array([edit_window_mask(*w) for w in win_ix], dtype=np.int64)[wi]
    return d


def _tuning_struct(tuning: Optional[dict]) -> Optional[_lib.Tuning]:
    if not tuning:
        return None
    t = _lib.Tuning()
    for k, v in tuning.items():
        if k == "reserved":
            continue
        if k == "interleave":
            t.reserved[0] = int(v)
            continue
        if k == "direct_max":
            t.reserved[1] = int(v)
            continue
        setattr(t, k, int(v))
    return t


class DeviceGenome:
    """Both strand arrays of a genome resident in HBM (bgo_genome)."""

    def __init__(self, genome: dict, n_mask_fwd: Optional[np.ndarray] = None,
                 n_mask_rc: Optional[np.ndarray] = None):
        lib = _lib.load()
        fwd = np.ascontiguousarray(genome["genome_forward"], dtype=np.uint32)
        rc = genome.get("genome_rc")
        rc = None if rc is None else np.ascontiguousarray(rc, dtype=np.uint32)
        if rc is not None and rc.size != fwd.size:
            raise ValueError("genome_rc and genome_forward differ in size")
        self.total_bp = int(genome["total_bp"])
        nbytes = 0
        if n_mask_fwd is not None:
            n_mask_fwd = np.ascontiguousarray(n_mask_fwd, dtype=np.uint8)
            nbytes = n_mask_fwd.size
            if rc is not None:
                if n_mask_rc is None:
                    raise ValueError("an N mask for the reverse-complement array is required")
                n_mask_rc = np.ascontiguousarray(n_mask_rc, dtype=np.uint8)
                if n_mask_rc.size != nbytes:
                    raise ValueError("forward and reverse N masks differ in size")
        h = C.c_void_p()
        _lib.check(lib.bgo_genome_create(
            _lib.ptr(fwd), _lib.ptr(rc) if rc is not None else C.c_void_p(0), fwd.size,
            _lib.ptr(n_mask_fwd) if nbytes else C.c_void_p(0),
            _lib.ptr(n_mask_rc) if (nbytes and rc is not None) else C.c_void_p(0), nbytes,
            self.total_bp, C.byref(h)))
        self._h = h
        self.has_rc = rc is not None
        self.has_n_mask = nbytes > 0

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise RuntimeError("genome released")
        return self._h

    def device_bytes(self) -> int:
        n = C.c_int64(0)
        _lib.check(_lib.load().bgo_genome_device_bytes(self.handle, C.byref(n)))
        return int(n.value)

    def close(self) -> None:
        if self._h is not None:
            _lib.load().bgo_genome_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """A prepared batched scan with its own HIP stream and hit buffer."""

    def __init__(self, genome: DeviceGenome, specs: Sequence[GuideSpec],
                 tuning: Optional[dict] = None, hit_capacity: int = 1 << 20):
        lib = _lib.load()
        self.genome = genome
        self.specs = list(specs)
        self.n_guides = len(self.specs)
        self.descs = guide_descs(self.specs)
        self.score_table = np.array(cctop_table(), dtype=np.float64)
        self.hit_capacity = int(hit_capacity)
        h = C.c_void_p()
        t = _tuning_struct(tuning)
        _lib.check(lib.bgo_plan_create(genome.handle, _lib.ptr(self.descs), self.n_guides,
                                       _lib.ptr(self.score_table),
                                       C.byref(t) if t is not None else None,
                                       self.hit_capacity, C.byref(h)))
        self._h = h

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise RuntimeError("plan released")
        return self._h

    def set_guides(self, specs: Sequence[GuideSpec], descs: Optional[np.ndarray] = None) -> None:
        """Swap the guide set (waits for the previous launch).  `descs` may pass
        the specs' descriptors when the caller already has them."""
        self.specs = list(specs)
        self.n_guides = len(self.specs)
        self.descs = guide_descs(self.specs) if descs is None else np.ascontiguousarray(descs)
        if self.descs.size != self.n_guides:
            raise ValueError("descriptor count differs from the guide count")
        _lib.check(_lib.load().bgo_plan_set_guides(self.handle, _lib.ptr(self.descs),
                                                   self.n_guides))

    def set_end_guard(self, end_guard: int) -> None:
        """Scan positions < total_bp - end_guard (default 26, the reference's
        extensive range)."""
        _lib.check(_lib.load().bgo_plan_set_end_guard(self.handle, int(end_guard)))

    def set_hit_capacity(self, capacity: int) -> None:
        _lib.check(_lib.load().bgo_plan_set_hit_capacity(self.handle, int(capacity)))
        self.hit_capacity = int(capacity)

    def launch(self, pos_begin: int = 0, pos_end: Optional[int] = None, strands: int = 3) -> None:
        end = self.genome.total_bp if pos_end is None else int(pos_end)
        _lib.check(_lib.load().bgo_plan_launch(self.handle, int(pos_begin), end, int(strands)))

    def graph_launch(self, strands: int = 3) -> None:
        _lib.check(_lib.load().bgo_plan_graph_launch(self.handle, int(strands)))

    def wait(self) -> Tuple[int, float]:
        n = C.c_int64(0)
        ms = C.c_float(0.0)
        _lib.check(_lib.load().bgo_plan_wait(self.handle, C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    def fetch(self, n: Optional[int] = None) -> np.ndarray:
        cap = self.hit_capacity if n is None else min(int(n), self.hit_capacity)
        out = np.empty(max(cap, 0), dtype=HIT_DTYPE)
        got = C.c_int64(0)
        _lib.check(_lib.load().bgo_plan_fetch(self.handle, _lib.ptr(out), out.size,
                                              C.byref(got)))
        return out[: got.value]

    def fetch_view(self, n: Optional[int] = None) -> np.ndarray:
        """Like fetch, but the records land in the plan's page-locked buffer by
        direct DMA and the array is a view of it: valid until the plan's next
        fetch_view or close (copy it to keep it longer)."""
        ptr = C.c_void_p()
        got = C.c_int64(0)
        _lib.check(_lib.load().bgo_plan_fetch_pinned(self.handle, -1 if n is None else int(n),
                                                     C.byref(ptr), C.byref(got)))
        if got.value == 0:
            return np.zeros(0, dtype=HIT_DTYPE)
        buf = (C.c_char * (got.value * HIT_DTYPE.itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=HIT_DTYPE, count=got.value)

    def tuning(self) -> dict:
        t = _lib.Tuning()
        _lib.check(_lib.load().bgo_plan_get_tuning(self.handle, C.byref(t)))
        return t.as_dict()

    def family_stats(self) -> List[Tuple[int, int]]:
        """[(guides, PAM-passing windows compared in the last launch)] per family."""
        nf = C.c_int32(0)
        _lib.check(_lib.load().bgo_plan_family_stats(self.handle, 0, C.byref(nf), None, None))
        w = np.zeros(max(nf.value, 1), dtype=np.int64)
        g = np.zeros(max(nf.value, 1), dtype=np.int32)
        _lib.check(_lib.load().bgo_plan_family_stats(self.handle, nf.value, C.byref(nf),
                                                     _lib.ptr(w), _lib.ptr(g)))
        return [(int(g[i]), int(w[i])) for i in range(nf.value)]

    def seed_stats(self) -> List[int]:
        """Seed engine: guide-table entries verified per family in the last launch."""
        nf = len(self.family_stats())
        v = np.zeros(max(nf, 1), dtype=np.int64)
        _lib.check(_lib.load().bgo_plan_seed_stats(self.handle, nf, _lib.ptr(v)))
        return [int(x) for x in v[:nf]]

    def close(self) -> None:
        if self._h is not None:
            _lib.load().bgo_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ScanResult:
    hits: np.ndarray                 # HIT_DTYPE records, unordered
    n_hits: int
    kernel_ms: float
    tuning: dict
    families: List[Tuple[int, int]] = field(default_factory=list)
    launches: int = 1


def default_capacity(n_guides: int) -> int:
    return max(1 << 20, 4096 * n_guides)


# Most site records one launch may hold (32 B each: 512 MB of device memory
# and as much page-locked host memory).  Bounds the exact-size re-run of a
# launch whose sites overflowed its first buffer.
DEFAULT_MAX_RECORDS = 1 << 24


class RecordLimitExceeded(RuntimeError):
    """A launch found more sites than the caller's record limit allows."""

    def __init__(self, n_sites: int, limit: int):
        super().__init__(f"{n_sites} sites exceed the record limit of {limit}")
        self.n_sites = int(n_sites)
        self.limit = int(limit)


def scan(genome: DeviceGenome, specs: Sequence[GuideSpec], tuning: Optional[dict] = None,
         hit_capacity: Optional[int] = None, pos_begin: int = 0,
         pos_end: Optional[int] = None, strands: int = 3,
         end_guard: Optional[int] = None,
         max_records: Optional[int] = DEFAULT_MAX_RECORDS) -> ScanResult:
    """Scan and return every site; re-runs once with an exact-size buffer when
    the first buffer was too small (the kernel counts past its capacity).
    Raises RecordLimitExceeded when the sites exceed `max_records` (None: no
    limit); :func:`scan_capped` keeps such scans within a limit."""
    cap = int(hit_capacity or default_capacity(len(specs)))
    if max_records is not None:
        cap = min(cap, int(max_records))
    launches = 0
    while True:
        plan = Plan(genome, specs, tuning=tuning, hit_capacity=cap)
        try:
            if end_guard is not None:
                plan.set_end_guard(end_guard)
            plan.launch(pos_begin, pos_end, strands)
            n, ms = plan.wait()
            launches += 1
            if n <= cap:
                return ScanResult(plan.fetch(n), n, ms, plan.tuning(), plan.family_stats(),
                                  launches)
            if max_records is not None and n > max_records:
                raise RecordLimitExceeded(n, max_records)
            cap = n
        finally:
            plan.close()


@dataclass
class CappedScan:
    hits: np.ndarray      # kept records, unordered; `guide` indexes the caller's specs
    counts: np.ndarray    # int64[n_guides, 2]: exact sites per guide on '+' and '-'
    launches: int
    kernel_ms: float      # summed over launches

    @property
    def truncated(self) -> np.ndarray:
        """Per guide: more sites on a strand than were kept."""
        n = self.counts.shape[0]
        grp = self.hits["guide"].astype(np.int64) * 2 + self.hits["strand"].astype(np.int64)
        kept = np.bincount(grp, minlength=2 * n).reshape(n, 2)
        return (self.counts > kept).any(axis=1)


def _keep_lowest(hits: np.ndarray, cap: int, sizes: Optional[np.ndarray] = None) -> np.ndarray:
    """Records of each (guide, strand) with the `cap` lowest positions
    (`sizes`, when given, are the records per guide * 2 + strand)."""
    if hits.size == 0 or (sizes is not None and sizes.max(initial=0) <= cap):
        return hits
    grp = hits["guide"].astype(np.int64) * 2 + hits["strand"].astype(np.int64)
    order = np.lexsort((hits["pos"], grp))
    hits, grp = hits[order], grp[order]
    first = np.r_[True, grp[1:] != grp[:-1]]
    start = np.maximum.accumulate(np.where(first, np.arange(grp.size), 0))
    return hits[np.arange(grp.size) - start < cap]


def scan_capped(genome: DeviceGenome, specs: Sequence[GuideSpec], cap_per_strand: int,
                tuning: Optional[dict] = None, strands: int = 3,
                max_records: int = DEFAULT_MAX_RECORDS) -> CappedScan:
    """Every site of every guide, except that a guide keeps only its
    `cap_per_strand` lowest-position sites on a strand (positions in the
    scanned strand array), with exact per-strand counts.  No launch holds more
    than `max_records` records: a guide set whose sites exceed it is split in
    halves, and a single guide whose sites exceed it is scanned per strand
    over position ranges, lowest first, until `cap_per_strand` sites are
    kept.  The reference bounds the same work by its per-strand ceiling
    (finder.py:568-638) and its decode budget of 4,000,000 records
    (finder.py:116, 157-165)."""
    cap_per_strand = int(cap_per_strand)
    max_records = max(int(max_records), 1024)
    specs = list(specs)
    n = len(specs)
    counts = np.zeros((n, 2), dtype=np.int64)
    parts: List[np.ndarray] = []
    stat = {"launches": 0, "ms": 0.0}

    def launch(plan, b, e, mask):
        plan.launch(b, e, mask)
        m, ms = plan.wait()
        stat["launches"] += 1
        stat["ms"] += ms
        return m

    def fetch_exact(plan, b, e, mask):
        """All records of [b, e), or None when they exceed max_records."""
        m = launch(plan, b, e, mask)
        if m > plan.hit_capacity:
            if m > max_records:
                return None
            plan.set_hit_capacity(m)
            m = launch(plan, b, e, mask)
        return plan.fetch(m)

    def group(idx: np.ndarray) -> bool:
        sub = [specs[i] for i in idx.tolist()]
        plan = Plan(genome, sub, tuning=tuning,
                    hit_capacity=min(max_records, default_capacity(len(sub))))
        try:
            h = fetch_exact(plan, 0, None, strands)
        finally:
            plan.close()
        if h is None:
            return False
        g = idx[h["guide"].astype(np.int64)]
        h["guide"] = g
        sizes = np.bincount(g * 2 + h["strand"].astype(np.int64), minlength=2 * n)
        counts[:] += sizes.reshape(n, 2)
        parts.append(_keep_lowest(h, cap_per_strand, sizes))
        return True

    def single(i: int) -> None:
        plan = Plan(genome, [specs[i]], tuning=tuning, hit_capacity=max_records)
        try:
            for s in (0, 1):
                if not strands & (1 << s):
                    continue
                total = launch(plan, 0, None, 1 << s)
                counts[i, s] = total
                kept: List[np.ndarray] = []
                if total <= max_records:
                    kept.append(plan.fetch(total))
                else:  # lowest positions first, ranges sized to the record limit
                    end = genome.total_bp
                    lo, got = 0, 0
                    step = max(64, end // 64)
                    while lo < end and got < cap_per_strand:
                        hi = min(end, lo + step)
                        m = launch(plan, lo, hi, 1 << s)
                        if m > max_records:  # (64 positions hold <= 64 sites)
                            step = max(64, (hi - lo) // 2)
                            continue
                        kept.append(plan.fetch(m))
                        got += m
                        lo = hi
                        if 4 * m < max_records:
                            step *= 2
                h = np.concatenate(kept) if kept else np.zeros(0, dtype=HIT_DTYPE)
                h["guide"] = i
                parts.append(_keep_lowest(h, cap_per_strand))
        finally:
            plan.close()

    pending = [np.arange(n, dtype=np.int64)] if n else []
    while pending:
        idx = pending.pop()
        if group(idx):
            continue
        if idx.size > 1:
            half = idx.size // 2
            pending += [idx[half:], idx[:half]]
        else:
            single(int(idx[0]))
    hits = np.concatenate(parts) if parts else np.zeros(0, dtype=HIT_DTYPE)
    return CappedScan(hits, counts, stat["launches"], stat["ms"])
