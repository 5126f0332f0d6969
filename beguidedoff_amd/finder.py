"""Extensive-mode off-target finder on MI355X.

Drop-in for the reference's ``beguidedoff.finder.beguidedoff``
(beguidedoff/finder.py:256-679): same constructor and ``search`` arguments,
same deduplication, on-target exclusion, coordinate conventions, escalation
and truncation policy, sort order and 12 output columns.

Two scan engines produce the same hits:

* ``engine="plain"`` runs the reference's control flow (strand loop, batches
  of <= MAX_GUIDES keys, per-guide result buffers, overflow re-scans) over the
  drop-in kernel behind :func:`beguidedoff_amd.cuda_search.search_genome_cuda`.
* ``engine="batched"`` (default) scans both strands for every key in one
  persistent launch of the batched kernel, which also returns the mismatch
  mask and CCTOP score of each site.

Per-hit decoding (coordinates, sequences, scores, editability, sorting) runs
over the hits of all keys at once with numpy and reproduces the reference's
per-hit Python loop (finder.py:193-244, 405-472, 545-566) value for value.
With the batched engine the site and PAM bases come from the bit planes in
the hit records; the reported rows are checked against the genome.
"""
from __future__ import annotations

import ctypes as C
import logging
import threading
from collections import defaultdict
from dataclasses import dataclass
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import pandas as pd

from . import _lib
from .base_editing import EDITABLE_BASE, edit_window_mask
from .encoding import encode_dna, encode_dna_rows, encode_guide_64bit
from .genome_converter import load_genome
from .io import OUTPUT_COLUMNS, load_guides_from_tsv
from .pam import encode_pam_mask
from .scoring import cctop_table
from .synth import nmask_le
from .types import GuideInput

_LOGGER = logging.getLogger("beguidedoff")

_MAX_GUIDES_PER_LAUNCH = 2048          # cuda_search.MAX_GUIDES
_DEFAULT_RESULTS_PER_GUIDE = 10240     # finder.py:103
_MAX_RESULTS_PER_GUIDE = 2_621_440     # finder.py:111 (read at call time)
_DECODE_DICT_BUDGET = 4_000_000        # finder.py:116
_PROFILES_TSV = Path(__file__).resolve().parent / "base_editor_profiles.tsv"
_W40 = (1 << 40) - 1


@dataclass(frozen=True)
class _ScanGuide:
    """Deduplicated scan key (beguidedoff/finder.py:77-85)."""

    guide_seq: str
    pam: str
    on_target_chrom_idx: Optional[int] = None
    on_target_local_pos: Optional[int] = None
    on_target_strand: Optional[str] = None


class RawHits:
    """Raw sites of one scan key: scanned strand (0 '+', 1 '-'), position in
    the scanned strand array, N-corrected mismatch count, and (batched engine)
    the fused-epilogue mismatch mask and score."""

    __slots__ = ("strand", "pos", "n_mm", "mm_mask", "score")

    def __init__(self, strand, pos, n_mm, mm_mask=None, score=None):
        self.strand = np.asarray(strand, dtype=np.int8)
        self.pos = np.asarray(pos, dtype=np.int64)
        self.n_mm = np.asarray(n_mm, dtype=np.int64)
        self.mm_mask = None if mm_mask is None else np.asarray(mm_mask, dtype=np.uint32)
        self.score = None if score is None else np.asarray(score, dtype=np.float64)

    def __len__(self) -> int:
        return int(self.pos.size)

    @staticmethod
    def empty() -> "RawHits":
        return RawHits(np.zeros(0), np.zeros(0), np.zeros(0))

    @staticmethod
    def concat(parts: Sequence["RawHits"]) -> "RawHits":
        parts = [p for p in parts if len(p)]
        if not parts:
            return RawHits.empty()
        fused = all(p.mm_mask is not None for p in parts)
        return RawHits(np.concatenate([p.strand for p in parts]),
                       np.concatenate([p.pos for p in parts]),
                       np.concatenate([p.n_mm for p in parts]),
                       np.concatenate([p.mm_mask for p in parts]) if fused else None,
                       np.concatenate([p.score for p in parts]) if fused else None)


@dataclass
class _AllHits:
    """Raw sites of all scan keys in flat arrays: key index, scanned strand,
    position in the scanned strand array, N-corrected mismatch count, and
    (batched engine) the fused-epilogue mismatch mask, score and the bit
    planes x0 / x1 of the 26 scanned bases from the site's start.  mm_out,
    when given, is the mask reported as mismatch positions (the fast mode's
    includes the N positions); by default the 2-bit mask is reported."""
    key: np.ndarray
    strand: np.ndarray
    pos: np.ndarray
    n_mm: np.ndarray
    mm_mask: Optional[np.ndarray] = None
    score: Optional[np.ndarray] = None
    x0: Optional[np.ndarray] = None
    x1: Optional[np.ndarray] = None
    mm_out: Optional[np.ndarray] = None

    @staticmethod
    def from_dict(scan_guides: Sequence["_ScanGuide"],
                  raw: Dict["_ScanGuide", RawHits]) -> "_AllHits":
        parts = [raw.get(sg, RawHits.empty()) for sg in scan_guides]
        sizes = np.array([len(p) for p in parts], dtype=np.int64)
        key = np.repeat(np.arange(len(parts), dtype=np.int64), sizes)
        full = [p for p in parts if len(p)]
        if not full:
            z = np.zeros(0, dtype=np.int64)
            return _AllHits(z, z.astype(np.int8), z, z)
        fused = all(p.mm_mask is not None for p in full)
        return _AllHits(key, np.concatenate([p.strand for p in full]),
                        np.concatenate([p.pos for p in full]),
                        np.concatenate([p.n_mm for p in full]),
                        np.concatenate([p.mm_mask for p in full]) if fused else None,
                        np.concatenate([p.score for p in full]) if fused else None)


def _parse_device(device) -> Tuple[str, int]:
    if device is None:
        return "cpu", 0
    if hasattr(device, "type"):
        return str(device.type), int(device.index or 0)
    s = str(device)
    kind, _, idx = s.partition(":")
    return kind, int(idx) if idx else 0


def _guide_planes_2bit(seq: str) -> int:
    return encode_guide_64bit(encode_dna(seq)) & _W40


class beguidedoff:
    """Extensive (exhaustive GPU) off-target search."""

    # Most site records one batched launch may hold (None: the library's
    # default, batched.DEFAULT_MAX_RECORDS).
    max_records: Optional[int] = None
    # Batched engine: build the table on the device from the records in HBM
    # (bgo_plan_assemble); False assembles every record on the host.
    device_assembly: bool = True

    def __init__(self, genome_path: Optional[str] = None, max_mismatches: int = 4,
                 scoring_method: str = "CCTOP", device: Optional[str] = None,
                 n_top_offtargets: int = 3, results_per_guide: int = 10240,
                 max_results_per_guide: Optional[int] = None, engine: str = "batched",
                 genome: Optional[dict] = None, tuning: Optional[dict] = None):
        self.genome_path = genome_path
        self.max_mismatches = max_mismatches
        if scoring_method.upper() != "CCTOP":
            raise ValueError("extensive mode only supports scoring_method='CCTOP'")
        self.scoring_method = "CCTOP"
        self.n_top_offtargets = n_top_offtargets
        self.results_per_guide = results_per_guide
        self.max_results_per_guide = max_results_per_guide
        if engine not in ("batched", "plain"):
            raise ValueError(f"engine must be 'batched' or 'plain', got {engine!r}")
        self.engine = engine
        self.tuning = dict(tuning or {})

        gpu_ok = True
        try:
            n_dev = _lib.device_count()
        except _lib.LibraryUnavailable:
            gpu_ok, n_dev = False, 0
        if device is None:
            device = "cuda:0" if n_dev > 0 else "cpu"
        kind, index = _parse_device(device)
        self.device = f"{kind}:{index}" if kind in ("cuda", "hip") else kind
        if kind not in ("cuda", "hip") or not gpu_ok or index >= n_dev:
            raise RuntimeError(
                f"Extensive mode requires an MI355X GPU (device '{self.device}', "
                f"{n_dev} HIP device(s), kernel library "
                f"{'available' if gpu_ok else 'unavailable'}). Use the fast CPU mode "
                "instead: `beguidedoff(mode='fast', index_path=...)`.")
        self.device_index = index
        _lib.init(index)

        if genome is None:
            if genome_path is None:
                raise ValueError("genome_path (or a preloaded genome dict) is required")
            genome = load_genome(genome_path)
        self.genome_data = genome
        self.total_bp = int(genome["total_bp"])
        self.chrom_names = [str(c) for c in genome["chrom_names"]]
        self.chrom_offsets = np.asarray(genome["chrom_offsets"], dtype=np.int64)
        self.chrom_lengths = np.asarray(genome["chrom_lengths"], dtype=np.int64)
        self._chrom_names_arr = np.array(self.chrom_names, dtype=object)
        self._prepare_n_masks()
        self._genome_handle = None   # batched engine: device genome, built lazily
        # the device-table path's plan, reused by the next search (its buffers
        # and stream stay allocated); the lock keeps one search on it at a time
        self._table_plan = None
        self._table_lock = threading.Lock()

    @classmethod
    def host_view(cls, genome: dict, max_mismatches: int = 4, n_top_offtargets: int = 3,
                  engine: str = "batched") -> "beguidedoff":
        """A finder without a device, for host-side decoding of hits produced
        elsewhere (e.g. replaying recorded raw hits); it cannot scan."""
        self = cls.__new__(cls)
        self.genome_path = None
        self.max_mismatches = max_mismatches
        self.scoring_method = "CCTOP"
        self.n_top_offtargets = n_top_offtargets
        self.results_per_guide = _DEFAULT_RESULTS_PER_GUIDE
        self.max_results_per_guide = None
        self.engine = engine
        self.tuning = {}
        self.device = "none"
        self.device_index = -1
        self.genome_data = genome
        self.total_bp = int(genome["total_bp"])
        self.chrom_names = [str(c) for c in genome["chrom_names"]]
        self.chrom_offsets = np.asarray(genome["chrom_offsets"], dtype=np.int64)
        self.chrom_lengths = np.asarray(genome["chrom_lengths"], dtype=np.int64)
        self._chrom_names_arr = np.array(self.chrom_names, dtype=object)
        self._prepare_n_masks()
        self._genome_handle = None
        self._table_plan = None
        self._table_lock = threading.Lock()
        return self

    def assemble(self, guides: List[GuideInput], raw_by_key: Dict[tuple, RawHits],
                 return_all: bool = False,
                 count_score_threshold: Optional[float] = None) -> pd.DataFrame:
        """Table of `guides` from raw hits given per (guide_seq, pam) key."""
        scan_guides, mapping = self._deduplicate_guides(guides)
        raw = {sg: raw_by_key.get((sg.guide_seq, sg.pam), RawHits.empty()) for sg in scan_guides}
        return assemble_results(self, guides, scan_guides, mapping, raw, count_score_threshold,
                                return_all)

    # ---------------------------------------------------------------- setup

    def _prepare_n_masks(self) -> None:
        """Little-endian N masks of both strand arrays, or None when the
        artefact has no N (beguidedoff/finder.py:310-345)."""
        g = self.genome_data
        if "n_mask_le_fwd" in g and "n_mask_le_rc" in g:
            fwd, rc = g["n_mask_le_fwd"], g["n_mask_le_rc"]
            has_n = fwd is not None and bool(np.any(fwd))
            self._n_mask_fwd = fwd if has_n else None
            self._n_mask_rc = rc if has_n else None
            return
        self._n_mask_fwd, self._n_mask_rc = nmask_le(g)

    @property
    def genome_info(self) -> dict:
        return {"build": "unknown", "total_bp": self.total_bp,
                "n_chromosomes": len(self.chrom_names), "chromosomes": self.chrom_names}

    def _strand_arrays(self):
        out = [(0, "+", self.genome_data["genome_forward"], self._n_mask_fwd)]
        if "genome_rc" in self.genome_data:
            out.append((1, "-", self.genome_data["genome_rc"], self._n_mask_rc))
        return out

    def _deduplicate_guides(self, guides: List[GuideInput]):
        """Scan keys and key -> inputs mapping (beguidedoff/finder.py:361-403)."""
        mapping: Dict[_ScanGuide, List[GuideInput]] = defaultdict(list)
        for g in guides:
            ot_idx = ot_local = ot_strand = None
            if g.on_target_chrom and g.on_target_pos is not None:
                ot_idx = next((i for i, n in enumerate(self.chrom_names)
                               if n == g.on_target_chrom), None)
                if ot_idx is not None:
                    if g.on_target_strand is None:
                        raise ValueError(
                            f"Guide {g.guide_id} has on-target chromosome and position but no "
                            "strand. Please provide on_target_strand ('+' or '-').")
                    window = 20 + len(g.pam)
                    ot_local = (g.on_target_pos - 1 if g.on_target_strand == "+"
                                else g.on_target_pos - window)
                    ot_strand = g.on_target_strand
            mapping[_ScanGuide(g.guide_seq, g.pam, ot_idx, ot_local, ot_strand)].append(g)
        return list(mapping.keys()), mapping

    # ---------------------------------------------------------------- search

    def search(self, input_file: Optional[str] = None, guides: Optional[List[GuideInput]] = None,
               return_all: bool = False, count_score_threshold: Optional[float] = None,
               output_file: Optional[str] = None, include_non_editable: bool = False,
               presets_path: Optional[str] = None, verbose: bool = True,
               logger: Optional[logging.Logger] = None) -> pd.DataFrame:
        """Search the genome for every guide (beguidedoff/finder.py:474-566).

        ``include_non_editable`` is accepted and ignored, like the reference:
        non-editable sites are kept exactly when ``return_all`` is set.
        """
        if (input_file is None) == (guides is None):
            raise ValueError("Exactly one of input_file or guides must be provided")
        if input_file is not None:
            guides = load_guides_from_tsv(
                input_file, presets_path=presets_path or str(_PROFILES_TSV))
        if not guides:
            raise ValueError("No guides provided")
        log = logger if logger is not None else _LOGGER
        for g in guides:
            if g.pam_position.lower() != "3prime":
                raise ValueError(f"Guide {g.guide_id} has pam_position '{g.pam_position}'. "
                                 "Only 3prime PAM is supported.")
        scan_guides, mapping = self._deduplicate_guides(guides)
        if verbose:
            log.info("Guides after deduplication: %d (from %d input rows)", len(scan_guides),
                     len(guides))
            log.info("Running GPU search on %s (%s engine) ...", self.device, self.engine)
        if self.engine == "batched" and self.device_assembly:
            df = self._search_device_table(guides, scan_guides, mapping, count_score_threshold,
                                           return_all)
            if df is not None:
                if output_file:
                    df.to_csv(output_file, sep="\t", index=False)
                    if verbose:
                        log.info("Wrote results to %s", output_file)
                return df
        if self.engine == "plain":
            raw, overflow, _counts = self._search_gpu(scan_guides, log, verbose)
        else:
            raw, overflow, _counts = self._search_batched(scan_guides, log, verbose)
        truncated = [sg for sg, o in overflow.items() if o]
        if truncated:
            log.warning("%d guide(s) returned truncated off-target lists (the re-scan "
                        "escalation ceiling was reached); their off-target counts are lower "
                        "bounds. See the per-guide warnings above.", len(truncated))
        df = assemble_results(self, guides, scan_guides, mapping, raw, count_score_threshold,
                              return_all)
        if output_file:
            df.to_csv(output_file, sep="\t", index=False)
            if verbose:
                log.info("Wrote results to %s", output_file)
        return df

    # ----------------------------------------------------- plain engine

    def _search_strand(self, strand_idx: int, words, n_mask, scan_guides: List[_ScanGuide],
                       results_per_guide: int):
        """One strand, keys in launches of <= MAX_GUIDES (finder.py:119-249)."""
        from .cuda_search import search_genome_cuda

        if results_per_guide <= _DEFAULT_RESULTS_PER_GUIDE:
            batch_limit = _MAX_GUIDES_PER_LAUNCH
        else:
            batch_limit = max(1, _DECODE_DICT_BUDGET // results_per_guide)
        raw, ovf, counts = {}, {}, {}
        for b0 in range(0, len(scan_guides), batch_limit):
            batch = scan_guides[b0:b0 + batch_limit]
            n = len(batch)
            vals = np.zeros(n, dtype=np.uint64)
            masks = np.zeros((n, 6), dtype=np.uint8)
            lens = np.zeros(n, dtype=np.uint8)
            for i, sg in enumerate(batch):
                vals[i] = encode_guide_64bit(encode_dna(sg.guide_seq)) & ((1 << 64) - 1)
                m = encode_pam_mask(sg.pam)
                masks[i, :] = m
                lens[i] = len(sg.pam)
            res, o, c = search_genome_cuda(
                words, self.total_bp, vals, masks, lens, np.ones(n, dtype=np.uint8),
                self.max_mismatches, return_overflow=True, results_per_guide=results_per_guide,
                n_mask_packed=n_mask, n_encoded=0 if strand_idx == 0 else 3)
            for gi, sg in enumerate(batch):
                hits = res[gi]
                pos = np.fromiter((h[0] for h in hits), dtype=np.int64, count=len(hits))
                mm = np.fromiter((h[1] for h in hits), dtype=np.int64, count=len(hits))
                raw[sg] = RawHits(np.full(pos.size, strand_idx), pos, mm)
                ovf[sg] = bool(o.get(gi, False))
                counts[sg] = int(c.get(gi, 0))
        return raw, ovf, counts

    def _scan_strands(self, scan_guides, results_per_guide):
        raw = {sg: [] for sg in scan_guides}
        ovf = {sg: False for sg in scan_guides}
        counts = {sg: 0 for sg in scan_guides}
        for s, _label, words, n_mask in self._strand_arrays():
            r, o, c = self._search_strand(s, words, n_mask, scan_guides, results_per_guide)
            for sg in scan_guides:
                raw[sg].append(r[sg])
                ovf[sg] = ovf[sg] or o[sg]
                counts[sg] += c[sg]
        return {sg: RawHits.concat(v) for sg, v in raw.items()}, ovf, counts

    def _search_gpu(self, scan_guides, log, verbose):
        """Overflow escalation loop of beguidedoff/finder.py:568-638."""
        raw = {sg: RawHits.empty() for sg in scan_guides}
        overflow = {sg: False for sg in scan_guides}
        counts = {sg: 0 for sg in scan_guides}
        pending = list(scan_guides)
        rpg = self.results_per_guide
        while pending:
            r, o, c = self._scan_strands(pending, rpg)
            for sg in pending:
                raw[sg], overflow[sg], counts[sg] = r[sg], o[sg], c[sg]
            pending = [sg for sg in pending if o[sg]]
            if not pending:
                break
            cap = (self.max_results_per_guide if self.max_results_per_guide is not None
                   else _MAX_RESULTS_PER_GUIDE)
            if rpg >= cap:
                for sg in sorted(pending, key=lambda s: s.guide_seq):
                    _warn_truncated(log, sg, counts[sg], rpg, len(raw[sg]))
                break
            rpg = min(rpg * 16, cap)
            if log is not None:
                names = ", ".join(sorted({sg.guide_seq for sg in pending})[:5])
                log.warning("Hit buffer overflow for %d guide(s) at results_per_guide=%s "
                            "(e.g. %s); re-scanning them with results_per_guide=%d.",
                            len(pending), self.results_per_guide, names, rpg)
        return raw, overflow, counts

    # --------------------------------------------------- batched engine

    def _device_genome(self):
        if self._genome_handle is None:
            from .batched import DeviceGenome

            self._genome_handle = DeviceGenome(self.genome_data, self._n_mask_fwd,
                                               self._n_mask_rc)
        return self._genome_handle

    def _search_batched(self, scan_guides, log, verbose):
        """All keys, both strands, one persistent launch per PAM-family group.

        Results are exact; a key whose sites on one strand exceed the
        escalation ceiling keeps the lowest-position `ceiling` of them and is
        reported as truncated (the reference keeps the first `ceiling` the
        kernel stored, finder.py:609-625).  No launch holds more than
        `max_records` site records (batched.scan_capped), so a heavy
        multi-mapper costs bounded device and host memory, as the reference's
        per-strand ceiling bounds it."""
        from .batched import GuideSpec, scan_capped

        specs = [GuideSpec(sg.guide_seq, sg.pam, self.max_mismatches) for sg in scan_guides]
        cap = (self.max_results_per_guide if self.max_results_per_guide is not None
               else _MAX_RESULTS_PER_GUIDE)
        kw = {} if self.max_records is None else {"max_records": self.max_records}
        res = scan_capped(self._device_genome(), specs, cap, tuning=self.tuning, **kw)
        hits = res.hits
        key = hits["guide"].astype(np.int64)
        n_keys = len(scan_guides)
        counts_k = res.counts.sum(axis=1)
        over_k = (res.counts > cap).any(axis=1)
        if over_k.any():
            kept_k = np.bincount(key, minlength=n_keys)
            for k in np.flatnonzero(over_k):
                _warn_truncated(log, scan_guides[k], int(counts_k[k]), cap, int(kept_k[k]))
        flat = _AllHits(key, hits["strand"].astype(np.int8), hits["pos"].astype(np.int64),
                        hits["n_mm"].astype(np.int64), hits["mm_mask"].astype(np.uint32),
                        hits["score"].astype(np.float64), hits["x0"].astype(np.uint32),
                        hits["x1"].astype(np.uint32))
        overflow = {sg: bool(over_k[k]) for k, sg in enumerate(scan_guides)}
        counts = {sg: int(counts_k[k]) for k, sg in enumerate(scan_guides)}
        return flat, overflow, counts

    def _search_device_table(self, guides, scan_guides, mapping, count_score_threshold,
                             return_all) -> Optional[pd.DataFrame]:
        """One launch over every key and the table assembled where the records
        lie (bgo_plan_assemble): coordinates, on-target exclusion,
        editability per input, below-threshold counts, per-id order and top
        N on the device; only the rows come back.  None when the launch's
        sites overflow its buffer or a key exceeds the escalation ceiling on a
        strand: the host path then runs (scan_capped + assemble_results)."""
        from .batched import DEFAULT_MAX_RECORDS, GuideSpec, default_capacity

        n_keys = len(scan_guides)
        if n_keys == 0:
            return None
        specs = [GuideSpec(sg.guide_seq, sg.pam, self.max_mismatches) for sg in scan_guides]
        limit = self.max_records if self.max_records is not None else DEFAULT_MAX_RECORDS
        cap = min(limit, default_capacity(n_keys))
        with self._table_lock:
            try:
                return self._device_table_on_plan(guides, scan_guides, mapping, specs, cap,
                                                  count_score_threshold, return_all)
            except BaseException:
                self._close_table_plan()  # its state after a failure is not reused
                raise

    def _close_table_plan(self) -> None:
        if self._table_plan is not None:
            self._table_plan.close()
            self._table_plan = None

    def _device_table_on_plan(self, guides, scan_guides, mapping, specs, cap,
                              count_score_threshold, return_all) -> Optional[pd.DataFrame]:
        from .batched import Plan

        n_keys = len(scan_guides)
        plan = self._table_plan
        if plan is None:
            plan = self._table_plan = Plan(self._device_genome(), specs, tuning=self.tuning,
                                           hit_capacity=cap)
        else:
            plan.set_guides(specs)
            if plan.hit_capacity != cap:
                plan.set_hit_capacity(cap)
        plan.launch()
        # the host tables of the assembly are built while the scan runs
        keys, inputs, gid, n_ids, input_index = self._assembly_tables(guides, scan_guides, mapping)
        ceiling = (self.max_results_per_guide if self.max_results_per_guide is not None
                   else _MAX_RESULTS_PER_GUIDE)
        names = self.chrom_names
        rank_of = {nm: i for i, nm in enumerate(sorted(set(names)))}
        name_rank = np.array([rank_of[nm] for nm in names], dtype=np.uint32)
        top = max(int(self.n_top_offtargets), 0)  # <= 0: no rows, as the host path
        n, _ms = plan.wait()
        if n > cap:
            return None
        rows_cap = max(1, n_ids * top) if not return_all else max(1, n * len(inputs))
        below = np.zeros(max(n_ids, 1), dtype=np.int64)
        while True:
            rows = np.zeros(rows_cap, dtype=_lib.ASM_ROW_DTYPE)
            n_rows, mx = C.c_int64(0), C.c_int64(0)
            _lib.check(_lib.load().bgo_plan_assemble(
                plan.handle, _lib.ptr(keys), n_keys, _lib.ptr(inputs), len(inputs), n_ids,
                _lib.ptr(self.chrom_offsets), _lib.ptr(self.chrom_lengths),
                _lib.ptr(name_rank), len(names), top, int(bool(return_all)),
                float(count_score_threshold) if count_score_threshold is not None else 0.0,
                int(count_score_threshold is not None), int(ceiling), _lib.ptr(rows), rows_cap,
                C.byref(n_rows), _lib.ptr(below), C.byref(mx)))
            if mx.value > ceiling:
                return None
            if n_rows.value <= rows_cap:
                rows = rows[: n_rows.value]
                break
            rows_cap = int(n_rows.value)
        return self._rows_table(guides, scan_guides, rows, gid, below[:n_ids], input_index)

    def _assembly_tables(self, guides, scan_guides, mapping):
        """bgo_asm_key per scan key and bgo_asm_input per (key, input) pair,
        key-major with each key's inputs in input order; guide ids numbered
        by first appearance."""
        input_index = {id(g): i for i, g in enumerate(guides)}
        id_of: Dict[str, int] = {}
        gid = np.array([id_of.setdefault(g.guide_id, len(id_of)) for g in guides], dtype=np.int64)
        keys = np.zeros(len(scan_guides), dtype=_lib.ASM_KEY_DTYPE)
        gp0, gp1 = _guide_planes([sg.guide_seq for sg in scan_guides])
        keys["g0"], keys["g1"] = gp0, gp1
        keys["pam_len"] = [len(sg.pam) for sg in scan_guides]
        keys["ot_contig"] = [sg.on_target_chrom_idx if sg.on_target_chrom_idx is not None else -1
                             for sg in scan_guides]
        keys["ot_local"] = [sg.on_target_local_pos if sg.on_target_local_pos is not None else 0
                            for sg in scan_guides]
        keys["ot_strand"] = [0 if sg.on_target_strand == "+" else
                             (1 if sg.on_target_strand == "-" else -9) for sg in scan_guides]
        counts = np.array([len(mapping[sg]) for sg in scan_guides], dtype=np.int64)
        keys["n_inputs"] = counts
        keys["first_input"] = np.cumsum(counts) - counts
        inp_of_pair = np.fromiter((input_index[id(g)] for sg in scan_guides for g in mapping[sg]),
                                  dtype=np.int64, count=int(counts.sum()))
        # per input guide: its window mask and editable base (0 = A, 1 = C)
        win_of: Dict[tuple, int] = {}
        mask_in = np.zeros(len(guides), dtype=np.uint32)
        base_in = np.zeros(len(guides), dtype=np.uint8)
        for i, g in enumerate(guides):
            base = EDITABLE_BASE.get(g.editor_type)
            if base is None:
                raise ValueError(f"Unknown editor_type: {g.editor_type}")
            w = (g.edit_window_start, g.edit_window_end)
            m = win_of.get(w)
            if m is None:
                m = win_of[w] = edit_window_mask(*w)
            mask_in[i] = m
            base_in[i] = 0 if base == "A" else 1
        inputs = np.zeros(inp_of_pair.size, dtype=_lib.ASM_INPUT_DTYPE)
        inputs["id"] = gid[inp_of_pair]
        inputs["window"] = mask_in[inp_of_pair]
        inputs["target"] = base_in[inp_of_pair]
        inputs["order"] = np.arange(inp_of_pair.size)
        return keys, inputs, gid, len(id_of), inp_of_pair

    def _rows_table(self, guides, scan_guides, rows, gid, below, inp_of_pair) -> pd.DataFrame:
        """The 12-column table from device rows ordered by (id, rank): each
        input guide contributes its id's rows, in input order
        (finder.py:545-566)."""
        if rows.size == 0:
            return pd.DataFrame()
        h = rows["hit"]
        rid = gid[inp_of_pair[rows["input"].astype(np.int64)]]
        n_ids = int(gid.max()) + 1 if gid.size else 0
        per_id = np.bincount(rid, minlength=n_ids)
        starts = np.cumsum(per_id) - per_id
        # each input guide's block of its id's rows, in input order
        n_in = per_id[gid]
        first = np.cumsum(n_in) - n_in
        sel = np.repeat(starts[gid] - first, n_in) + np.arange(int(n_in.sum()))
        if sel.size == 0:
            return pd.DataFrame()
        r = rows[sel]
        h = r["hit"]
        key = h["guide"].astype(np.int64)
        strand = h["strand"].astype(np.int64)
        pos = h["pos"].astype(np.int64)
        x0, x1 = h["x0"].astype(np.uint32), h["x1"].astype(np.uint32)
        plen_k = np.array([len(sg.pam) for sg in scan_guides], dtype=np.int64)
        shifts = np.arange(26, dtype=np.uint32)
        codes = (((x0[:, None] >> shifts) & 1) | (((x1[:, None] >> shifts) & 1) << 1)).astype(
            np.uint8)
        for s, _label, words, _nm in self._strand_arrays():  # the rows' bases against the genome
            m = strand == s
            if m.any():
                ref = _codes_at(words, pos[m], 26)
                width = 20 + plen_k[key[m]]
                if np.any((ref != codes[m]) & (np.arange(26)[None, :] < width[:, None])):
                    raise RuntimeError("site bases in the hit records disagree with the genome")
        gseq_k = np.array([sg.guide_seq for sg in scan_guides], dtype=object)
        inp = inp_of_pair[r["input"].astype(np.int64)]
        editor_in = np.array([g.base_editor or "" for g in guides], dtype=object)
        pam_str = np.empty(sel.size, dtype=object)
        for pl in np.unique(plen_k[key]):
            m = plen_k[key] == pl
            pam_str[m] = _strings(codes[m, 20:20 + pl])
        cols = {
            "guide_rna_sequence": gseq_k[key],
            "base_editor": editor_in[inp],
            "ot_rank": r["rank"].astype(np.int64) + 1,
            "ot_chromosome": self._chrom_names_arr[r["contig"].astype(np.int64)],
            "ot_position": r["local"].astype(np.int64) + 1,
            "ot_strand": np.where(strand == 0, "+", "-").astype(object),
            "ot_sequence": _strings(codes[:, :20]),
            "ot_pam": pam_str,
            "n_mismatches": h["n_mm"].astype(np.int64),
            "mismatch_positions": np.array([_mm_string(m) for m in h["mm_mask"].tolist()],
                                           dtype=object),
            "score": np.array([_round6(x) for x in h["score"].tolist()], dtype=np.float64),
            "n_offtargets_below_threshold": below[gid[inp]],
        }
        return pd.DataFrame(cols, columns=OUTPUT_COLUMNS)

    def close(self) -> None:
        with self._table_lock:
            self._close_table_plan()
        if self._genome_handle is not None:
            self._genome_handle.close()
            self._genome_handle = None


def _warn_truncated(log, sg, count, rpg, kept):
    log.warning("Truncated off-target enumeration for guide %s (PAM=%s): %d raw genome-wide "
                "hits exceed the per-strand buffer of %d; keeping the first %d (genome order). "
                "The off-target count for this guide is a lower bound.",
                sg.guide_seq, sg.pam, count, rpg, kept)


# ------------------------------------------------------------ epilogue


def _codes_at(words: np.ndarray, pos: np.ndarray, width: int) -> np.ndarray:
    """2-bit codes of bases [pos, pos + width) for every pos: (n, width) uint8.
    Bases past the end of the array read as its last word's (callers ignore
    them)."""
    if pos.size == 0:
        return np.zeros((0, width), dtype=np.uint8)
    words = np.asarray(words)
    idx = pos[:, None] + np.arange(width, dtype=np.int64)[None, :]
    w = words[np.minimum(idx >> 4, words.size - 1)].astype(np.uint32)
    return ((w >> ((idx & 15) * 2).astype(np.uint32)) & 3).astype(np.uint8)


_ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def _strings(codes: np.ndarray) -> np.ndarray:
    """(n, w) codes -> object array of strings."""
    n, w = codes.shape
    if n == 0:
        return np.zeros(0, dtype=object)
    b = _ACGT[codes].tobytes().decode("ascii")
    out = np.empty(n, dtype=object)
    out[:] = [b[i:i + w] for i in range(0, n * w, w)]
    return out


_MM_STR_CACHE: Dict[int, str] = {}


def _mm_string(mask: int) -> str:
    s = _MM_STR_CACHE.get(mask)
    if s is None:
        s = ",".join(str(i + 1) for i in range(20) if (mask >> i) & 1)
        _MM_STR_CACHE[mask] = s
    return s


def _guide_planes(seqs: List[str]) -> Tuple[np.ndarray, np.ndarray]:
    """Bit planes (b0, b1) of the first 20 bases of each guide, base i at bit i."""
    p0 = np.zeros(len(seqs), dtype=np.uint32)
    p1 = np.zeros(len(seqs), dtype=np.uint32)
    lens = np.array([len(x) for x in seqs], dtype=np.int64)
    for L in np.unique(lens).tolist():
        k = min(L, 20)
        if k == 0:
            continue
        idx = np.flatnonzero(lens == L)
        codes = encode_dna_rows([seqs[i] for i in idx], L)[:, :k].astype(np.uint32)
        sh = np.arange(k, dtype=np.uint32)
        p0[idx] = ((codes & 1) << sh).sum(axis=1, dtype=np.uint32)
        p1[idx] = ((codes >> 1) << sh).sum(axis=1, dtype=np.uint32)
    return p0, p1


def _top_candidates(pid: np.ndarray, score: np.ndarray, n_ids: int, top: int) -> np.ndarray:
    """Indices of the hits whose score is at most their id's top-th smallest
    (all of an id's hits when it has fewer): every hit that can be among an
    id's first `top` rows in (score, ...) order, ties at the cut included."""
    if top <= 0 or pid.size == 0:
        return np.zeros(0, dtype=np.int64)
    o = np.argsort(score)  # ties in any order: only each id's cut value is used
    p = pid[o]
    o = o[np.argsort(p.astype(np.uint16) if n_ids <= 0xFFFF else p, kind="stable")]
    per = np.bincount(pid, minlength=n_ids)
    starts = np.cumsum(per) - per
    has = per > 0
    thr = np.full(n_ids, -np.inf)
    thr[has] = score[o[starts[has] + np.minimum(per[has], top) - 1]]
    return np.flatnonzero(score <= thr[pid])


def cctop_scores(mm_mask: np.ndarray) -> np.ndarray:
    """sum(1.2 ** p) over set positions, accumulated in ascending p: the same
    IEEE additions as calculate_cctop_score (x + 0.0 leaves x unchanged)."""
    tab = cctop_table()
    s = np.zeros(mm_mask.size, dtype=np.float64)
    for p in range(1, 21):
        s = s + np.where((mm_mask >> np.uint32(p - 1)) & 1, tab[p], 0.0)
    return s


def _position_to_chrom(offsets: np.ndarray, pos: np.ndarray):
    """Vector form of finder.py:36-40; index -1 wraps like Python indexing."""
    idx = np.searchsorted(offsets, pos, side="right") - 1
    return idx, pos - offsets[idx]


_W20 = (1 << 20) - 1


def _bits26(bits2d: np.ndarray) -> np.ndarray:
    """(n, 26) 0/1 values -> uint32 with bit i = column i."""
    return (bits2d.astype(np.uint32) << np.arange(26, dtype=np.uint32)[None, :]).sum(
        axis=1, dtype=np.uint64).astype(np.uint32)


def assemble_results(finder: "beguidedoff", guides: List[GuideInput],
                     scan_guides: List[_ScanGuide], mapping, raw,
                     count_score_threshold: Optional[float], return_all: bool) -> pd.DataFrame:
    """Coordinates, on-target exclusion, sequences, scores, editability,
    per-id sort, top-N and table for all keys at once
    (beguidedoff/finder.py:193-244, 405-472, 545-566).

    `raw` is an :class:`_AllHits` or a dict of per-key :class:`RawHits`.
    Hits are collected per guide_id (inputs sharing an id are merged, as the
    reference's id-keyed dicts merge them) and ordered by (score, chromosome
    name, position, strand); remaining ties keep the reference's insertion
    order (scan key, then input order).  Each input guide contributes its
    id's rows, in input order.
    """
    h = raw if isinstance(raw, _AllHits) else _AllHits.from_dict(scan_guides, raw)
    offs, lens = finder.chrom_offsets, finder.chrom_lengths
    names = finder._chrom_names_arr
    rank_of = {n: i for i, n in enumerate(sorted(set(finder.chrom_names)))}
    chrom_rank = np.array([rank_of[n] for n in finder.chrom_names], dtype=np.int64)
    plen_k = np.array([len(sg.pam) for sg in scan_guides], dtype=np.int64)

    # coordinates: '-' hits map back to the forward leftmost and report
    # positions (finder.py:193-215)
    key, strand, pos = h.key, h.strand, h.pos
    window = 20 + plen_k[key]
    leftmost, report = pos.copy(), pos.copy()
    minus = strand == 1
    if minus.any():
        k, local_rc = _position_to_chrom(offs, pos[minus])
        lm = offs[k] + (lens[k] - local_rc - window[minus])
        leftmost[minus] = lm
        report[minus] = lm + window[minus] - 1
    # on-target exclusion by (chromosome, local leftmost, strand) (finder.py:217-219)
    has_ot = np.array([sg.on_target_chrom_idx is not None for sg in scan_guides], dtype=bool)
    keep = np.ones(pos.size, dtype=bool)
    if has_ot.size and has_ot.any():
        ot_chrom = np.array([sg.on_target_chrom_idx if sg.on_target_chrom_idx is not None else -1
                             for sg in scan_guides], dtype=np.int64)
        ot_local = np.array([sg.on_target_local_pos if sg.on_target_local_pos is not None else -1
                             for sg in scan_guides], dtype=np.int64)
        ot_code = np.array([0 if sg.on_target_strand == "+" else
                            (1 if sg.on_target_strand == "-" else -9) for sg in scan_guides],
                           dtype=np.int64)
        k, local = _position_to_chrom(offs, leftmost)
        keep = ~(has_ot[key] & (k == ot_chrom[key]) & (local == ot_local[key])
                 & (strand == ot_code[key]))
    if not keep.all():
        key, strand, pos, report = key[keep], strand[keep], pos[keep], report[keep]
    n_mm = h.n_mm[keep] if not keep.all() else h.n_mm
    n = int(pos.size)

    # 20-nt site and genomic PAM from the scanned strand (finder.py:221-244),
    # as bit planes of its 26 bases: from the batched kernel's records, or read
    # from the strand arrays
    if h.x0 is not None:
        x0 = h.x0[keep] if not keep.all() else h.x0
        x1 = h.x1[keep] if not keep.all() else h.x1
    else:
        codes = np.zeros((n, 26), dtype=np.uint8)
        for s, _label, words, _nm in finder._strand_arrays():
            sel = strand == s
            if sel.any():
                codes[sel] = _codes_at(words, pos[sel], 26)
        x0, x1 = _bits26(codes & 1), _bits26(codes >> 1)
    gp0, gp1 = _guide_planes([sg.guide_seq for sg in scan_guides])
    mm_mask = ((x0 ^ gp0[key]) | (x1 ^ gp1[key])) & np.uint32(_W20)
    if h.score is not None:
        score = h.score[keep] if not keep.all() else h.score
        fused = h.mm_mask[keep] if not keep.all() else h.mm_mask
        if not np.array_equal(fused, mm_mask):
            raise RuntimeError("fused epilogue mismatch mask disagrees with the site's bases")
    else:
        score = cctop_scores(mm_mask)
    if h.mm_out is not None:
        mm_report = h.mm_out[keep] if not keep.all() else h.mm_out
    else:
        mm_report = mm_mask
    is_a = ~(x0 | x1) & np.uint32(_W20)   # code 0
    is_c = x0 & ~x1 & np.uint32(_W20)     # code 1

    # (hit, input guide) pairs: every input mapped to the hit's key (finder.py:424-439)
    input_index = {id(g): i for i, g in enumerate(guides)}
    n_in = np.array([len(mapping[sg]) for sg in scan_guides], dtype=np.int64)
    in_flat = np.array([input_index[id(g)] for sg in scan_guides for g in mapping[sg]],
                       dtype=np.int64)
    in_start = np.cumsum(n_in) - n_in
    rep_n = n_in[key]
    hid = np.repeat(np.arange(n, dtype=np.int64), rep_n)
    j = np.arange(hid.size, dtype=np.int64) - np.repeat(np.cumsum(rep_n) - rep_n, rep_n)
    inp = in_flat[in_start[key[hid]] + j]
    # per-input editor target and window, and guide_id index (first appearance)
    used = np.zeros(len(guides), dtype=bool)
    used[inp] = True
    target = np.zeros(len(guides), dtype=np.uint8)
    wmask = np.zeros(len(guides), dtype=np.uint32)
    for i in np.flatnonzero(used):
        g = guides[i]
        base = EDITABLE_BASE.get(g.editor_type)
        if base is None:
            raise ValueError(f"Unknown editor_type: {g.editor_type}")
        target[i] = {"A": 0, "C": 1}[base]
        wmask[i] = edit_window_mask(g.edit_window_start, g.edit_window_end)
    id_of: Dict[str, int] = {}
    gid = np.array([id_of.setdefault(g.guide_id, len(id_of)) for g in guides], dtype=np.int64)
    n_ids = len(id_of)
    editable = (np.where(target[inp] == 0, is_a[hid], is_c[hid]) & wmask[inp]) != 0
    pid = gid[inp]
    below = np.zeros(n_ids, dtype=np.int64)
    if count_score_threshold is not None:
        below = np.bincount(pid[editable & (score[hid] <= count_score_threshold)],
                            minlength=n_ids).astype(np.int64)
    take = np.flatnonzero(editable | bool(return_all))
    hid, inp, pid = hid[take], inp[take], pid[take]

    # per-id order (score, chromosome name, position, strand, key, input)
    chrom, local = _position_to_chrom(offs, report[hid])
    # (chromosome rank, position) and (strand, key, input) packed into int64 keys
    place = (chrom_rank[chrom] << 33) + local
    tie = (strand[hid].astype(np.int64) << 62) + (key[hid] << 31) + inp
    sc = score[hid]
    if return_all:
        order = np.lexsort((tie, place, sc, pid))
    else:
        # only each id's top-N scores (ties included) can be reported
        order = _top_candidates(pid, sc, n_ids, finder.n_top_offtargets)
        order = order[np.lexsort((tie[order], place[order], sc[order], pid[order]))]
    pid_s = pid[order]
    per_id = np.bincount(pid_s, minlength=n_ids)
    starts = np.cumsum(per_id) - per_id
    rank = np.arange(order.size, dtype=np.int64) - starts[pid_s]
    if not return_all:
        top = rank < finder.n_top_offtargets
        order, pid_s, rank = order[top], pid_s[top], rank[top]
        per_id = np.bincount(pid_s, minlength=n_ids)
        starts = np.cumsum(per_id) - per_id
    # rows: each input guide contributes its id's block, in input order
    rows = np.concatenate([np.arange(starts[gid[i]], starts[gid[i]] + per_id[gid[i]])
                           for i in range(len(guides))]) if len(guides) else np.zeros(0, np.int64)
    if rows.size == 0:
        return pd.DataFrame()
    r = order[rows]
    rh, rk, rc = hid[r], key[hid[r]], chrom[r]
    shifts = np.arange(26, dtype=np.uint32)
    codes = (((x0[rh][:, None] >> shifts) & 1) | (((x1[rh][:, None] >> shifts) & 1) << 1)).astype(
        np.uint8)
    if h.x0 is not None:  # the reported sites' bases against the strand arrays
        for s, _label, words, _nm in finder._strand_arrays():
            sel = strand[rh] == s
            if sel.any():
                ref = _codes_at(words, pos[rh[sel]], 26)
                width = 20 + plen_k[rk[sel]]
                if np.any((ref != codes[sel]) & (np.arange(26)[None, :] < width[:, None])):
                    raise RuntimeError("site bases in the hit records disagree with the genome")
    gseq_k = np.array([sg.guide_seq for sg in scan_guides], dtype=object)
    editor_in = np.array([g.base_editor or "" for g in guides], dtype=object)
    pam_str = np.empty(rows.size, dtype=object)
    for pl in np.unique(plen_k[rk]):
        m = plen_k[rk] == pl
        pam_str[m] = _strings(codes[m, 20:20 + pl])
    cols = {
        "guide_rna_sequence": gseq_k[rk],
        "base_editor": editor_in[inp[r]],
        "ot_rank": rank[rows] + 1,
        "ot_chromosome": names[rc],
        "ot_position": (local[r] + 1).astype(np.int64),
        "ot_strand": np.where(strand[rh] == 0, "+", "-").astype(object),
        "ot_sequence": _strings(codes[:, :20]),
        "ot_pam": pam_str,
        "n_mismatches": n_mm[rh].astype(np.int64),
        "mismatch_positions": np.array([_mm_string(int(m)) for m in mm_report[rh]],
                                       dtype=object),
        "score": np.array([_round6(float(x)) for x in score[rh]], dtype=np.float64),
        "n_offtargets_below_threshold": below[pid_s[rows]],
    }
    return pd.DataFrame(cols, columns=OUTPUT_COLUMNS)


_ROUND_CACHE: Dict[float, float] = {}


def _round6(x: float) -> float:
    """Python's round(x, 6) (the reference's rounding), memoised per score."""
    r = _ROUND_CACHE.get(x)
    if r is None:
        r = round(x, 6)
        _ROUND_CACHE[x] = r
    return r
