"""Fast mode on MI355X: the reference's anchored CCTOP search
(beguidedoff/fast/finder.py:90-441, regions.py:80-123) computed by one
exhaustive batched scan of the resident genome instead of an FM-index.

A fast-mode candidate of a guide is a window inside one contig whose PAM
matches, whose "cold" block (positions p with 1.2**p >= cutoff) equals the
anchor on the 2-bit genome, and whose CCTOP score over its 2-bit mismatches
plus the N positions that count as mismatches is <= cutoff
(fast/finder.py:163-277).  Such a window has at most `budget` mismatches,
where budget is the largest m with 1.2 + ... + 1.2**m <= cutoff (3 at cutoff
5), so the batched kernel at that per-guide budget reports every candidate.
The host then applies the anchor, N and score rules exactly:

* '+' anchor: guide[c0:20] on the site's positions c0..19; '-' anchor:
  rc(guide)[:20 - c0], i.e. guide[len - (20 - c0):] on the reverse-complement
  site's positions c0..19 (the same block for 20-nt guides, shifted by one
  for 21-nt guides, as in fast/finder.py:186-189);
* N: an N at a site position whose guide base is the N code (A on '+', T on
  '-') is one more mismatch, in the positions and the score;
* the window, anchor included, lies inside one contig; windows up to the
  genome's end are scanned (the extensive mode's 26-position tail guard is
  lowered to 20 + the shortest PAM).

Two engines find the candidates, with identical results:
* "index" (default): an anchor index of the forward genome built in HBM
  (``bgo_fast_index_*``: every position keyed by its 12 bases, grouped by key
  with a counting sort on the GPU); a query reads only the anchors'
  occurrences and checks their windows on the GPU;
* "scan": the batched kernel at the cutoff's mismatch budget, filtered on the
  host.

The table is the extensive finder's (same 12 columns, sort and top-N).

Anchor occurrence cap (fast/finder.py:149-159, 368-375): like the reference,
an anchor with more than `max_hits_per_anchor` occurrences keeps only the
first ones of its suffix-array interval in the reference's FM-index text, and
the guide is reported with the reference's warning.  The index counts every
anchor's occurrences; for an anchor over the cap, its occurrences are listed
from the index and the cap of them whose suffixes come first are selected on
the host (bgo_sa_text_select), and only candidates anchored at those are
kept.  `max_hits_per_anchor=0` keeps every occurrence.
"""
from __future__ import annotations

import ctypes as C
import logging
import time
from typing import List, Optional, Tuple

import numpy as np
import pandas as pd

from .finder import (_AllHits, _W20, _LOGGER, _PROFILES_TSV, assemble_results, beguidedoff,
                     cctop_scores)
from . import _lib
from . import genome_converter
from .encoding import encode_dna, encode_dna_rows, reverse_complement_rows
from .io import load_guides_from_tsv
from .pam import encode_pam_mask
from .scoring import cctop_table
from .types import GuideInput

_GUIDE_LENGTH = 20
_META_HEADER = "#beguidedoff-fmindex\t1"


def _n_runs(seq: str) -> np.ndarray:
    """Half-open [start, end) runs of 'N' (after upper-casing) as an (n, 2) array."""
    raw = np.frombuffer(seq.upper().encode("latin-1", "replace"), dtype=np.uint8)
    isn = np.concatenate(([0], (raw == ord("N")).astype(np.int8), [0]))
    d = np.diff(isn)
    return np.stack([np.flatnonzero(d == 1), np.flatnonzero(d == -1)], axis=1)


def build_index(fasta_path: str, output_path: str, verbose: bool = False,
                engine: str = "numpy", device: int = 0) -> None:
    """The fast mode's index files (fast/fmindex.py:63-128), written under the
    reference's names:

    * ``output_path``: the encoded genome artefact (the genome_converter
      arrays, npz content at exactly this path), which the fast finder loads;
      the anchor index itself is built in HBM from it at start-up;
    * ``output_path + ".meta"``: the chromosome table in the reference's
      format (csrc/fmindex_binding.cpp:47-56), a header line then
      ``name<TAB>length<TAB>offset`` with offsets in the reference's index
      text, where one separator symbol follows every contig but the last;
    * ``output_path + ".nmask"``: the N runs (upper-cased 'N') as half-open
      ``start<TAB>end`` lines in the same coordinates (fast/fmindex.py:94-111).
    """
    contigs = list(genome_converter.read_fasta(fasta_path))
    arrays = genome_converter.encode_contigs(contigs, include_rc=True, include_n_mask=True,
                                             engine=engine, device=device)
    with open(output_path, "wb") as fh:
        np.savez_compressed(fh, **arrays)
    offset = 0
    with open(str(output_path) + ".meta", "w") as meta, \
            open(str(output_path) + ".nmask", "w") as nmask:
        meta.write(_META_HEADER + "\n")
        for k, (name, seq) in enumerate(contigs):
            if k:
                offset += 1  # the separator between contigs
            meta.write(f"{name}\t{len(seq)}\t{offset}\n")
            for start, end in _n_runs(seq).tolist():
                nmask.write(f"{offset + start}\t{offset + end}\n")
            offset += len(seq)
            if verbose:
                _LOGGER.info("  indexed %s (%d bp)", name, len(seq))
    if verbose:
        _LOGGER.info("Wrote %s (+ .meta, .nmask)", output_path)


def cold_region(cutoff: float) -> Tuple[int, int]:
    """(first 0-based cold position, cold length) for a cutoff (regions.py:80-123);
    ValueError when nothing is cold or the cold positions are not contiguous
    (fast/finder.py:166-176)."""
    if cutoff < 0:
        raise ValueError(f"cutoff must be non-negative, got {cutoff}")
    cold = [p for p in range(1, _GUIDE_LENGTH + 1) if 1.2 ** p >= cutoff]
    if not cold:
        raise ValueError(
            f"search cutoff {cutoff} leaves no exact 'cold' region; fast mode cannot enumerate "
            "that many tolerated mismatches. Use a lower cutoff (the PAM-proximal 20 bp caps it "
            "at 1.2**20 ~= 38.3) or the extensive mode.")
    if cold != list(range(cold[0], cold[-1] + 1)):
        raise ValueError("fast mode requires a contiguous cold region")
    return cold[0] - 1, len(cold)


def mismatch_budget(cutoff: float) -> int:
    """Largest m with 1.2 + 1.2**2 + ... + 1.2**m <= cutoff: no site with more
    mismatches can score <= cutoff."""
    total, m = 0.0, 0
    while m < _GUIDE_LENGTH and total + 1.2 ** (m + 1) <= cutoff:
        m += 1
        total += 1.2 ** m
    return m


def _planes(codes: np.ndarray) -> Tuple[int, int]:
    sh = np.arange(codes.size, dtype=np.uint32)
    c = codes.astype(np.uint32)
    return int(((c & 1) << sh).sum()), int(((c >> 1) << sh).sum())


_RC = str.maketrans("ACGTacgt", "TGCAtgca")


def _revcomp(s: str) -> str:
    return s.translate(_RC)[::-1]


def _n_bits20(nmask_le: Optional[np.ndarray], pos: np.ndarray) -> np.ndarray:
    """Bits pos..pos+19 of a little-endian N mask (0 without a mask)."""
    if nmask_le is None or pos.size == 0:
        return np.zeros(pos.size, dtype=np.uint32)
    m = np.asarray(nmask_le, dtype=np.uint8)
    byte = pos >> 3
    v = np.zeros(pos.size, dtype=np.uint64)
    for i in range(4):  # 20 bits + up to 7 of offset fit in 4 bytes
        idx = byte + i
        ok = idx < m.size
        v |= np.where(ok, m[np.minimum(idx, m.size - 1)], 0).astype(np.uint64) << np.uint64(8 * i)
    return ((v >> (pos & 7).astype(np.uint64)) & np.uint64(_W20)).astype(np.uint32)


def anchor_table(scan_guides, cutoff: float) -> np.ndarray:
    """Two anchors per scan key ('+' then '-'), as bgo_fast_anchor records
    (fast/finder.py:179-189): '+' reads guide[c0:20] on the forward genome,
    '-' reads rc(guide)[:m] there, i.e. guide[len - m:] on the reverse
    strand's site positions c0..19.  The index key range holds the keys whose
    first min(m, 12) bases are the anchor's.  Built per guide length with
    table lookups over all keys at once."""
    c0, m = cold_region(cutoff)
    k = min(m, 12)
    n_keys = len(scan_guides)
    seqs = [sg.guide_seq.upper() for sg in scan_guides]
    lens = np.array([len(x) for x in seqs], dtype=np.int64)
    if n_keys and int(lens.min()) < _GUIDE_LENGTH:
        raise ValueError(f"fast mode needs guides of at least {_GUIDE_LENGTH} nt")
    plus = np.zeros((n_keys, _GUIDE_LENGTH), dtype=np.uint32)
    minus = np.zeros((n_keys, _GUIDE_LENGTH), dtype=np.uint32)
    first = np.zeros((n_keys, 2, k), dtype=np.uint32)  # the anchors' first k forward bases
    for L in np.unique(lens).tolist():
        idx = np.flatnonzero(lens == L)
        group = [seqs[i] for i in idx]
        fwd = encode_dna_rows(group, L).astype(np.uint32)
        plus[idx] = fwd[:, :_GUIDE_LENGTH]
        minus[idx, :c0] = fwd[:, :c0]
        minus[idx, c0:] = fwd[:, L - m:]
        first[idx, 0] = fwd[:, c0:c0 + k]
        first[idx, 1] = reverse_complement_rows(group, L)[:, :k]
    pams = [sg.pam for sg in scan_guides]
    nib_of = {}
    for pam in set(pams):
        pm = encode_pam_mask(pam)[: len(pam)].astype(np.uint32)
        nib_of[pam] = int(np.bitwise_or.reduce(pm << (4 * np.arange(pm.size, dtype=np.uint32))))
    sh = np.arange(_GUIDE_LENGTH, dtype=np.uint32)

    def planes(c):
        return (((c & 1) << sh).sum(axis=1).astype(np.uint32),
                ((c >> 1) << sh).sum(axis=1).astype(np.uint32))

    key_sh = (2 * (11 - np.arange(k))).astype(np.uint32)
    lo = (first << key_sh).sum(axis=2).astype(np.uint32)  # [key, strand]
    out = np.zeros(2 * n_keys, dtype=_lib.FAST_ANCHOR_DTYPE)
    g0, g1 = planes(plus)
    for s, site in ((0, plus), (1, minus)):
        r = out[s::2]
        r["key_lo"] = lo[:, s]
        r["key_hi"] = lo[:, s] + np.uint32(1 << (2 * (12 - k)))
        r["a0"], r["a1"] = planes(site)
        r["g0"], r["g1"] = g0, g1
        r["n_code"] = ((plus == (0 if s == 0 else 3)).astype(np.uint32) << sh).sum(axis=1)
        r["pam_nibs"] = np.array([nib_of[p] for p in pams], dtype=np.uint32)
        r["guide"] = np.arange(n_keys, dtype=np.uint32)
        r["strand"] = s
        r["pam_len"] = np.array([len(p) for p in pams], dtype=np.uint8)
        r["cold_first"] = c0
    return out


class FastIndexGPU:
    """Anchor index of a resident genome in HBM (bgo_fast_index_*)."""

    def __init__(self, device_genome, chrom_offsets, chrom_lengths):
        self._genome = device_genome  # must outlive the index
        offs = np.ascontiguousarray(chrom_offsets, dtype=np.int64)
        lens = np.ascontiguousarray(chrom_lengths, dtype=np.int64)
        h = C.c_void_p()
        t0 = time.perf_counter()
        _lib.check(_lib.load().bgo_fast_index_create(device_genome.handle, _lib.ptr(offs),
                                                     _lib.ptr(lens), int(offs.size), C.byref(h)))
        self.build_s = time.perf_counter() - t0
        self._h = h
        self.score_table = np.array(cctop_table(), dtype=np.float64)

    @property
    def nbytes(self) -> int:
        n = C.c_int64(0)
        _lib.check(_lib.load().bgo_fast_index_bytes(self._h, C.byref(n)))
        return int(n.value)

    def offsets(self) -> np.ndarray:
        """The bucket table (uint32[2^24 + 1]) in host memory, copied once."""
        if getattr(self, "_offsets", None) is None:
            out = np.empty((1 << 24) + 1, dtype=np.uint32)
            _lib.check(_lib.load().bgo_fast_index_offsets(self._h, _lib.ptr(out)))
            self._offsets = out
        return self._offsets

    def positions(self, begin: int, end: int) -> np.ndarray:
        out = np.empty(max(0, int(end) - int(begin)), dtype=np.uint32)
        _lib.check(_lib.load().bgo_fast_index_positions(self._h, int(begin), int(end),
                                                        _lib.ptr(out)))
        return out

    def query(self, anchors: np.ndarray, cutoff: float, capacity: int = 1 << 16):
        """Candidate records of the anchors and the kernel time (ms); re-runs
        once with an exact-size buffer when the first was too small."""
        anchors = np.ascontiguousarray(anchors, dtype=_lib.FAST_ANCHOR_DTYPE)
        while True:
            out = np.empty(capacity, dtype=_lib.HIT_DTYPE)
            n, ms = C.c_int64(0), C.c_float(0.0)
            _lib.check(_lib.load().bgo_fast_index_query(
                self._h, _lib.ptr(anchors), int(anchors.size), float(cutoff),
                _lib.ptr(self.score_table), _lib.ptr(out), int(out.size), C.byref(n),
                C.byref(ms)))
            if n.value <= out.size:
                return out[: n.value], float(ms.value)
            capacity = int(n.value)

    def close(self) -> None:
        if self._h is not None:
            _lib.load().bgo_fast_index_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SaText:
    """A forward genome prepared for anchor suffix-array selection
    (bgo_sa_text; host memory, the words stay referenced here)."""

    def __init__(self, genome: dict):
        self._words = np.ascontiguousarray(genome["genome_forward"], dtype=np.uint32)
        self._offs = np.ascontiguousarray(genome["chrom_offsets"], dtype=np.int64)
        self._lens = np.ascontiguousarray(genome["chrom_lengths"], dtype=np.int64)
        h = C.c_void_p()
        _lib.check(_lib.load().bgo_sa_text_create(
            _lib.ptr(self._words), self._words.size, int(genome["total_bp"]),
            _lib.ptr(self._offs), _lib.ptr(self._lens), int(self._offs.size), C.byref(h)))
        self._h = h

    def select(self, positions: np.ndarray, anchor: str, cap: int) -> Tuple[np.ndarray, int]:
        """(the anchor occurrences kept under the cap, ascending; the anchor's
        occurrence count) among `positions`."""
        pos = np.ascontiguousarray(positions, dtype=np.uint32)
        codes = np.ascontiguousarray(encode_dna(anchor), dtype=np.uint8)
        out = np.empty(int(cap), dtype=np.uint32)
        n_occ, n = C.c_int64(0), C.c_int64(0)
        _lib.check(_lib.load().bgo_sa_text_select(self._h, _lib.ptr(pos), pos.size,
                                                  _lib.ptr(codes), codes.size, int(cap),
                                                  _lib.ptr(out), C.byref(n_occ), C.byref(n)))
        return out[:n.value], int(n_occ.value)

    def close(self) -> None:
        if self._h is not None:
            _lib.load().bgo_sa_text_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _hits_subset(h: _AllHits, keep: np.ndarray) -> _AllHits:
    def sub(a):
        return None if a is None else a[keep]
    return _AllHits(h.key[keep], h.strand[keep], h.pos[keep], h.n_mm[keep], sub(h.mm_mask),
                    sub(h.score), sub(h.x0), sub(h.x1), sub(h.mm_out))


class beguidedofffast:
    """Fast (anchored, score-cutoff) search with the reference's constructor
    and ``search`` arguments (fast/finder.py:90-128, 303-441).

    ``index_path`` names an encoded genome artefact (.npz): the MI355X fast
    mode scans the resident genome and needs no FM-index.  ``index`` may pass a
    preloaded extensive finder (``beguidedoff_amd.beguidedoff(...)``) whose
    device-resident genome is shared, as the reference's server passes a
    preloaded index handle (fast/finder.py:114-119).  ``genome``, ``device``
    and ``tuning`` are as for the extensive finder.  ``engine`` is "index"
    (default: an anchor index in HBM, built here) or "scan" (no index).
    """

    def __init__(self, index_path: Optional[str] = None, search_cutoff: float = 5.0,
                 scoring_method: str = "CCTOP", n_top_offtargets: int = 3,
                 max_hits_per_anchor: int = 10000, verbose: bool = False,
                 index: Optional[beguidedoff] = None, genome: Optional[dict] = None,
                 device: Optional[str] = None, tuning: Optional[dict] = None,
                 engine: str = "index"):
        if scoring_method.upper() != "CCTOP":
            raise ValueError("fast mode only supports scoring_method='CCTOP'")
        self.index_path = str(index_path) if index_path is not None else ""
        self.search_cutoff = float(search_cutoff)
        self.scoring_method = "CCTOP"
        self.n_top_offtargets = int(n_top_offtargets)
        self.max_hits_per_anchor = int(max_hits_per_anchor)
        self.verbose = verbose
        if index is not None:
            self._ext = index
        else:
            if not self.index_path and genome is None:
                raise ValueError("fast mode requires either index_path or a preloaded index")
            self._ext = beguidedoff(genome_path=self.index_path or None, genome=genome,
                                    device=device, tuning=tuning,
                                    n_top_offtargets=self.n_top_offtargets)
        self._ext.n_top_offtargets = self.n_top_offtargets
        if engine not in ("index", "scan"):
            raise ValueError(f"engine must be 'index' or 'scan', got {engine!r}")
        self.engine = engine
        self._index: Optional[FastIndexGPU] = None
        self._sa_text: Optional[SaText] = None
        if engine == "index":
            self._index = FastIndexGPU(self._ext._device_genome(), self._ext.chrom_offsets,
                                       self._ext.chrom_lengths)

    @property
    def total_bp(self) -> int:
        return self._ext.total_bp

    @property
    def chrom_names(self) -> List[str]:
        return self._ext.chrom_names

    def _candidates(self, scan_guides, cutoff: float,
                    log: Optional[logging.Logger] = None) -> _AllHits:
        """Fast-mode candidates of every scan key, as flat hits whose score,
        mismatch mask and count include the N positions, under the anchor
        occurrence cap."""
        if self._index is not None:
            hits = self._candidates_index(scan_guides, cutoff)
        else:
            hits = self._candidates_scan(scan_guides, cutoff)
        return self._anchor_cap(scan_guides, cutoff, hits, log or _LOGGER)

    def _anchor_cap(self, scan_guides, cutoff: float, hits: _AllHits,
                    log: logging.Logger) -> _AllHits:
        """Keep only candidates anchored at the first max_hits_per_anchor
        occurrences of their anchor in the reference's suffix-array order
        (fast/finder.py:149-159, csrc/fmindex_binding.cpp:164-176), and warn
        for each guide whose anchor has more (fast/finder.py:368-375)."""
        cap = self.max_hits_per_anchor
        if cap <= 0 or not scan_guides:
            return hits
        if self._index is None:  # the scan engine counts occurrences with an index too
            self._index = FastIndexGPU(self._ext._device_genome(), self._ext.chrom_offsets,
                                       self._ext.chrom_lengths)
        anchors = anchor_table(scan_guides, cutoff)
        offs = self._index.offsets()
        lo = offs[anchors["key_lo"]].astype(np.int64)
        hi = offs[anchors["key_hi"]].astype(np.int64)
        over = np.flatnonzero(hi - lo > cap)  # bucket counts bound the occurrence counts
        if over.size == 0:
            return hits
        if self._sa_text is None:
            self._sa_text = SaText(self._ext.genome_data)
        c0, m = cold_region(cutoff)
        ext = self._ext
        keep = np.ones(hits.key.size, dtype=bool)
        warned = set()
        for a in over.tolist():
            k, strand = a // 2, a % 2
            sg = scan_guides[k]
            g = sg.guide_seq.upper()
            anchor = g[c0:c0 + m] if strand == 0 else _revcomp(g)[:m]
            kept, n_occ = self._sa_text.select(self._index.positions(lo[a], hi[a]), anchor, cap)
            if n_occ <= cap:
                continue
            if (sg.guide_seq, sg.pam) not in warned:
                warned.add((sg.guide_seq, sg.pam))
                log.warning("Anchor hit cap (%d) exceeded for guide %s (PAM=%s); results may "
                            "be incomplete.", cap, sg.guide_seq, sg.pam)
            rows = np.flatnonzero((hits.key == k) & (hits.strand == strand))
            if rows.size == 0:
                continue
            pos = hits.pos[rows]
            plen = len(sg.pam)
            if strand == 0:  # the '+' anchor starts at the site's cold position c0
                occ = pos + c0
            else:  # '-': the forward window's leftmost + PAM length
                ci = np.searchsorted(ext.chrom_offsets, pos, side="right") - 1
                left = (ext.chrom_offsets[ci] + ext.chrom_lengths[ci]
                        - (pos - ext.chrom_offsets[ci]) - (_GUIDE_LENGTH + plen))
                occ = left + plen
            keep[rows] = np.isin(occ, kept.astype(np.int64))
        return _hits_subset(hits, keep) if not keep.all() else hits

    def _candidates_index(self, scan_guides, cutoff: float) -> _AllHits:
        anchors = anchor_table(scan_guides, cutoff)
        h, self.last_kernel_ms = self._index.query(anchors, cutoff)
        key = h["guide"].astype(np.int64)
        x0, x1 = h["x0"].astype(np.uint32), h["x1"].astype(np.uint32)
        g0 = anchors["g0"][0::2][key] if key.size else np.zeros(0, np.uint32)
        g1 = anchors["g1"][0::2][key] if key.size else np.zeros(0, np.uint32)
        mm2 = ((x0 ^ g0) | (x1 ^ g1)) & np.uint32(_W20)
        return _AllHits(key, h["strand"].astype(np.int8), h["pos"].astype(np.int64),
                        h["n_mm"].astype(np.int64), mm2, h["score"].astype(np.float64), x0, x1,
                        h["mm_mask"].astype(np.uint32))

    def close(self) -> None:
        if self._index is not None:
            self._index.close()
            self._index = None
        if self._sa_text is not None:
            self._sa_text.close()
            self._sa_text = None

    def _candidates_scan(self, scan_guides, cutoff: float) -> _AllHits:
        from .batched import GuideSpec, scan

        ext = self._ext
        c0, _ = cold_region(cutoff)
        budget = mismatch_budget(cutoff)
        plen_k = np.array([len(sg.pam) for sg in scan_guides], dtype=np.int64)
        specs = [GuideSpec(sg.guide_seq, sg.pam, budget) for sg in scan_guides]
        res = scan(ext._device_genome(), specs, tuning=ext.tuning,
                   end_guard=_GUIDE_LENGTH + int(plen_k.min()))
        h = res.hits
        key = h["guide"].astype(np.int64)
        strand = h["strand"].astype(np.int8)
        pos = h["pos"].astype(np.int64)
        x0, x1 = h["x0"].astype(np.uint32), h["x1"].astype(np.uint32)
        mm2 = h["mm_mask"].astype(np.uint32)

        # window inside one contig (same contig layout on both strand arrays)
        offs, lens = ext.chrom_offsets, ext.chrom_lengths
        k = np.searchsorted(offs, pos, side="right") - 1
        inside = (k >= 0) & (pos + _GUIDE_LENGTH + plen_k[key] <= offs[k] + lens[k])

        # anchors on the site's cold positions c0..19, per key and strand
        cold = np.uint32(((1 << (_GUIDE_LENGTH - c0)) - 1) << c0)
        anchors = anchor_table(scan_guides, cutoff).reshape(len(scan_guides), 2)
        s_idx = strand.astype(np.int64)
        anc = anchors[key, s_idx]
        anchor_ok = (((x0 ^ anc["a0"]) | (x1 ^ anc["a1"])) & cold) == 0

        # N positions that count as mismatches (fast/finder.py:227-257)
        nbits = np.zeros(pos.size, dtype=np.uint32)
        for s, _label, _words, nm in ext._strand_arrays():
            sel = strand == s
            if sel.any():
                nbits[sel] = _n_bits20(nm, pos[sel])
        mm = mm2 | (nbits & anc["n_code"])
        score = cctop_scores(mm)
        keep = inside & anchor_ok & (score <= cutoff)
        n_mm = np.bitwise_count(mm[keep]).astype(np.int64)
        return _AllHits(key[keep], strand[keep], pos[keep], n_mm, mm2[keep], score[keep],
                        x0[keep], x1[keep], mm[keep])

    def search(self, input_file: Optional[str] = None, guides: Optional[List[GuideInput]] = None,
               return_all: bool = False, count_score_threshold: Optional[float] = None,
               output_file: Optional[str] = None, include_non_editable: bool = False,
               presets_path: Optional[str] = None, verbose: bool = True,
               logger: Optional[logging.Logger] = None) -> pd.DataFrame:
        """Fast search (fast/finder.py:303-441): same inputs and 12-column table."""
        if (input_file is None) == (guides is None):
            raise ValueError("Exactly one of input_file or guides must be provided")
        if input_file is not None:
            guides = load_guides_from_tsv(input_file, presets_path=presets_path or str(_PROFILES_TSV))
        if not guides:
            raise ValueError("No guides provided")
        log = logger if logger is not None else _LOGGER
        for g in guides:
            if g.pam_position.lower() != "3prime":
                raise ValueError(f"Guide {g.guide_id} has pam_position '{g.pam_position}'. "
                                 "Only 3prime PAM is supported.")
        cutoff = self.search_cutoff
        if count_score_threshold is not None and count_score_threshold > cutoff:
            log.warning("count_score_threshold=%s exceeds search_cutoff=%s; widening the search "
                        "region (this is slower).", count_score_threshold, cutoff)
            cutoff = float(count_score_threshold)
        scan_guides, mapping = self._ext._deduplicate_guides(guides)
        if verbose:
            log.info("Fast search: %d unique protospacers (from %d input rows), cutoff=%s",
                     len({(sg.guide_seq, sg.pam) for sg in scan_guides}), len(guides), cutoff)
        hits = self._candidates(scan_guides, cutoff, log)
        df = assemble_results(self._ext, guides, scan_guides, mapping, hits,
                              count_score_threshold, return_all)
        if output_file:
            df.to_csv(output_file, sep="\t", index=False)
            if verbose:
                log.info("Wrote results to %s", output_file)
        return df
