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

The table is the extensive finder's (same 12 columns, sort and top-N).
Differences from the reference, by construction: the anchor occurrence cap
(`max_hits_per_anchor`, first hits in suffix-array order) is not applied, so
results are complete; the reference's cap is reached only by anchors with more
than 10,000 genome occurrences.
"""
from __future__ import annotations

import logging
from typing import List, Optional, Tuple

import numpy as np
import pandas as pd

from .finder import (_AllHits, _W20, _LOGGER, _PROFILES_TSV, assemble_results, beguidedoff,
                     cctop_scores)
from .encoding import encode_dna
from .io import load_guides_from_tsv
from .types import GuideInput

_GUIDE_LENGTH = 20


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


class beguidedofffast:
    """Fast (anchored, score-cutoff) search with the reference's constructor
    and ``search`` arguments (fast/finder.py:90-128, 303-441).

    ``index_path`` names an encoded genome artefact (.npz): the MI355X fast
    mode scans the resident genome and needs no FM-index.  ``index`` may pass a
    preloaded extensive finder (``beguidedoff_amd.beguidedoff(...)``) whose
    device-resident genome is shared, as the reference's server passes a
    preloaded index handle (fast/finder.py:114-119).  ``genome``, ``device``
    and ``tuning`` are as for the extensive finder.
    """

    def __init__(self, index_path: Optional[str] = None, search_cutoff: float = 5.0,
                 scoring_method: str = "CCTOP", n_top_offtargets: int = 3,
                 max_hits_per_anchor: int = 10000, verbose: bool = False,
                 index: Optional[beguidedoff] = None, genome: Optional[dict] = None,
                 device: Optional[str] = None, tuning: Optional[dict] = None):
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

    @property
    def total_bp(self) -> int:
        return self._ext.total_bp

    @property
    def chrom_names(self) -> List[str]:
        return self._ext.chrom_names

    def _candidates(self, scan_guides, cutoff: float) -> _AllHits:
        """Fast-mode candidates of every scan key, as flat hits whose score,
        mismatch mask and count include the N positions."""
        from .batched import GuideSpec, scan

        ext = self._ext
        c0, m = cold_region(cutoff)
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
        anchor = np.zeros((len(scan_guides), 2, 2), dtype=np.uint32)  # [key, strand, plane]
        n_code = np.zeros((len(scan_guides), 2), dtype=np.uint32)      # guide == N code
        for i, sg in enumerate(scan_guides):
            g = encode_dna(sg.guide_seq.upper())
            plus = np.zeros(_GUIDE_LENGTH, dtype=np.uint8)
            plus[: min(g.size, _GUIDE_LENGTH)] = g[:_GUIDE_LENGTH]
            minus = plus.copy()
            minus[c0:] = g[g.size - m:]               # rc(guide)[:m] read on the site
            anchor[i, 0] = _planes(plus)
            anchor[i, 1] = _planes(minus)
            sh = np.arange(_GUIDE_LENGTH, dtype=np.uint32)
            n_code[i, 0] = int(((plus == 0).astype(np.uint32) << sh).sum())
            n_code[i, 1] = int(((plus == 3).astype(np.uint32) << sh).sum())
        s_idx = strand.astype(np.int64)
        a0, a1 = anchor[key, s_idx, 0], anchor[key, s_idx, 1]
        anchor_ok = (((x0 ^ a0) | (x1 ^ a1)) & cold) == 0

        # N positions that count as mismatches (fast/finder.py:227-257)
        nbits = np.zeros(pos.size, dtype=np.uint32)
        for s, _label, _words, nm in ext._strand_arrays():
            sel = strand == s
            if sel.any():
                nbits[sel] = _n_bits20(nm, pos[sel])
        mm = mm2 | (nbits & n_code[key, s_idx])
        score = cctop_scores(mm)
        keep = inside & anchor_ok & (score <= cutoff)
        n_mm = np.array([bin(int(v)).count("1") for v in mm[keep]], dtype=np.int64)
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
        hits = self._candidates(scan_guides, cutoff)
        df = assemble_results(self._ext, guides, scan_guides, mapping, hits,
                              count_score_threshold, return_all)
        if output_file:
            df.to_csv(output_file, sep="\t", index=False)
            if verbose:
                log.info("Wrote results to %s", output_file)
        return df
