"""ORACLE — finder-level restatement (test infrastructure, not product code).

Restates the reference's extensive ``search()`` host logic in its own
per-hit style, over the C scan oracle (oracle/bgo_oracle.c) instead of a GPU:

  dedup + on-target key        beguidedoff/finder.py:361-403
  per-strand scan + decode     beguidedoff/finder.py:119-249, 640-679
  N masks (LE, per-contig RC)  beguidedoff/finder.py:310-345
  per-hit epilogue             beguidedoff/finder.py:405-472
  sort / top-N / table         beguidedoff/finder.py:545-566, io.py:155-180
  codec / PAM / CCTOP / edit   encoding.py:6-102, pam.py:7-45, scoring.py:5-24,
                               base_editing.py:9-46

It is exact (no result buffers), so it has no overflow escalation.  It does
not import beguidedoff_amd: the small shared helpers are restated here.
"""
from __future__ import annotations

from collections import defaultdict
from typing import Dict, List, Optional

import numpy as np
import pandas as pd

from .scan import scan_strand

_NUC = {"A": 0, "C": 1, "G": 2, "T": 3}
_IUPAC = {"A": 1, "C": 2, "G": 4, "T": 8, "R": 5, "Y": 10, "M": 3, "K": 12, "S": 6, "W": 9,
          "H": 11, "B": 14, "V": 7, "D": 13, "N": 15}
_COLUMNS = ["guide_rna_sequence", "base_editor", "ot_rank", "ot_chromosome", "ot_position",
            "ot_strand", "ot_sequence", "ot_pam", "n_mismatches", "mismatch_positions", "score",
            "n_offtargets_below_threshold"]


def encode(seq: str) -> List[int]:
    return [_NUC.get(c, 0) for c in seq.upper()]


def pack(codes: List[int]) -> int:
    v = 0
    for i, c in enumerate(codes):
        v |= c << (2 * i)
    return v


def mismatch_positions(a: int, b: int) -> List[int]:
    x = a ^ b
    bits = (x | (x >> 1)) & 0x5555555555
    return [i + 1 for i in range(20) if (bits >> (2 * i)) & 1]


def cctop(positions: List[int]) -> float:
    if not positions:
        return 0.0
    return sum(1.2 ** p for p in positions)


def editable(seq: str, editor_type: str, start: int, end: int) -> List[int]:
    target = {"CBE": "C", "ABE": "A", "CGE": "C"}.get(editor_type)
    if target is None:
        raise ValueError(f"Unknown editor_type: {editor_type}")
    return [i + 1 for i in range(start - 1, end) if i < len(seq) and seq[i] == target]


def pam_mask(pam: str) -> np.ndarray:
    pam = pam.upper()
    if not 1 <= len(pam) <= 6:
        raise ValueError(f"PAM must be 1-6 characters, got {len(pam)}")
    m = np.zeros(6, dtype=np.uint8)
    for i, c in enumerate(pam):
        m[i] = _IUPAC[c]
    return m


def base(words: np.ndarray, i: int) -> int:
    return (int(words[i // 16]) >> ((i % 16) * 2)) & 3


class FinderOracle:
    def __init__(self, genome: dict, max_mismatches: int = 4, n_top_offtargets: int = 3,
                 threads: Optional[int] = None):
        self.g = genome
        self.max_mismatches = max_mismatches
        self.n_top = n_top_offtargets
        self.threads = threads
        self.total_bp = int(genome["total_bp"])
        self.names = [str(c) for c in genome["chrom_names"]]
        self.offsets = np.asarray(genome["chrom_offsets"], dtype=np.int64)
        self.lengths = np.asarray(genome["chrom_lengths"], dtype=np.int64)
        self._masks()

    def _masks(self):
        big = self.g.get("n_mask")
        self.nm_fwd = self.nm_rc = None
        if big is None or np.asarray(big).size == 0:
            return
        bits = np.unpackbits(np.asarray(big, dtype=np.uint8), bitorder="big", count=self.total_bp)
        if not bits.any():
            return
        self.nm_fwd = np.pad(np.packbits(bits, bitorder="little"), (0, 16))
        rc = np.empty(self.total_bp, dtype=np.uint8)
        for off, ln in zip(self.offsets.tolist(), self.lengths.tolist()):
            rc[off:off + ln] = bits[off:off + ln][::-1]
        self.nm_rc = np.pad(np.packbits(rc, bitorder="little"), (0, 16))

    def _pos_to_chrom(self, pos: int):
        idx = int(np.searchsorted(self.offsets, pos, side="right") - 1)
        return idx, int(pos - self.offsets[idx])

    def _keys(self, guides):
        mapping = defaultdict(list)
        for g in guides:
            ot_idx = ot_local = ot_strand = None
            if g.on_target_chrom and g.on_target_pos is not None:
                ot_idx = next((i for i, n in enumerate(self.names) if n == g.on_target_chrom), None)
                if ot_idx is not None:
                    if g.on_target_strand is None:
                        raise ValueError(f"Guide {g.guide_id} has on-target chromosome and "
                                         "position but no strand.")
                    w = 20 + len(g.pam)
                    ot_local = g.on_target_pos - 1 if g.on_target_strand == "+" else g.on_target_pos - w
                    ot_strand = g.on_target_strand
            mapping[(g.guide_seq, g.pam, ot_idx, ot_local, ot_strand)].append(g)
        return list(mapping.keys()), mapping

    def raw_hits(self, keys):
        """key -> list of per-hit dicts, '+' strand first (finder.py:640-679)."""
        out = {k: [] for k in keys}
        strands = [("+", self.g["genome_forward"], self.nm_fwd, 0)]
        if "genome_rc" in self.g:
            strands.append(("-", self.g["genome_rc"], self.nm_rc, 3))
        gvals = np.array([pack(encode(k[0])) & ((1 << 64) - 1) for k in keys], dtype=np.uint64)
        masks = np.stack([pam_mask(k[1]) for k in keys]) if keys else np.zeros((0, 6), np.uint8)
        lens = np.array([len(k[1]) for k in keys], dtype=np.uint8)
        for label, words, nm, n_enc in strands:
            gi, pos, mm = scan_strand(words, self.total_bp, gvals, masks, lens,
                                      np.ones(len(keys), np.uint8), self.max_mismatches, nm,
                                      n_enc, threads=self.threads)
            for g, p, m in zip(gi.tolist(), pos.tolist(), mm.tolist()):
                key = keys[g]
                plen = len(key[1])
                window = 20 + plen
                if label == "+":
                    leftmost = report = p
                else:
                    idx, local_rc = self._pos_to_chrom(p)
                    leftmost = int(self.offsets[idx] + self.lengths[idx] - local_rc - window)
                    report = leftmost + window - 1
                if key[2] is not None:
                    ci, cl = self._pos_to_chrom(leftmost)
                    if ci == key[2] and cl == key[3] and label == key[4]:
                        continue
                seq = "".join("ACGT"[base(words, p + i)] for i in range(20))
                pam = "".join("ACGT"[base(words, p + 20 + i)] for i in range(plen))
                out[key].append({"position": report, "n_mismatches": m, "offtarget_seq": seq,
                                 "genomic_pam": pam, "strand": label, "scanned_pos": p})
        return out

    def search(self, guides, return_all: bool = False,
               count_score_threshold: Optional[float] = None) -> pd.DataFrame:
        keys, mapping = self._keys(guides)
        raw = self.raw_hits(keys)
        groups: Dict[str, list] = defaultdict(list)
        below: Dict[str, int] = defaultdict(int)
        for key, hits in raw.items():
            gval = pack(encode(key[0]))
            for h in hits:
                idx, local = self._pos_to_chrom(h["position"])
                mm = mismatch_positions(gval, pack(encode(h["offtarget_seq"])))
                score = cctop(mm)
                for g in mapping[key]:
                    ed = editable(h["offtarget_seq"], g.editor_type, g.edit_window_start,
                                  g.edit_window_end)
                    if not return_all and not ed:
                        continue
                    groups[g.guide_id].append({
                        "chrom": self.names[idx], "position": local + 1, "strand": h["strand"],
                        "seq": h["offtarget_seq"], "pam": h["genomic_pam"], "guide": key[0],
                        "base_editor": g.base_editor or "", "n_mm": h["n_mismatches"],
                        "mm": mm, "score": score})
                    if count_score_threshold is not None and ed and score <= count_score_threshold:
                        below[g.guide_id] += 1
        dfs = []
        for g in guides:
            hits = groups.get(g.guide_id, [])
            hits.sort(key=lambda x: (x["score"], x["chrom"], x["position"], x["strand"]))
            if not return_all:
                hits = hits[: self.n_top]
            if hits:
                dfs.append(pd.DataFrame([{
                    "guide_rna_sequence": h["guide"], "base_editor": h["base_editor"],
                    "ot_rank": r, "ot_chromosome": h["chrom"], "ot_position": h["position"],
                    "ot_strand": h["strand"], "ot_sequence": h["seq"], "ot_pam": h["pam"],
                    "n_mismatches": h["n_mm"], "mismatch_positions": ",".join(map(str, h["mm"])),
                    "score": round(h["score"], 6),
                    "n_offtargets_below_threshold": below.get(g.guide_id, 0),
                } for r, h in enumerate(hits, start=1)], columns=_COLUMNS))
        return pd.concat(dfs, ignore_index=True) if dfs else pd.DataFrame()
