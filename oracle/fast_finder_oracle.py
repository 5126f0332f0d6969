"""ORACLE — fast-mode table (test infrastructure, not product code).

Restates the reference's fast search (beguidedoff/fast/finder.py:303-441) in
its own per-candidate style over the C restatement of the anchored candidate
search (oracle/bgo_fast.c via oracle/fastmode.py, which follows
fast/finder.py:163-277): per unique (guide, PAM) the candidates, then per
input guide the on-target skip (:280-298), the editable filter, the
below-threshold count, the (score, chromosome, position, strand) sort, top-N
and the 12-column table (io.py:155-180).  20-nt guides only, as the
candidate oracle.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

import numpy as np
import pandas as pd

from .fastmode import FastIndex, site_string
from .finder_oracle import _COLUMNS, base, editable

_COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def _rc(s: str) -> str:
    return "".join(_COMP[c] for c in reversed(s))


class FastFinderOracle:
    def __init__(self, genome: dict, search_cutoff: float = 5.0, n_top_offtargets: int = 3,
                 threads: int = 2):
        self.genome = genome
        self.cutoff = float(search_cutoff)
        self.n_top = int(n_top_offtargets)
        self.names = [str(n) for n in genome["chrom_names"]]
        self.offs = [int(x) for x in genome["chrom_offsets"]]
        self.lens = [int(x) for x in genome["chrom_lengths"]]
        self.words = genome["genome_forward"]
        self.index = FastIndex(genome, threads=threads)

    def _chrom(self, pos: int) -> int:
        return int(np.searchsorted(np.array(self.offs), pos, side="right") - 1)

    def _bases(self, start: int, n: int) -> str:
        return "".join("ACGT"[base(self.words, start + i)] for i in range(n))

    def _on_target(self, g) -> Optional[Tuple[int, int, str]]:
        if not g.on_target_chrom or g.on_target_pos is None:
            return None
        if g.on_target_chrom not in self.names:
            return None
        if g.on_target_strand is None:
            raise ValueError("on-target strand missing")
        idx = self.names.index(g.on_target_chrom)
        window = 20 + len(g.pam)
        local = g.on_target_pos - 1 if g.on_target_strand == "+" else g.on_target_pos - window
        return idx, local, g.on_target_strand

    def search(self, guides, return_all: bool = False,
               count_score_threshold: Optional[float] = None) -> pd.DataFrame:
        cutoff = self.cutoff
        if count_score_threshold is not None and count_score_threshold > cutoff:
            cutoff = float(count_score_threshold)
        groups: Dict[Tuple[str, str], list] = {}
        for g in guides:
            groups.setdefault((g.guide_seq, g.pam), []).append(g)
        keys = list(groups)
        found = self.index.search(keys, cutoff=cutoff)
        rows: Dict[str, list] = {g.guide_id: [] for g in guides}
        below: Dict[str, int] = {g.guide_id: 0 for g in guides}
        for (seq, pam), hits in zip(keys, found):
            plen = len(pam)
            window = 20 + plen
            cands = []
            for lm, st, sv, mm, sc in zip(hits.leftmost.tolist(), hits.strand.tolist(),
                                          hits.site.tolist(), hits.mm.tolist(),
                                          hits.score.tolist()):
                k = self._chrom(lm)
                local = lm - self.offs[k]
                strand = "+" if st == 0 else "-"
                if strand == "+":
                    pam_seq = self._bases(lm + 20, plen)
                    position = local + 1
                else:
                    pam_seq = _rc(self._bases(lm, plen))
                    position = local + window
                pos_list = [i + 1 for i in range(20) if (mm >> i) & 1]
                cands.append((k, local, strand, position, site_string(sv), pam_seq, pos_list, sc))
            for g in groups[(seq, pam)]:
                ot = self._on_target(g)
                for k, local, strand, position, site, pam_seq, pos_list, sc in cands:
                    if ot is not None and (k, local, strand) == ot:
                        continue
                    ed = editable(site, g.editor_type, g.edit_window_start, g.edit_window_end)
                    if not return_all and not ed:
                        continue
                    rows[g.guide_id].append({
                        "chrom": self.names[k], "position": position, "strand": strand,
                        "seq": site, "pam": pam_seq, "guide": seq,
                        "base_editor": g.base_editor or "", "n_mm": len(pos_list),
                        "mm": pos_list, "score": sc})
                    if count_score_threshold is not None and ed and sc <= count_score_threshold:
                        below[g.guide_id] += 1
        dfs = []
        for g in guides:
            hits = rows[g.guide_id]
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
                    "n_offtargets_below_threshold": below[g.guide_id],
                } for r, h in enumerate(hits, start=1)], columns=_COLUMNS))
        return pd.concat(dfs, ignore_index=True) if dfs else pd.DataFrame()
