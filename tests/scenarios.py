"""Known-answer scenarios of the reference's test suite, as genome builders.

Each function returns (contigs, guides, kwargs) where contigs are
(name, sequence) pairs; expectations are asserted by the tests that use them.
Sources (the reference's tests, read for their inputs and expected values):
  repeat_genome          tests/test_extensive_gpu.py:34-107
  n_gate_genome          tests/test_extensive_gpu.py:116-156
  n_adjacent_genome      tests/test_extensive_gpu.py:158-182
  no_n_mask_genome       tests/test_extensive_gpu.py:184-204
  auto_chunk             tests/test_extensive_gpu.py:207-233
  exotic_pam             tests/test_extensive_gpu.py:302-383
  search positions       tests/test_search.py:12-118
  count thresholds       tests/test_count_score_threshold.py:18-102
"""
import random
from itertools import product

import numpy as np

from beguidedoff_amd.types import GuideInput
from brute import revcomp

GUIDE = "ACGTACGTACGTACGTACGT"
N_COPIES = 12_000


def repeat_genome():
    unit = "T" + GUIDE[1:] + "AGG"
    return [("chr1", unit * N_COPIES + "A" * 30)]


def single_guide():
    return [GuideInput("g1", GUIDE, "NGG", "3prime", 4, 8, "CBE")]


def n_gate_genome():
    return [("chr1", "G" * 40 + "T" * 20 + "TTT" + "G" * 40 + "N" * 500
             + "G" * 40 + "A" * 20 + "AAA" + "G" * 40)]


def poly_guides():
    return [GuideInput("polyA", "A" * 20, "AAA", "3prime", 4, 8, "CBE"),
            GuideInput("polyT", "T" * 20, "TTT", "3prime", 4, 8, "CBE")]


def n_adjacent_genome():
    return [("chr1", "G" * 30 + "N" * 10 + "CCC" + "T" * 20 + "G" * 30)]


def no_n_mask_genome():
    return [("chr1", "G" * 40 + "A" * 20 + "AAA" + "G" * 40)]


def auto_chunk():
    rng = np.random.default_rng(7)
    genome = list(rng.choice(list("ACGT"), size=30_000))
    guides = [GuideInput(f"g{i}", "ACGTACGTACGTAC" + "".join(p), "NGG", "3prime", 4, 8, "CBE")
              for i, p in enumerate(product("ACGT", repeat=6))][:2100]
    planted = guides[137]
    for pos in (5_000, 15_000, 25_000):
        genome[pos:pos + 20] = list(planted.guide_seq)
        genome[pos + 20:pos + 23] = list("AGG")
    return [("chr1", "".join(genome))], guides, planted


RESOLVE = {"A": "A", "C": "C", "G": "G", "T": "T", "R": "A", "Y": "C", "M": "A", "K": "G",
           "S": "C", "W": "A", "H": "A", "B": "C", "V": "A", "D": "A", "N": "A"}
VIOLATE = {"V": "T", "S": "A", "W": "C", "B": "A", "D": "C", "H": "G"}


def exotic_pam(seed: int = 11):
    """Sites for exotic-code PAMs planted in a random genome; the expected
    answer is the per-contig brute force with cutoff 5.0 and the CBE 4-8
    editability filter (tests/test_extensive_gpu.py:268-299)."""
    rnd = random.Random(seed)
    specs = [("TTTV", "CBE"), ("SGG", "CBE"), ("WGG", "CBE"), ("NGG", "CBE")]
    guides, c_idx = [], []
    for _ in specs:
        g = [rnd.choice("ACGT") for _ in range(20)]
        c = rnd.randrange(3, 8)
        g[c] = "C"
        c_idx.append(c)
        guides.append("".join(g))
    seq = "".join(rnd.choice("ACGT") for _ in range(24_000))
    pos = 200
    for guide, c0, (pam, _) in zip(guides, c_idx, specs):
        mut_range = [i for i in range(5) if i != c0]
        for n_mm in (0, 0, 1, 2):
            site = list(guide)
            for p in rnd.sample(mut_range, n_mm):
                site[p] = rnd.choice([b for b in "ACGT" if b != site[p]])
            seq = seq[:pos] + "".join(site) + "".join(RESOLVE[c] for c in pam) + seq[pos + 24:]
            pos += 60
        bad = next((c for c in pam if c in VIOLATE), None)
        if bad is not None:
            pam_bad = "".join(VIOLATE[c] if c == bad else RESOLVE[c] for c in pam)
            seq = seq[:pos] + guide + pam_bad + seq[pos + 24:]
            pos += 60
    inputs = [GuideInput(f"g{i}", g, pam, "3prime", 4, 8, et)
              for i, (g, (pam, et)) in enumerate(zip(guides, specs))]
    return [("chr1", seq)], inputs


def minus_strand_genome():
    rc_target = revcomp(GUIDE + "AGG")
    return [("chr1", "A" * 14 + rc_target + "A" * 23 + rc_target + "A" * 17)]


def offtarget_pos31_genome():
    ot = GUIDE[:2] + "A" + GUIDE[3:] + "AGG"
    return [("chr1", "A" * 30 + ot + "A" * 100)]


def _mutate(seq: str, pos: int) -> str:
    alt = {"A": "C", "C": "A", "G": "T", "T": "G"}[seq[pos - 1]]
    return seq[:pos - 1] + alt + seq[pos:]


def threshold_genome():
    return [("chr1", GUIDE + "AGG" + "A" * 30 + _mutate(GUIDE, 1) + "AGG" + "A" * 30
             + _mutate(GUIDE, 4) + "AGG" + "A" * 30 + _mutate(GUIDE, 8) + "AGG" + "A" * 100)]


def threshold_guide():
    return GuideInput("g1", GUIDE, "NGG", "3prime", 4, 8, "CBE", on_target_chrom="chr1",
                      on_target_pos=1, on_target_strand="+")


TIE_GUIDE = "GATTCAGCTAGGCTCATGCA"


def chrom_name_tie():
    """Twelve contigs chr1..chr12 in numeric order, each with an exact copy of
    TIE_GUIDE + AGG at local 1-based position 100 + 7k (chr k), plus a second
    exact copy on chr10 at 1,500.  Every copy scores 0, so the order of the
    rows is decided by the chromosome name as a string, then the position
    (finder.py:552): chr1, chr10, chr10, chr11, chr12, chr2, ..., chr9.
    Returns (contigs, expected [(chromosome, position)] of the 0-score rows in
    table order)."""
    rng = np.random.default_rng(52)
    contigs, zero = [], []
    for k in range(1, 13):
        seq = list(rng.choice(list("ACGT"), size=2400))
        sites = [100 + 7 * k] + ([1500] if k == 10 else [])
        for p in sites:
            seq[p - 4:p - 1] = list("TTT")
            seq[p - 1:p + 22] = list(TIE_GUIDE + "AGG")
            zero.append((f"chr{k}", p))
        contigs.append((f"chr{k}", "".join(seq)))
    return contigs, sorted(zero)


def tie_guide():
    return GuideInput("tie", TIE_GUIDE, "NGG", "3prime", 4, 8, "CBE")
