"""Independent brute-force scanners over plain strings (test helpers).

`kernel_sites` reproduces the reference kernel's semantics over the
concatenated strand strings: windows may span contig junctions and positions
>= total_bp - 26 are never scanned (beguidedoff/cuda_search.py:208-213);
the '-' strand string is the concatenation of per-contig reverse complements
(genome_converter.py:66-71); an N reads as A on '+' and T on '-' for the PAM
and counts as one mismatch in the 20-nt window (cuda_search.py:262-308).

`contig_sites` is the per-contig semantics of the reference's own test
oracles (tests/_fastutil.py:56-106, tests/test_extensive_gpu.py:268-299):
only windows that lie inside one contig, both strands, positions reported
like the finder (1-based leftmost on '+', rightmost on '-').
"""
from typing import Dict, List, Sequence, Tuple

IUPAC = {"A": "A", "C": "C", "G": "G", "T": "T", "R": "AG", "Y": "CT", "M": "AC", "K": "GT",
         "S": "CG", "W": "AT", "H": "ACT", "B": "CGT", "V": "ACG", "D": "AGT", "N": "ACGT"}
COMP = {"A": "T", "C": "G", "G": "C", "T": "A", "N": "N"}  # genomes here are ACGTN only


def revcomp(s: str) -> str:
    return "".join(COMP[c] for c in reversed(s))


def strand_strings(contigs: Sequence[Tuple[str, str]]):
    fwd = "".join(s.upper() for _, s in contigs)
    rc = "".join(revcomp(s.upper()) for _, s in contigs)
    return fwd, rc


def _guide20(seq: str) -> str:
    g = "".join(c if c in "ACGT" else "A" for c in seq.upper()[:20])
    return g + "A" * (20 - len(g))


def _count(window: str, g: str, n_as_mismatch: bool, n_code: str) -> int:
    """Mismatches of a 20-nt window (ACGTN only): an N counts as a mismatch
    when the genome has an N mask, else it reads as the strand's N code."""
    mm = 0
    for w, b in zip(window, g):
        if w == "N":
            mm += 1 if n_as_mismatch else (n_code != b)
        else:
            mm += w != b
    return mm


def _pam_ok(pam_seq: str, pam: str, n_code: str) -> bool:
    for c, code in zip(pam_seq, pam.upper()):
        b = n_code if c == "N" else c
        if b not in IUPAC[code]:
            return False
    return True


def kernel_sites(contigs, guides: Sequence[Tuple[str, str]], max_mm: int,
                 n_mask: bool = True) -> set:
    """{(guide index, strand 0/1, scanned position, n_mm)} with kernel quirks."""
    fwd, rc = strand_strings(contigs)
    L = len(fwd)
    out = set()
    for s, S, n_code in ((0, fwd, "A"), (1, rc, "T")):
        for p in range(0, max(0, L - 26)):
            for gi, (seq, pam) in enumerate(guides):
                plen = len(pam)
                if not _pam_ok(S[p + 20:p + 20 + plen], pam, n_code):
                    continue
                mm = _count(S[p:p + 20], _guide20(seq), n_mask, n_code)
                if mm <= max_mm:
                    out.add((gi, s, p, mm))
    return out


def contig_sites(contigs, guides: Sequence[Tuple[str, str]], max_mm: int) -> set:
    """{(guide index, chrom, 1-based position, strand, n_mm)} inside contigs."""
    out = set()
    for name, seq in contigs:
        seq = seq.upper()
        rc = revcomp(seq)
        for gi, (g, pam) in enumerate(guides):
            win = 20 + len(pam)
            for i in range(len(seq) - win + 1):
                if _pam_ok(seq[i + 20:i + win], pam, "A"):
                    mm = _count(seq[i:i + 20], _guide20(g), True, "A")
                    if mm <= max_mm:
                        out.add((gi, name, i + 1, "+", mm))
                if _pam_ok(rc[i + 20:i + win], pam, "T"):
                    mm = _count(rc[i:i + 20], _guide20(g), True, "T")
                    if mm <= max_mm:
                        leftmost = len(seq) - i - win
                        out.add((gi, name, leftmost + win, "-", mm))
    return out


def random_contigs(rng, lengths: Sequence[int], n_frac: float = 0.0, gc: float = 0.41,
                   n_run: Tuple[int, int] = (5, 40)) -> List[Tuple[str, str]]:
    at = (1 - gc) / 2
    p = [at, gc / 2, gc / 2, at]
    contigs = []
    for k, L in enumerate(lengths):
        s = list(rng.choice(list("ACGT"), size=L, p=p))
        n_left = int(n_frac * L)
        while n_left > 0 and L > 50:
            ln = int(rng.integers(*n_run))
            st = int(rng.integers(0, max(1, L - ln)))
            s[st:st + ln] = ["N"] * min(ln, L - st)
            n_left -= ln
        contigs.append((f"c{k}", "".join(s)))
    return contigs


def plant(rng, contigs, guide: str, pam_seq: str, n_mm: int, strand: str):
    """Insert a near-copy of guide+pam at a random position of a random contig."""
    k = int(rng.integers(0, len(contigs)))
    name, s = contigs[k]
    win = 20 + len(pam_seq)
    pos = int(rng.integers(0, len(s) - win))
    site = list(_guide20(guide))
    for i in rng.choice(20, size=n_mm, replace=False):
        site[i] = "ACGT"[("ACGT".index(site[i]) + int(rng.integers(1, 4))) % 4]
    oriented = "".join(site) + pam_seq
    ins = oriented if strand == "+" else revcomp(oriented)
    contigs[k] = (name, s[:pos] + ins + s[pos + win:])
    return k, pos
