"""Pin the C scan oracle (oracle/bgo_oracle.c) against an independent
string brute force, and the reference's quirks F3-F6 against per-contig
semantics.  CPU only."""
import numpy as np
import pytest

from beguidedoff_amd import synth
from beguidedoff_amd.encoding import encode_dna, encode_guide_64bit
from beguidedoff_amd.pam import encode_pam_mask
from brute import contig_sites, kernel_sites, plant, random_contigs, revcomp
from oracle.scan import scan_strand

PAMS = ["NGG", "NRG", "NNGRRT", "NNNRRT", "NG", "TTTV", "SGG", "WGG", "A", "NNNNNN", "BDHV"]


def oracle_sites(genome, guides, max_mm, n_mask=True, threads=4):
    vals = np.array([encode_guide_64bit(encode_dna(s)) & ((1 << 64) - 1) for s, _ in guides],
                    dtype=np.uint64)
    masks = np.stack([encode_pam_mask(p) for _, p in guides])
    lens = np.array([len(p) for _, p in guides], np.uint8)
    out = set()
    for s, key, nm, enc in ((0, "genome_forward", "n_mask_le_fwd", 0),
                            (1, "genome_rc", "n_mask_le_rc", 3)):
        m = genome.get(nm) if n_mask else None
        g, p, mm = scan_strand(genome[key], genome["total_bp"], vals, masks, lens,
                               np.ones(len(guides), np.uint8), max_mm, m, enc, threads=threads)
        out |= set(zip(g.tolist(), [s] * g.size, p.tolist(), mm.tolist()))
    return out


def build_case(seed, n_contigs=3, length=(200, 700), n_frac=0.04, n_guides=4, max_mm=4):
    rng = np.random.default_rng(seed)
    lengths = [int(rng.integers(*length)) for _ in range(n_contigs)]
    contigs = random_contigs(rng, lengths, n_frac=n_frac)
    pams = [PAMS[int(i)] for i in rng.choice(len(PAMS), size=n_guides)]
    seqs = ["".join(rng.choice(list("ACGT"), size=int(rng.choice([20, 21]))))
            for _ in range(n_guides)]
    guides = list(zip(seqs, pams))
    for gs, pam in guides:
        for _ in range(3):
            pam_seq = "".join(rng.choice(list(__import__("brute").IUPAC[c])) for c in pam)
            plant(rng, contigs, gs, pam_seq, int(rng.integers(0, max_mm + 1)),
                  "+" if rng.random() < 0.5 else "-")
    # sites straddling every junction and the genome tail
    for k in range(n_contigs - 1):
        name, s = contigs[k]
        gs, pam = guides[k % n_guides]
        pam_seq = "".join(__import__("brute").IUPAC[c][0] for c in pam)
        site = gs[:20] + pam_seq
        cut = int(rng.integers(3, len(site) - 3))
        contigs[k] = (name, s[:-cut] + site[:cut])
        nm2, s2 = contigs[k + 1]
        contigs[k + 1] = (nm2, site[cut:] + s2[len(site) - cut:])
    name, s = contigs[-1]
    gs, pam = guides[0]
    tail = gs[:20] + "".join(__import__("brute").IUPAC[c][0] for c in pam)
    contigs[-1] = (name, s[:-len(tail) - 2] + tail + "AC")
    return contigs, guides, max_mm


@pytest.mark.parametrize("seed", range(12))
def test_oracle_equals_kernel_semantics_bruteforce(seed):
    contigs, guides, max_mm = build_case(seed)
    genome = synth.genome_from_strings(contigs)
    got = oracle_sites(genome, guides, max_mm)
    want = kernel_sites(contigs, guides, max_mm, n_mask="N" in "".join(s for _, s in contigs))
    assert got == want


@pytest.mark.parametrize("seed", range(6))
def test_quirks_are_only_junction_and_tail_windows(seed):
    """Sites of windows inside one contig == per-contig brute force minus the
    windows in the global tail skip (F5: junction windows are extra; F6:
    positions >= total_bp - 26 are never scanned)."""
    contigs, guides, max_mm = build_case(100 + seed)
    genome = synth.genome_from_strings(contigs)
    offs = genome["chrom_offsets"]
    lens = genome["chrom_lengths"]
    names = [str(n) for n in genome["chrom_names"]]
    L = int(genome["total_bp"])
    inside, junction = set(), 0
    for gi, s, p, mm in oracle_sites(genome, guides, max_mm):
        win = 20 + len(guides[gi][1])
        k = int(np.searchsorted(offs, p, side="right") - 1)
        if p + win > offs[k] + lens[k]:
            junction += 1
            continue
        local = p - offs[k]
        if s == 0:
            inside.add((gi, names[k], local + 1, "+", mm))
        else:
            inside.add((gi, names[k], int(lens[k] - local - win) + win, "-", mm))
    ref = contig_sites(contigs, guides, max_mm)
    tail = set()
    for gi, name, pos, strand, mm in ref:
        k = names.index(name)
        win = 20 + len(guides[gi][1])
        scanned = (offs[k] + pos - 1) if strand == "+" else (offs[k] + lens[k] - pos)
        if scanned >= L - 26:
            tail.add((gi, name, pos, strand, mm))
    assert inside == ref - tail
    assert junction > 0 or seed % 2  # the planted junction sites are usually hits


def test_oracle_without_n_mask_reads_n_as_code():
    contigs, guides, max_mm = build_case(7, n_frac=0.1)
    genome = synth.genome_from_strings(contigs)
    got = oracle_sites(genome, guides, max_mm, n_mask=False)
    assert got == kernel_sites(contigs, guides, max_mm, n_mask=False)


def test_oracle_threads_and_ranges_are_consistent():
    contigs, guides, max_mm = build_case(3, length=(3000, 5000), n_guides=6)
    genome = synth.genome_from_strings(contigs)
    vals = np.array([encode_guide_64bit(encode_dna(s)) & ((1 << 64) - 1) for s, _ in guides],
                    dtype=np.uint64)
    masks = np.stack([encode_pam_mask(p) for _, p in guides])
    lens = np.array([len(p) for _, p in guides], np.uint8)
    args = (genome["genome_forward"], genome["total_bp"], vals, masks, lens,
            np.ones(len(guides), np.uint8), 5, genome["n_mask_le_fwd"], 0)
    one = scan_strand(*args, threads=1)
    many = scan_strand(*args, threads=8)
    for a, b in zip(one, many):
        assert np.array_equal(a, b)
    parts = [scan_strand(*args, pos_begin=b, pos_end=b + 1777, threads=2)
             for b in range(0, int(genome["total_bp"]), 1777)]
    for i in range(3):
        assert np.array_equal(np.concatenate([p[i] for p in parts]), one[i])


def test_oracle_five_prime_pam():
    """pam_3prime = 0: PAM at pos, protospacer at pos + plen
    (cuda_search.py:254-260)."""
    rng = np.random.default_rng(5)
    contigs = random_contigs(rng, [900])
    g = "".join(rng.choice(list("ACGT"), size=20))
    s = contigs[0][1]
    contigs = [("c0", s[:300] + "TTTA" + g + s[324:])]
    genome = synth.genome_from_strings(contigs)
    vals = np.array([encode_guide_64bit(encode_dna(g))], np.uint64)
    masks = encode_pam_mask("TTTV")[None, :]
    gi, pos, mm = scan_strand(genome["genome_forward"], 900, vals, masks, np.array([4], np.uint8),
                              np.zeros(1, np.uint8), 3, None, 0)
    S = contigs[0][1]
    want = []
    for p in range(900 - 26):
        if p + 4 + 20 > 900:
            continue
        if S[p:p + 3] == "TTT" and S[p + 3] in "ACG":
            m = sum(a != b for a, b in zip(S[p + 4:p + 24], g))
            if m <= 3:
                want.append((p, m))
    assert list(zip(pos.tolist(), mm.tolist())) == want and (300, 0) in want
