"""The GPU fast mode (beguidedoff_amd/fast.py) against the fast-mode oracles:
the table of oracle/fast_finder_oracle.py (20-nt guides, over the C
restatement of the anchored candidate search) and a literal brute force of
fast/finder.py:163-277 (20- and 21-nt guides), for both candidate engines (the
HBM anchor index and the budget scan).  Genomes carry N runs, sites planted at
both genome tails, across contig junctions and with N inside; one genome is
built of long N and poly-A runs and tandem copies of a site, so index buckets
hold thousands of positions and runs of one key."""
import numpy as np
import pandas as pd
import pytest

from beguidedoff_amd import beguidedoff as make_finder
from beguidedoff_amd import synth
from beguidedoff_amd.fast import FastIndexGPU, anchor_table, cold_region, mismatch_budget
from beguidedoff_amd.types import GuideInput
from brute import IUPAC, plant, random_contigs, revcomp
from oracle.fast_finder_oracle import FastFinderOracle

IUPAC_BITS = {"A": 1, "C": 2, "G": 4, "T": 8, "R": 5, "Y": 10, "S": 6, "W": 9, "K": 12, "M": 3,
              "B": 14, "D": 13, "H": 11, "V": 7, "N": 15}
CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


def test_regions_and_budget_known_answers():
    # reference tests/test_regions.py:7-14: cutoff 5 -> cold positions 9-20
    assert cold_region(5.0) == (8, 12) and cold_region(3.0) == (6, 14)
    # 1.2 + 1.44 + 1.728 = 4.368 <= 5 < 4.368 + 2.0736
    assert [mismatch_budget(t) for t in (5.0, 4.368, 4.36, 3.0, 1.2, 1.19)] == [3, 3, 2, 2, 1, 0]
    with pytest.raises(ValueError, match="cold"):
        cold_region(40.0)


def test_anchor_table_known_answers():
    """Anchor records decoded back to strings: keys are the anchors' first
    min(m, 12) bases (fast/finder.py:179-189), sites carry the anchor on
    positions c0..19, N codes mark A ('+') and T ('-') guide bases."""
    from beguidedoff_amd.finder import _ScanGuide

    def bases(p0, p1, n=20):
        return "".join("ACGT"[((p0 >> i) & 1) | (((p1 >> i) & 1) << 1)] for i in range(n))

    def key(s):
        return sum(CODE[c] << (2 * (11 - j)) for j, c in enumerate(s))

    g20, g21 = "ACGTTGCAACGTTGCAAGGC", "TTGACCAGTACGGATCCATGA"
    for cutoff, c0, m in ((5.0, 8, 12), (3.0, 6, 14), (7.0, 10, 10)):
        a = anchor_table([_ScanGuide(g20, "NGRRT"), _ScanGuide(g21.lower(), "NG")], cutoff)
        assert a["strand"].tolist() == [0, 1, 0, 1] and a["guide"].tolist() == [0, 0, 1, 1]
        for r, g, st in zip(a, (g20, g20, g21, g21), (0, 1, 0, 1)):
            anchor = g[c0:20] if st == 0 else revcomp(g)[:m]
            k = min(m, 12)
            assert (r["key_lo"], r["key_hi"]) == (key(anchor[:k]), key(anchor[:k]) + 4 ** (12 - k))
            site = bases(r["a0"], r["a1"])
            assert site[:c0] == g[:c0]
            assert site[c0:] == (g[c0:20] if st == 0 else g[len(g) - m:])
            assert bases(r["g0"], r["g1"]) == g[:20]
            n_sym = "A" if st == 0 else "T"
            assert r["n_code"] == sum(1 << i for i in range(20) if g[i] == n_sym)
            assert r["cold_first"] == c0
        assert a["pam_nibs"].tolist() == [0x8554F, 0x8554F, 0x4F, 0x4F]
        assert a["pam_len"].tolist() == [5, 5, 2, 2]
    with pytest.raises(ValueError, match="at least 20"):
        anchor_table([_ScanGuide("ACGT" * 4, "NGG")], 5.0)


def brute_candidates(contigs, guide, pam, cutoff):
    """fast/finder.py:163-277 literally: anchors on the A-coded forward
    genome, window and anchor inside one contig, IUPAC PAM, 2-bit mismatches
    of the first 20 bases plus the N positions, score <= cutoff."""
    full = "".join(s for _, s in contigs).upper()
    enc = "".join(c if c in "ACGT" else "A" for c in full)
    isn = [c == "N" for c in full]
    offs = np.cumsum([0] + [len(s) for _, s in contigs])[:-1]
    lens = [len(s) for _, s in contigs]
    c0, m = cold_region(cutoff)
    plen, window = len(pam), 20 + len(pam)
    out = set()
    for strand, anchor, boff in (("+", guide[c0:20], c0), ("-", revcomp(guide)[:m], 0)):
        start = enc.find(anchor)
        while start >= 0:
            gpos = start
            start = enc.find(anchor, start + 1)
            lm = gpos - boff if strand == "+" else gpos - plen
            k = int(np.searchsorted(offs, lm, side="right") - 1)
            off, ln = int(offs[k]), lens[k]
            if gpos < off or gpos + len(anchor) > off + ln or lm < off or lm + window > off + ln:
                continue
            w = enc[lm:lm + window]
            site, spam = (w[:20], w[20:]) if strand == "+" else (revcomp(w[plen:]), revcomp(w[:plen]))
            if not all((IUPAC_BITS[p] >> CODE[b]) & 1 for b, p in zip(spam, pam)):
                continue
            mm = {i + 1 for i in range(20) if guide[i] != site[i]}
            g0 = lm if strand == "+" else lm + plen
            nsym = "A" if strand == "+" else "T"
            for i in range(20):
                if (isn[g0 + i] if strand == "+" else isn[g0 + 19 - i]) and guide[i] == nsym:
                    mm.add(i + 1)
            pos = sorted(mm)
            score = sum(1.2 ** p for p in pos) if pos else 0.0
            if score <= cutoff:
                out.add((lm, strand, tuple(pos), score))
    return out


def put(contigs, k, pos, text):
    name, s = contigs[k]
    contigs[k] = (name, s[:pos] + text + s[pos + len(text):])


def near(rng, guide, n_hot, n_cold, c0=8):
    s = list(guide[:20])
    idx = list(rng.choice(c0, size=n_hot, replace=False)) + \
        list(c0 + rng.choice(20 - c0, size=n_cold, replace=False))
    for i in idx:
        s[i] = "ACGT"[("ACGT".index(s[i]) + int(rng.integers(1, 4))) % 4]
    return "".join(s)


def fast_case(seed, lengths=(20, 20, 21, 20, 21, 20)):
    rng = np.random.default_rng(seed)
    contigs = random_contigs(rng, [int(x) for x in rng.integers(3000, 7000, size=4)], n_frac=0.02)
    contigs = [(f"chr{i + 1}", s) for i, (_, s) in enumerate(contigs)]
    pams = ["NGG", "NG", "NNGRRT", "NGG", "NGA", "NAG"]
    guides = []
    for i, gl in enumerate(lengths):
        seq = "".join(rng.choice(list("ACGT"), size=gl))
        pam = pams[i % len(pams)]
        guides.append((seq, pam))
        for _ in range(10):
            site = near(rng, seq, int(rng.integers(0, 4)), int(rng.random() < 0.2))
            pseq = "".join(rng.choice(list(IUPAC[c])) for c in pam)
            plant(rng, contigs, site + seq[20:], pseq, 0, "+" if rng.random() < 0.5 else "-")
    # N inside planted sites, at hot and cold positions
    for i in range(4):
        seq, pam = guides[i]
        site = list(near(rng, seq, 1, 0))
        site[int(rng.integers(0, 20))] = "N"
        plant(rng, contigs, "".join(site), "".join(rng.choice(list(IUPAC[c])) for c in pam), 0, "+")
    # tails: '+' window ending the genome, '-' window starting the last contig
    seq, pam = guides[0]
    tail = seq[:20] + "AGG"
    put(contigs, len(contigs) - 1, len(contigs[-1][1]) - len(tail), tail)
    put(contigs, len(contigs) - 1, 0, revcomp(near(rng, seq, 1, 0) + "TGG"))
    # across the junction of chr1 and chr2
    j = near(rng, seq, 1, 0) + "CGG"
    put(contigs, 0, len(contigs[0][1]) - 10, j[:10])
    put(contigs, 1, 0, j[10:])
    return contigs, guides


def product_candidates(finder, contigs, guides, cutoff):
    """{guide index: {(leftmost, strand, positions, score)}} from the fast finder."""
    ext = finder._ext
    inputs = [GuideInput(f"g{i}", s, p, "3prime", 1, 20, "ABE") for i, (s, p) in enumerate(guides)]
    keys, _ = ext._deduplicate_guides(inputs)
    h = finder._candidates(keys, cutoff)
    offs, lens = ext.chrom_offsets, ext.chrom_lengths
    out = {i: set() for i in range(len(keys))}
    for kk, st, pos, mm, sc in zip(h.key.tolist(), h.strand.tolist(), h.pos.tolist(),
                                   h.mm_out.tolist(), h.score.tolist()):
        window = 20 + len(keys[kk].pam)
        if st == 0:
            lm = pos
        else:
            c = int(np.searchsorted(offs, pos, side="right") - 1)
            lm = int(offs[c] + lens[c] - (pos - offs[c]) - window)
        pl = tuple(i + 1 for i in range(20) if (mm >> i) & 1)
        out[kk].add((lm, "+-"[st], pl, sc))
    return [out[i] for i in range(len(keys))]


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ["index", "scan"])
@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("cutoff", [5.0, 3.0, 7.0])
def test_candidates_equal_literal_bruteforce(gpu, seed, cutoff, engine):
    """Cutoffs 5 and 3 have a cold block of 12 and 14 bases (one index key,
    then a key plus two checked bases), 7 one of 10 (a range of 16 keys)."""
    contigs, guides = fast_case(seed)
    genome = synth.genome_from_strings(contigs)
    f = make_finder(mode="fast", genome=genome, search_cutoff=cutoff, engine=engine)
    got = product_candidates(f, contigs, guides, cutoff)
    total = 0
    for (seq, pam), cand in zip(guides, got):
        want = brute_candidates(contigs, seq, pam, cutoff)
        assert cand == want, (seq, pam)
        total += len(want)
    assert total >= 10


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ["index", "scan"])
def test_21nt_minus_anchor_is_shifted(gpu, engine):
    """For 21-nt guides the reference's '-' anchor is rc(guide)[:12], the rc
    of bases 10-21: a '-' site exists only where those equal bases 9-20 (a
    run), and then the brute force and the product agree."""
    rng = np.random.default_rng(8)
    contigs = random_contigs(rng, [4000, 3000], n_frac=0.0)
    contigs = [(f"chr{i + 1}", s) for i, (_, s) in enumerate(contigs)]
    run = "".join(rng.choice(list("ACGT"), size=8)) + "C" * 13        # bases 9-21 all C
    plain = "".join(rng.choice(list("ACGT"), size=21))
    for seq in (run, plain):
        for strand in "+-":
            for _ in range(3):
                plant(rng, contigs, near(rng, seq, 1, 0), "AGG", 0, strand)
    genome = synth.genome_from_strings(contigs)
    f = make_finder(mode="fast", genome=genome, engine=engine)
    guides = [(run, "NGG"), (plain, "NGG")]
    got = product_candidates(f, contigs, guides, 5.0)
    want = [brute_candidates(contigs, s, p, 5.0) for s, p in guides]
    assert got == want
    assert any(st == "-" for _, st, _, _ in want[0])
    assert not any(st == "-" for _, st, _, _ in want[1])


def table_guides(rng, guides, contigs):
    out = []
    for i, (seq, pam) in enumerate(guides):
        if len(seq) != 20:
            continue
        et, ws, we = [("ABE", 4, 8), ("CBE", 3, 9), ("ABE", 1, 20)][i % 3]
        out.append(GuideInput(f"g{i}", seq, pam, "3prime", ws, we, et,
                              base_editor=["ABE8e", None][i % 2]))
    # an input sharing its id and sequence with another editor, and one with an on-target
    out.append(GuideInput(out[0].guide_id, out[0].guide_seq, out[0].pam, "3prime", 1, 20, "CBE"))
    out.append(GuideInput("ot", out[1].guide_seq, out[1].pam, "3prime", 1, 20, "ABE",
                          on_target_chrom="chr3", on_target_pos=101, on_target_strand="+"))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("return_all,thr,cutoff", [(False, None, 5.0), (True, 5.0, 5.0),
                                                   (False, 4.0, 3.0)])
def test_table_equals_oracle(gpu, seed, return_all, thr, cutoff):
    rng = np.random.default_rng(50 + seed)
    contigs, guides = fast_case(seed, lengths=(20, 20, 20, 20, 20, 20))
    inputs = table_guides(rng, guides, contigs)
    # make the on-target input's site real: plant it at chr3:101 '+'
    seq, pam = inputs[-1].guide_seq, inputs[-1].pam
    put(contigs, 2, 100, seq + "".join(rng.choice(list(IUPAC[c])) for c in pam))
    genome = synth.genome_from_strings(contigs)
    f = make_finder(mode="fast", genome=genome, search_cutoff=cutoff, n_top_offtargets=4)
    got = f.search(guides=inputs, return_all=return_all, count_score_threshold=thr, verbose=False)
    want = FastFinderOracle(genome, search_cutoff=cutoff, n_top_offtargets=4).search(
        inputs, return_all=return_all, count_score_threshold=thr)
    assert len(want) > 5
    pd.testing.assert_frame_equal(got, want)


def repeat_case(seed):
    """Contigs of random sequence, a 60 kb N run, a 4 kb poly-A and poly-T run
    and 400 tandem copies of a near-site (bucket sizes past a block's 256
    threads, many positions of one key in a wave)."""
    rng = np.random.default_rng(seed)
    unit_guide = "".join(rng.choice(list("ACGT"), size=20))
    unit = near(rng, unit_guide, 1, 0) + "TGG"
    contigs = random_contigs(rng, [5000, 4000, 3000], n_frac=0.0)
    contigs = [("chr1", contigs[0][1] + "N" * 60000 + contigs[1][1]),
               ("chr2", "A" * 4000 + unit * 400 + "T" * 4000),
               ("chr3", contigs[2][1])]
    guides = [(unit_guide, "NGG"), ("A" * 20, "NGG"), ("T" * 20, "NGG"), ("A" * 19 + "G", "NNN"),
              ("T" * 19 + "C", "NN"), (unit_guide[::-1], "NG")]
    for seq, pam in guides[3:]:
        for strand in "+-":
            plant(rng, contigs, near(rng, seq, 1, 0), "".join(rng.choice(list(IUPAC[c])) for c in pam),
                  0, strand)
    return contigs, guides


@pytest.mark.gpu
@pytest.mark.parametrize("cutoff", [5.0, 7.0])
def test_index_on_repeats_and_n_runs(gpu, cutoff):
    """Every occurrence kept (max_hits_per_anchor=0): the A-rich anchors have
    ~64,000 occurrences here, over the reference's default cap."""
    contigs, guides = repeat_case(3)
    genome = synth.genome_from_strings(contigs)
    f = make_finder(mode="fast", genome=genome, search_cutoff=cutoff, max_hits_per_anchor=0)
    got = product_candidates(f, contigs, guides, cutoff)
    want = [brute_candidates(contigs, s, p, cutoff) for s, p in guides]
    assert got == want
    assert len(want[0]) == 400 and sum(len(w) for w in want) > 400


@pytest.mark.gpu
def test_index_contents_equal_a_cpu_counting_sort(gpu):
    """The bucket table equals a CPU bincount of the 12-mer keys (base p in
    the top bits, N and bases past the end as A) and the position array is a
    permutation of all positions grouped in key order.  The genome spans four
    build tiles of 2^20 positions, and its all-A key (a 1.3 Mbp N run) fills
    two tiles of its high-digit bucket."""
    from beguidedoff_amd.batched import DeviceGenome
    rng = np.random.default_rng(11)

    def rnd(n):
        return "".join(rng.choice(list("ACGT"), size=n))
    contigs = [("chr1", rnd(1_200_000) + "N" * 1_300_000 + rnd(400_000)),
               ("chr2", "T" * 50_000 + rnd(900_001)),
               ("chr3", rnd(777))]
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    lens = np.array([len(s) for _, s in contigs], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    idx = FastIndexGPU(dg, offs, lens)
    try:
        lut = np.zeros(256, dtype=np.uint32)
        for b, c in CODE.items():
            lut[ord(b)] = c
        codes = lut[np.frombuffer("".join(s for _, s in contigs).encode(), dtype=np.uint8)]
        n = codes.size
        padded = np.concatenate([codes, np.zeros(11, dtype=np.uint32)])
        keys = np.zeros(n, dtype=np.uint32)
        for i in range(12):
            keys |= padded[i:i + n] << np.uint32(22 - 2 * i)
        want = np.zeros((1 << 24) + 1, dtype=np.int64)
        want[1:] = np.cumsum(np.bincount(keys.astype(np.int64), minlength=1 << 24))
        assert want[4096] > 1 << 20  # high digit 0 (the N run): two pass-2 tiles
        assert np.array_equal(idx.offsets().astype(np.int64), want)
        pos = idx.positions(0, n)
        assert np.array_equal(np.sort(pos), np.arange(n, dtype=np.uint32))
        assert np.all(np.diff(keys[pos].astype(np.int64)) >= 0)
    finally:
        idx.close()
        dg.close()


@pytest.mark.gpu
def test_index_query_grows_its_buffer_and_sizes(gpu):
    contigs, guides = repeat_case(4)
    genome = synth.genome_from_strings(contigs)
    f = make_finder(mode="fast", genome=genome)
    ext = f._ext
    keys, _ = ext._deduplicate_guides(
        [GuideInput(f"g{i}", s, p, "3prime", 1, 20, "ABE") for i, (s, p) in enumerate(guides)])
    anchors = anchor_table(keys, 5.0)
    big, _ = f._index.query(anchors, 5.0)
    small, ms = f._index.query(anchors, 5.0, capacity=1)
    assert len(big) >= 400 and ms >= 0.0

    def rows(h):
        return sorted(zip(h["guide"].tolist(), h["strand"].tolist(), h["pos"].tolist(),
                          h["mm_mask"].tolist(), h["score"].tolist()))
    assert rows(big) == rows(small)
    n = sum(len(s) for _, s in contigs)
    assert f._index.nbytes == 4 * ((1 << 24) + 1) + 4 * n + 16 * len(contigs)
    assert f._index.build_s > 0.0
    # a second index over the same resident genome answers the same
    other = FastIndexGPU(ext._device_genome(), ext.chrom_offsets, ext.chrom_lengths)
    assert rows(other.query(anchors, 5.0)[0]) == rows(big)
    other.close()
    f.close()


def cap_case(seed):
    """A small genome with anchors far over a small cap: N blocks (indexed as
    A; one ends the last contig, so its suffixes meet the terminal 0), poly-A
    and poly-T runs, 120 tandem copies of a near-site, random sequence."""
    rng = np.random.default_rng(seed)
    unit_guide = "".join(rng.choice(list("ACGT"), size=20))
    unit = near(rng, unit_guide, 1, 0) + "TGG"
    r = random_contigs(rng, [2000, 1500, 1000, 800, 1500], n_frac=0.0)
    contigs = [("chr1", r[0][1] + "N" * 600 + r[1][1] + "A" * 300 + r[2][1]),
               ("chr2", "A" * 150 + unit * 120 + "T" * 200 + r[3][1]),
               ("chr3", r[4][1] + "N" * 250)]
    guides = [(unit_guide, "NGG"), ("A" * 20, "NNN"), ("T" * 20, "NNN"),
              ("ACGTAC" + "A" * 14, "NGG"), ("GATTACA" + "A" * 13, "NG")]
    return contigs, guides


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ["index", "scan"])
@pytest.mark.parametrize("cap", [1, 7, 40, 0])
def test_anchor_cap_keeps_the_first_in_suffix_array_order(gpu, engine, cap, caplog):
    """The reference keeps the first max_hits_per_anchor occurrences of an
    anchor's suffix-array interval (csrc/fmindex_binding.cpp:164-176) and
    warns (fast/finder.py:368-375): the candidates equal the literal brute
    force restricted to the occurrences a brute-force suffix sort of the
    reference's index text keeps (oracle/sa_order.py)."""
    import logging

    from oracle.sa_order import capped_occurrences

    contigs, guides = cap_case(11)
    genome = synth.genome_from_strings(contigs)
    f = make_finder(mode="fast", genome=genome, engine=engine, max_hits_per_anchor=cap)
    with caplog.at_level(logging.WARNING):
        got = product_candidates(f, contigs, guides, 5.0)
    c0, m = cold_region(5.0)
    over = set()
    for i, ((seq, pam), cand) in enumerate(zip(guides, got)):
        want = brute_candidates(contigs, seq, pam, 5.0)
        for strand, anchor in (("+", seq[c0:20]), ("-", revcomp(seq)[:m])):
            kept, total = capped_occurrences(contigs, anchor, cap)
            if cap > 0 and total > cap:
                over.add(seq)
                boff = c0 if strand == "+" else len(pam)
                want = {w for w in want if w[1] != strand or w[0] + boff in set(kept)}
        assert cand == want, (seq, pam, cap)
    msgs = [r.getMessage() for r in caplog.records if "Anchor hit cap" in r.getMessage()]
    assert {s for s in over if any(s in x for x in msgs)} == over
    assert len(msgs) == len(over)
    if cap in (1, 7):
        assert len(over) >= 4  # the tandem unit and the A-rich anchors
    f.close()
