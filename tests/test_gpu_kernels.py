"""GPU parity: the plain drop-in kernel and the batched engine against the
scan oracle, bit-exact as sorted sets."""
import numpy as np
import pytest

from beguidedoff_amd import synth
from beguidedoff_amd.batched import DeviceGenome, GuideSpec, Plan, guide_descs, scan
from beguidedoff_amd.cuda_search import MAX_GUIDES, search_genome_cuda
from beguidedoff_amd.encoding import encode_dna, encode_guide_64bit
from beguidedoff_amd.finder import cctop_scores
from beguidedoff_amd.pam import encode_pam_mask
from oracle.scan import scan_strand
from test_oracle_pins import build_case

pytestmark = pytest.mark.gpu

STRANDS = ((0, "genome_forward", "n_mask_le_fwd", 0), (1, "genome_rc", "n_mask_le_rc", 3))


def arrays(guides):
    vals = np.array([encode_guide_64bit(encode_dna(s)) & ((1 << 64) - 1) for s, _ in guides],
                    dtype=np.uint64)
    masks = np.stack([encode_pam_mask(p) for _, p in guides])
    lens = np.array([len(p) for _, p in guides], np.uint8)
    return vals, masks, lens


def oracle_set(genome, guides, max_mm, three=None, strands=(0, 1), pos=(0, None), mm_per=None):
    vals, masks, lens = arrays(guides)
    tp = np.ones(len(guides), np.uint8) if three is None else three
    out = set()
    for s, key, nm, enc in STRANDS:
        if s not in strands or key not in genome:
            continue
        if mm_per is None:
            g, p, m = scan_strand(genome[key], genome["total_bp"], vals, masks, lens, tp, max_mm,
                                  genome.get(nm), enc, pos_begin=pos[0], pos_end=pos[1])
            out |= set(zip(g.tolist(), [s] * g.size, p.tolist(), m.tolist()))
        else:
            for gi, k in enumerate(mm_per):
                g, p, m = scan_strand(genome[key], genome["total_bp"], vals[gi:gi + 1],
                                      masks[gi:gi + 1], lens[gi:gi + 1], tp[gi:gi + 1], k,
                                      genome.get(nm), enc, pos_begin=pos[0], pos_end=pos[1])
                out |= {(gi, s, pp, mm) for pp, mm in zip(p.tolist(), m.tolist())}
    return out


def plain_set(genome, guides, max_mm, three=None, chunk=200_000_000, rpg=10240):
    vals, masks, lens = arrays(guides)
    tp = np.ones(len(guides), np.uint8) if three is None else three
    out = set()
    for s, key, nm, enc in STRANDS:
        res = search_genome_cuda(genome[key], genome["total_bp"], vals, masks, lens, tp, max_mm,
                                 chunk_size=chunk, results_per_guide=rpg,
                                 n_mask_packed=genome.get(nm), n_encoded=enc)
        out |= {(g, s, p, m) for g, hits in res.items() for p, m in hits}
    return out


def hit_set(hits):
    return set(zip(hits["guide"].tolist(), hits["strand"].tolist(), hits["pos"].tolist(),
                   hits["n_mm"].tolist()))


@pytest.mark.parametrize("seed", range(8))
def test_plain_matches_oracle_small(gpu, seed):
    contigs, guides, max_mm = build_case(seed)
    genome = synth.genome_from_strings(contigs)
    assert plain_set(genome, guides, max_mm) == oracle_set(genome, guides, max_mm)


def test_plain_chunks_five_prime_and_no_mask(gpu):
    contigs, guides, max_mm = build_case(21, length=(3000, 6000), n_guides=8)
    genome = synth.genome_from_strings(contigs)
    three = np.array([1, 0] * 4, np.uint8)
    want = oracle_set(genome, guides, max_mm, three=three)
    assert plain_set(genome, guides, max_mm, three=three, chunk=1000) == want
    nomask = {k: v for k, v in genome.items() if not k.startswith("n_mask")}
    assert plain_set(nomask, guides, max_mm) == oracle_set(nomask, guides, max_mm)


def test_plain_result_buffer_overflow(gpu):
    contigs = [("chr1", ("T" + "ACGTACGTACGTACGTACGT"[1:] + "AGG") * 500 + "A" * 30)]
    genome = synth.genome_from_strings(contigs)
    vals, masks, lens = arrays([("ACGTACGTACGTACGTACGT", "NGG")])
    full, ovf, cnt = search_genome_cuda(genome["genome_forward"], genome["total_bp"], vals, masks,
                                        lens, np.ones(1, np.uint8), 4, return_overflow=True,
                                        results_per_guide=10240)
    assert not ovf[0] and cnt[0] == len(full[0]) == 500
    part, ovf, cnt = search_genome_cuda(genome["genome_forward"], genome["total_bp"], vals, masks,
                                        lens, np.ones(1, np.uint8), 4, return_overflow=True,
                                        results_per_guide=100)
    assert ovf[0] and cnt[0] == 500 and len(part[0]) == 100
    assert set(part[0]) <= set(full[0])
    with pytest.raises(RuntimeError, match="Too many guides"):
        search_genome_cuda(genome["genome_forward"], genome["total_bp"],
                           np.zeros(MAX_GUIDES + 1, np.uint64), np.zeros((MAX_GUIDES + 1, 6), np.uint8),
                           np.full(MAX_GUIDES + 1, 3, np.uint8), np.ones(MAX_GUIDES + 1, np.uint8), 4)


def test_plain_small_config(gpu):
    w = synth.workload("small")
    seqs = [g.guide_seq for g in w.guides]
    guides = [(s, "NGG") for s in seqs]
    assert plain_set(w.genome, guides, 3) == oracle_set(w.genome, guides, 3)


# Compare engines as tunings: 1 VALU xor/popcount (by default families of up
# to 8 guides of dense PAMs and 4 of sparse ones compare in place; direct_max
# -1 sends every family through the window buffer, 1000 is capped at 8),
# 2 int8 MFMA, 3 fp4 MFMA.
ENGINE_TUNINGS = {"valu": {"engine": 1}, "valu_buffered": {"engine": 1, "direct_max": -1},
                  "valu_direct": {"engine": 1, "direct_max": 1000}, "i8": {"engine": 2},
                  "fp4": {"engine": 3}}
ENGINES = [pytest.param(t, id=k) for k, t in ENGINE_TUNINGS.items()]


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("seed", range(8))
def test_batched_matches_oracle_small(gpu, seed, engine):
    contigs, guides, max_mm = build_case(seed)
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"]) if "N" in "".join(
        s for _, s in contigs) else DeviceGenome(genome)
    nm = genome if dg.has_n_mask else {k: v for k, v in genome.items() if not k.startswith("n_mask")}
    r = scan(dg, [GuideSpec(s, p, max_mm) for s, p in guides], tuning=dict(engine))
    assert hit_set(r.hits) == oracle_set(nm, guides, max_mm)


@pytest.mark.parametrize("engine", ENGINES)
def test_batched_mixed_budgets_many_families_ranges(gpu, engine):
    """>8 PAM families (two launch groups), per-guide budgets, 20/21-nt
    guides, position ranges and single-strand launches."""
    rng = np.random.default_rng(3)
    contigs, guides, _ = build_case(77, length=(20_000, 40_000), n_guides=12)
    pams = ["NGG", "NRG", "NNGRRT", "NNNRRT", "NG", "TTTV", "SGG", "WGG", "A", "NNNNNN", "BDHV", "GG"]
    guides = [(s, pams[i]) for i, (s, _) in enumerate(guides)]
    budgets = [int(rng.integers(0, 6)) for _ in guides]
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    specs = [GuideSpec(s, p, k) for (s, p), k in zip(guides, budgets)]
    r = scan(dg, specs, tuning=dict(engine))
    assert len(r.families) == 12
    assert hit_set(r.hits) == oracle_set(genome, guides, None, mm_per=budgets)
    L = genome["total_bp"]
    for (b, e), strands in [((1000, 17_000), 3), ((0, L // 2), 1), ((L // 3, L), 2)]:
        r = scan(dg, specs, pos_begin=b, pos_end=e, strands=strands, tuning=dict(engine))
        want = oracle_set(genome, guides, None, mm_per=budgets, pos=(b, e),
                          strands=[s for s in (0, 1) if strands & (1 << s)])
        assert hit_set(r.hits) == want


@pytest.mark.parametrize("engine", ENGINES)
def test_batched_capacity_rerun_and_fused_epilogue(gpu, engine):
    rng = np.random.default_rng(9)
    contigs, guides, max_mm = build_case(31, length=(50_000, 80_000), n_guides=6, max_mm=5)
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    specs = [GuideSpec(s, p, 6, ["ABE", "CBE"][i % 2], (int(rng.integers(1, 6)), int(rng.integers(6, 21))))
             for i, (s, p) in enumerate(guides)]
    r = scan(dg, specs, hit_capacity=3, tuning=dict(engine))
    assert r.launches == 2 and r.n_hits == r.hits.size > 3
    assert hit_set(r.hits) == oracle_set(genome, guides, 6)
    h = r.hits
    for s, key in ((0, "genome_forward"), (1, "genome_rc")):
        sel = h["strand"] == s
        pos = h["pos"][sel].astype(np.int64)
        codes = np.stack([synth.extract_codes(genome[key], int(p), 26) for p in pos]) if pos.size else np.zeros((0, 26), np.uint8)
        b0 = (codes & 1).astype(np.uint32) << np.arange(26, dtype=np.uint32)
        b1 = (codes >> 1).astype(np.uint32) << np.arange(26, dtype=np.uint32)
        assert np.array_equal(h["x0"][sel], b0.sum(1).astype(np.uint32))
        assert np.array_equal(h["x1"][sel], b1.sum(1).astype(np.uint32))
        gcodes = np.stack([np.pad(encode_dna(specs[g].seq)[:20], (0, 0)) for g in h["guide"][sel]]) if pos.size else np.zeros((0, 20), np.uint8)
        mm = ((codes[:, :20] != gcodes).astype(np.uint32) << np.arange(20, dtype=np.uint32)).sum(1)
        assert np.array_equal(h["mm_mask"][sel], mm.astype(np.uint32))
        assert np.array_equal(h["score"][sel], cctop_scores(mm.astype(np.uint32)))
        from beguidedoff_amd.base_editing import editable_base_in_window
        for row, c in zip(h[sel], codes):
            sp = specs[row["guide"]]
            seq = "".join("ACGT"[x] for x in c[:20])
            assert bool(row["editable"]) == bool(editable_base_in_window(seq, sp.editor_type, *sp.window))


@pytest.mark.parametrize("engine", ENGINES)
def test_batched_family_window_counts(gpu, engine):
    contigs, guides, max_mm = build_case(5, length=(5000, 9000))
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    pams = ["NGG", "NNNRRT", "NG", "TTTV"]
    specs = [GuideSpec(guides[0][0], p, 3) for p in pams]
    r = scan(dg, specs, tuning=dict(engine))
    L = genome["total_bp"]
    from beguidedoff_amd.pam import IUPAC_MASK
    for (n_g, windows), pam in zip(r.families, pams):
        want = 0
        for key in ("genome_forward", "genome_rc"):
            c = synth.extract_codes(genome[key], 0, L)
            ok = np.ones(L - 26, bool)
            for i, ch in enumerate(pam):
                ok &= ((IUPAC_MASK[ch] >> c[20 + i:20 + i + L - 26]) & 1).astype(bool)
            want += int(ok.sum())
        assert n_g == 1 and windows == want


@pytest.mark.parametrize("engine", ENGINES)
def test_graph_replays_match_eager(gpu, engine):
    contigs, guides, max_mm = build_case(8, length=(30_000, 50_000), n_guides=8)
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    specs = [GuideSpec(s, p, max_mm) for s, p in guides]
    plan = Plan(dg, specs, hit_capacity=1 << 16, tuning=dict(engine))
    plan.launch()
    n, _ = plan.wait()
    eager = hit_set(plan.fetch(n))
    for _ in range(4):
        plan.graph_launch()
        n2, _ = plan.wait()
        assert n2 == n and hit_set(plan.fetch(n2)) == eager
        plan.launch(0, genome["total_bp"] // 2, 1)   # eager work between replays
        plan.wait()
    plan.close()
    assert eager == oracle_set(genome, guides, max_mm)


def test_tunings_give_identical_sites(gpu):
    """Tile length, waves per block, occupancy and strand interleave change
    only the schedule, never the sites."""
    contigs, guides, max_mm = build_case(12, length=(60_000, 90_000), n_guides=10)
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    specs = [GuideSpec(s, p, max_mm) for s, p in guides]
    want = oracle_set(genome, guides, max_mm)
    for t in ({"chunk_positions": 64}, {"chunk_positions": 4096, "interleave": 1},
              {"chunk_positions": 1 << 20, "waves_per_block": 1, "blocks_per_cu": 1},
              {"waves_per_block": 2, "blocks_per_cu": 3, "interleave": 1},
              {"engine": 2, "chunk_positions": 64}, {"engine": 2, "waves_per_block": 1},
              {"engine": 2, "chunk_positions": 1 << 20, "interleave": 1},
              {"engine": 3, "chunk_positions": 64}, {"engine": 3, "waves_per_block": 1},
              {"engine": 3, "chunk_positions": 1 << 20, "interleave": 1}):
        r = scan(dg, specs, tuning=t)
        assert hit_set(r.hits) == want, t
    from beguidedoff_amd.autotune import autotune
    import os
    os.environ["BGO_TUNE_CACHE"] = "/tmp/bgo_tune_test.json"
    best, trials = autotune(dg, specs, sample_positions=100_000, use_cache=False,
                            space={"chunk_positions": [4096, 65536], "waves_per_block": [4],
                                   "blocks_per_cu": [0], "interleave": [0, 1]})
    assert len(trials) == 4 and hit_set(scan(dg, specs, tuning=best).hits) == want


def test_mfma_engine_many_guides_high_budget(gpu):
    """Hundreds of guides per family (many 32-guide B tiles), budgets up to 8
    so candidates are frequent, N runs and planted sites on both strands."""
    rng = np.random.default_rng(17)
    contigs, guides, _ = build_case(99, length=(30_000, 50_000), n_guides=4)
    seqs = ["".join(rng.choice(list("ACGT"), size=20)) for _ in range(150)]
    guides = [(q, ["NGG", "NG", "NNNRRT"][i % 3]) for i, q in enumerate(seqs)] + guides
    budgets = [int(rng.integers(0, 9)) for _ in guides]
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    specs = [GuideSpec(s, p, k) for (s, p), k in zip(guides, budgets)]
    want = oracle_set(genome, guides, None, mm_per=budgets)
    for engine in ENGINE_TUNINGS.values():
        assert hit_set(scan(dg, specs, tuning=dict(engine)).hits) == want


@pytest.mark.parametrize("engine", ENGINES)
def test_budget_and_composition_extremes(gpu, engine):
    """Budgets 0..20 against guides from all-A to no-A bases: the MFMA sign
    tests see every bias from -20 to +20 (budget - non-A guide bases), and a
    budget of 20 makes every PAM-passing window a site."""
    rng = np.random.default_rng(5)
    contigs, _, _ = build_case(5, length=(2_000, 4_000), n_guides=1)
    seqs = ["A" * 20, "C" * 20, "T" * 20, "A" * 10 + "G" * 10, "AAAAAAAAAAAAAAAAAAAC",
            "".join(rng.choice(list("ACGT"), size=20))]
    guides, budgets = [], []
    for q in seqs:
        for k in (0, 1, 4, 10, 19, 20):
            guides.append((q, "NG"))
            budgets.append(k)
    guides.append((contigs[0][1][100:120].replace("N", "A"), "NNN"))
    budgets.append(0)
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    specs = [GuideSpec(s, p, k) for (s, p), k in zip(guides, budgets)]
    want = oracle_set(genome, guides, None, mm_per=budgets)
    assert len(want) > 1000
    assert hit_set(scan(dg, specs, tuning=dict(engine)).hits) == want


@pytest.mark.parametrize("engine", ENGINES)
def test_odd_guide_tile_counts_with_planted_sites(gpu, engine):
    """Families of 1, 2, 3, 4 and 5 guide tiles (20, 40, 70, 100 and 130
    guides).  In full 128-window batches the fp4 engine tests three guide
    tiles per chain and chains the products of the one or two tiles left
    over across window tiles (4 tiles: a triple, then one tile; 5: a triple,
    then two; 1 and 2: leftovers only).  Near-copies of guides from every
    tile are planted, so candidates fall in every chain position."""
    from brute import IUPAC, plant
    rng = np.random.default_rng(23)
    contigs, _, _ = build_case(23, length=(60_000, 90_000), n_guides=1)
    guides = []
    for pam, n in (("NG", 40), ("NGG", 70), ("NNNRRT", 130), ("NNGRRT", 100), ("NAG", 20)):
        guides += [("".join(rng.choice(list("ACGT"), size=20)), pam) for _ in range(n)]
    for gi in rng.choice(len(guides), size=90, replace=False):
        q, pam = guides[int(gi)]
        pam_seq = "".join(rng.choice(list(IUPAC[c])) for c in pam)
        plant(rng, contigs, q, pam_seq, int(rng.integers(0, 4)), "+" if rng.random() < 0.5 else "-")
    budgets = [int(rng.integers(2, 6)) for _ in guides]
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    specs = [GuideSpec(s, p, k) for (s, p), k in zip(guides, budgets)]
    want = oracle_set(genome, guides, None, mm_per=budgets)
    assert len({g for g, *_ in want}) >= 40
    assert hit_set(scan(dg, specs, tuning=dict(engine)).hits) == want


def test_fetch_view_equals_fetch(gpu):
    """The page-locked fetch returns the same records as the copying fetch,
    also after the buffer grows and when capped."""
    contigs, guides, max_mm = build_case(21, length=(40_000, 60_000), n_guides=6)
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    p = Plan(dg, [GuideSpec(s, pm, 6) for s, pm in guides], hit_capacity=1 << 20)
    p.launch()
    n, _ = p.wait()
    assert n > 10
    a = p.fetch(n)
    for cap in (7, None):
        v = p.fetch_view(cap)
        assert np.array_equal(v, a[: v.size]) and v.size == (7 if cap else n)
    p.close()


def test_direct_compare_every_guide_count(gpu):
    """PAM families of 1..8 guides, dense and sparse PAMs, per-guide budgets and
    planted near-copies: the in-place compare (every guide count), the window
    buffer and fp4 report the oracle's sites."""
    from brute import IUPAC, plant
    rng = np.random.default_rng(41)
    contigs, _, _ = build_case(41, length=(40_000, 60_000), n_guides=1)
    pams = ["NG", "NNNRRT", "NRG", "NNG", "TG", "NNNNRT", "YG", "NGG"]
    guides = []
    for n, pam in enumerate(pams, start=1):
        guides += [("".join(rng.choice(list("ACGT"), size=20)), pam) for _ in range(n)]
    for gi in rng.choice(len(guides), size=30, replace=False):
        q, pam = guides[int(gi)]
        pam_seq = "".join(rng.choice(list(IUPAC[c])) for c in pam)
        plant(rng, contigs, q, pam_seq, int(rng.integers(0, 5)), "+" if rng.random() < 0.5 else "-")
    budgets = [int(rng.integers(0, 9)) for _ in guides]
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    specs = [GuideSpec(s, p, k) for (s, p), k in zip(guides, budgets)]
    want = oracle_set(genome, guides, None, mm_per=budgets)
    assert len(want) > 100 and len({g for g, *_ in want}) >= 25
    for t in ({"engine": 1}, {"engine": 1, "direct_max": 8}, {"engine": 1, "direct_max": -1},
              {"engine": 1, "direct_max": 3}, {"engine": 3}):
        r = scan(dg, specs, tuning=t)
        assert sorted(n for n, _ in r.families) == list(range(1, 9)), t
        assert hit_set(r.hits) == want, t


def test_auto_engine_choice(gpu):
    """Engine 0 takes the cheaper engine by the per-family cost model: the VALU
    compare for small sets (in-place families), fp4 for larger families, and
    the seed engine for families of thousands of guides when every budget is
    at most 4."""
    contigs, _, _ = build_case(3, length=(5_000, 8_000), n_guides=1)
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    rng = np.random.default_rng(2)

    def engine_for(counts, budget=4):
        specs = [GuideSpec("".join(rng.choice(list("ACGT"), size=20)), pam, budget)
                 for pam, n in counts.items() for _ in range(n)]
        p = Plan(dg, specs, hit_capacity=1 << 12)
        try:
            return p.tuning()["engine"]
        finally:
            p.close()

    assert engine_for({"NGG": 1, "NNNRRT": 1, "NNGRRT": 1, "NG": 1}) == 1
    assert engine_for({"NGG": 36, "NNNRRT": 21, "NNGRRT": 16, "NG": 7}) == 1
    assert engine_for({"NGG": 64, "NNNRRT": 64, "NNGRRT": 64, "NG": 64}) == 3
    assert engine_for({"NGG": 512, "NNNRRT": 256, "NNGRRT": 128, "NG": 128}) == 3
    assert engine_for({"NGG": 2048}) == 4
    assert engine_for({"NG": 1024, "NNNRRT": 1024}) == 4
    assert engine_for({"NGG": 2048}, budget=5) == 3


def block_genome(seed=3, blocks=24, block_len=1500):
    """Random contigs interleaved with poly-A and poly-T blocks and N runs: a
    poly-A guide has tens of thousands of sites on each strand (poly-T blocks
    read as poly-A on the reverse-complement array)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i in range(blocks):
        parts.append("".join(rng.choice(list("ACGT"), size=int(rng.integers(1000, 3000)))))
        parts.append(("A" if i % 2 else "T") * block_len)
        if i % 5 == 0:
            parts.append("N" * int(rng.integers(200, 900)))
    seq = "".join(parts)
    cut = len(seq) // 3
    return [("c0", seq[:cut]), ("c1", seq[cut:])]


def keep_lowest_set(sites, cap):
    by = {}
    for g, s, p, m in sites:
        by.setdefault((g, s), []).append((p, m))
    out = set()
    for (g, s), v in by.items():
        out |= {(g, s, p, m) for p, m in sorted(v)[:cap]}
    return out


@pytest.mark.parametrize("cap", [20, 700])
@pytest.mark.parametrize("engine", [pytest.param({"engine": 1}, id="valu"),
                                    pytest.param({"engine": 3}, id="fp4")])
def test_scan_capped_bounds_records_and_keeps_lowest(gpu, monkeypatch, engine, cap):
    """A heavy multi-mapper (poly-A, PAM NNN, <= 5 mm) among ordinary guides:
    with a record limit of 1024 the guide set is split until the multi-mapper
    is alone, whose strands are then scanned over position ranges; no fetch
    exceeds the limit, the per-strand counts are exact and each (guide,
    strand) keeps its lowest-position `cap` sites, all against the oracle.
    The multi-mapper has 17,919 / 17,944 sites on '+' / '-' and another guide
    25 / 34, so cap 20 also truncates inside a group fetch."""
    from beguidedoff_amd import batched as B

    contigs = block_genome()
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    rng = np.random.default_rng(8)
    guides = [("".join(rng.choice(list("ACGT"), size=20)), p) for p in ("NGG", "NG", "NNNRRT") * 3]
    guides.insert(4, ("A" * 20, "NNN"))
    guides.append(("AAAAAAAAAAAAAAAAAAAC", "NG"))
    specs = [GuideSpec(s, p, 5) for s, p in guides]
    want = oracle_set(genome, guides, 5)
    per = {}
    for g, s, _, _ in want:
        per[(g, s)] = per.get((g, s), 0) + 1
    assert per[(4, 0)] > 10_000 and per[(4, 1)] > 10_000
    fetched = []
    orig = B.Plan.fetch

    def spy(self, n=None):
        out = orig(self, n)
        fetched.append(out.size)
        return out

    monkeypatch.setattr(B.Plan, "fetch", spy)
    r = B.scan_capped(dg, specs, cap, tuning=engine, max_records=1024)
    assert fetched and max(fetched) <= 1024
    assert hit_set(r.hits) == keep_lowest_set(want, cap)
    counts = np.zeros((len(guides), 2), np.int64)
    for (g, s), c in per.items():
        counts[g, s] = c
    assert np.array_equal(r.counts, counts)
    assert r.truncated.tolist() == [bool((counts[i] > cap).any()) for i in range(len(guides))]
    # without a cap and limit the same launch sequence returns every site
    full = B.scan_capped(dg, specs, 1 << 30, tuning=engine)
    assert hit_set(full.hits) == want and full.launches == 1
    with pytest.raises(B.RecordLimitExceeded):
        scan(dg, specs, tuning=engine, max_records=1024)


# ----------------------------------------------------------- seed engine
#
# Engine 4 verifies only the guides that share one of five exact 4-base blocks
# with a window (a site within 4 mismatches always does).  Its sites must equal
# the oracle's like every other engine's, at budgets 0..4.

SEED = {"engine": 4}


@pytest.mark.parametrize("seed", range(8))
def test_seed_engine_matches_oracle_small(gpu, seed):
    contigs, guides, max_mm = build_case(seed)
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"]) if "N" in "".join(
        s for _, s in contigs) else DeviceGenome(genome)
    nm = genome if dg.has_n_mask else {k: v for k, v in genome.items() if not k.startswith("n_mask")}
    r = scan(dg, [GuideSpec(s, p, max_mm) for s, p in guides], tuning=dict(SEED))
    assert hit_set(r.hits) == oracle_set(nm, guides, max_mm)


def test_seed_engine_families_budgets_ranges_and_replays(gpu):
    """12 PAM families (several launch groups), per-guide budgets 0..4,
    20/21-nt guides, position ranges, single strands and graph replays."""
    rng = np.random.default_rng(3)
    contigs, guides, _ = build_case(77, length=(20_000, 40_000), n_guides=12)
    pams = ["NGG", "NRG", "NNGRRT", "NNNRRT", "NG", "TTTV", "SGG", "WGG", "A", "NNNNNN", "BDHV", "GG"]
    guides = [(s, pams[i]) for i, (s, _) in enumerate(guides)]
    budgets = [int(rng.integers(0, 5)) for _ in guides]
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    specs = [GuideSpec(s, p, k) for (s, p), k in zip(guides, budgets)]
    want = oracle_set(genome, guides, None, mm_per=budgets)
    r = scan(dg, specs, tuning=dict(SEED))
    assert len(r.families) == 12 and hit_set(r.hits) == want
    L = genome["total_bp"]
    for (b, e), strands in [((1000, 17_000), 3), ((0, L // 2), 1), ((L // 3, L), 2)]:
        r = scan(dg, specs, pos_begin=b, pos_end=e, strands=strands, tuning=dict(SEED))
        assert hit_set(r.hits) == oracle_set(genome, guides, None, mm_per=budgets, pos=(b, e),
                                             strands=[s for s in (0, 1) if strands & (1 << s)])
    plan = Plan(dg, specs, hit_capacity=1 << 16, tuning=dict(SEED))
    for _ in range(3):
        plan.graph_launch()
        n, _ = plan.wait()
        assert hit_set(plan.fetch(n)) == want
        plan.launch(0, L // 2, 1)   # eager work between replays
        plan.wait()
    plan.close()


def test_seed_engine_large_family_split_duplicates_and_plants(gpu):
    """A family of 4,500 guides (two seed tables of at most 4,096 guides, two
    launches), exact duplicates and near-duplicates of guides (pairs matching
    in several blocks are recorded once per guide), planted sites at 0..4
    mismatches on both strands, budgets 0..4, and the capacity re-run."""
    from brute import IUPAC, plant
    rng = np.random.default_rng(41)
    contigs, _, _ = build_case(41, length=(40_000, 60_000), n_guides=1)
    seqs = ["".join(rng.choice(list("ACGT"), size=20)) for _ in range(4400)]
    seqs += seqs[:50]                                    # exact duplicates
    for q in seqs[50:100]:                               # one-base variants
        i = int(rng.integers(0, 20))
        seqs.append(q[:i] + "ACGT"[("ACGT".index(q[i]) + 1) % 4] + q[i + 1:])
    guides = [(q, "NGG") for q in seqs]
    for gi in rng.choice(len(guides), size=120, replace=False):
        q, pam = guides[int(gi)]
        pam_seq = "".join(rng.choice(list(IUPAC[c])) for c in pam)
        plant(rng, contigs, q, pam_seq, int(rng.integers(0, 5)), "+" if rng.random() < 0.5 else "-")
    budgets = [int(rng.integers(0, 5)) for _ in guides]
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    specs = [GuideSpec(s, p, k) for (s, p), k in zip(guides, budgets)]
    want = oracle_set(genome, guides, None, mm_per=budgets)
    assert len(want) >= 50
    r = scan(dg, specs, tuning=dict(SEED), hit_capacity=16)
    assert r.launches == 2 and len(r.families) == 2
    assert hit_set(r.hits) == want
    assert r.hits.size == len(want)   # no pair recorded twice


def test_seed_engine_verify_counts(gpu):
    """The verified-entry counts equal, per family, the sum over PAM-passing
    windows and blocks of the number of the family's guides with the
    window's 4 bases in that block (counted on the host)."""
    from beguidedoff_amd.pam import IUPAC_MASK
    rng = np.random.default_rng(8)
    contigs, _, _ = build_case(8, length=(30_000, 50_000), n_guides=1)
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    guides = [("".join(rng.choice(list("ACGT"), size=20)), pam)
              for pam in ("NGG", "NG", "NNNRRT") for _ in range(300)]
    specs = [GuideSpec(s, p, 3) for s, p in guides]
    plan = Plan(dg, specs, hit_capacity=1 << 16, tuning=dict(SEED))
    plan.launch()
    plan.wait()
    got = plan.seed_stats()
    plan.close()
    L = genome["total_bp"]
    for f, pam in enumerate(("NGG", "NG", "NNNRRT")):
        gcodes = np.stack([encode_dna(s)[:20] for s, p in guides if p == pam]).astype(np.int64)
        want = 0
        for key in ("genome_forward", "genome_rc"):
            c = synth.extract_codes(genome[key], 0, L).astype(np.int64)
            n = L - 26
            ok = np.ones(n, bool)
            for i, ch in enumerate(pam):
                ok &= ((IUPAC_MASK[ch] >> c[20 + i:20 + i + n]) & 1).astype(bool)
            idx = np.nonzero(ok)[0]
            for j in range(5):
                wk = sum(c[idx + 4 * j + b] << (2 * b) for b in range(4))
                gk = sum(gcodes[:, 4 * j + b] << (2 * b) for b in range(4))
                want += int(np.bincount(gk, minlength=256)[wk].sum())
        assert got[f] == want, (pam, got[f], want)


def test_seed_engine_needs_budgets_up_to_4(gpu):
    contigs, guides, _ = build_case(2)
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome)
    with pytest.raises(RuntimeError, match="budgets <= 4"):
        scan(dg, [GuideSpec(s, p, 5) for s, p in guides], tuning=dict(SEED))


@pytest.mark.parametrize("engine", [pytest.param({"engine": e}, id=str(e)) for e in (1, 2, 3, 4)])
@pytest.mark.parametrize("seed", [3, 7])
def test_strand_symmetry_and_guide_order(gpu, engine, seed):
    """Two invariants of the scan: the − strand sites of a genome are the +
    strand sites of the genome built from its contigs' reverse complements
    (same contig order: the reverse-complement array is reversed per contig),
    and a permuted guide order gives every guide the same sites."""
    from beguidedoff_amd.encoding import reverse_complement_seq

    contigs, guides, max_mm = build_case(seed, length=(3_000, 9_000), n_guides=6)
    g1 = synth.genome_from_strings(contigs)
    g2 = synth.genome_from_strings([(n, reverse_complement_seq(s)) for n, s in contigs])
    specs = [GuideSpec(s, p, max_mm) for s, p in guides]
    d1 = DeviceGenome(g1, g1["n_mask_le_fwd"], g1["n_mask_le_rc"])
    d2 = DeviceGenome(g2, g2["n_mask_le_fwd"], g2["n_mask_le_rc"])
    try:
        minus = {(g, p, m) for g, s, p, m in hit_set(scan(d1, specs, tuning=engine).hits) if s == 1}
        plus = {(g, p, m) for g, s, p, m in hit_set(scan(d2, specs, tuning=engine,
                                                          strands=1).hits)}
        assert minus == plus and len(minus) > 0
        perm = np.random.default_rng(seed).permutation(len(specs))
        got = {(int(perm[g]), s, p, m)
               for g, s, p, m in hit_set(scan(d1, [specs[i] for i in perm], tuning=engine).hits)}
        assert got == hit_set(scan(d1, specs, tuning=engine).hits)
    finally:
        d1.close()
        d2.close()


def test_seed_engine_three_even_parts(gpu):
    """A family of 9,000 guides splits into three tables of 3,000 (each guide
    in the part whose key lists for it are shortest so far).  The verified
    entries summed over the parts do not depend on the split: they equal the
    host count over all 9,000 guides; the sites equal the oracle's."""
    from beguidedoff_amd.pam import IUPAC_MASK
    rng = np.random.default_rng(12)
    contigs, _, _ = build_case(12, length=(12_000, 20_000), n_guides=1)
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    seqs = ["".join(rng.choice(list("ACGT"), size=20)) for _ in range(9000)]
    guides = [(q, "NGG") for q in seqs]
    specs = [GuideSpec(q, "NGG", 3) for q in seqs]
    plan = Plan(dg, specs, hit_capacity=1 << 16, tuning=dict(SEED))
    plan.launch()
    n, _ = plan.wait()
    assert [g for g, _ in plan.family_stats()] == [3000, 3000, 3000]
    got = sum(plan.seed_stats())
    hits = plan.fetch(n)
    plan.close()
    dg.close()
    L = genome["total_bp"]
    gcodes = np.stack([encode_dna(q)[:20] for q in seqs]).astype(np.int64)
    want = 0
    for key in ("genome_forward", "genome_rc"):
        c = synth.extract_codes(genome[key], 0, L).astype(np.int64)
        m = L - 26
        ok = np.ones(m, bool)
        for i, ch in enumerate("NGG"):
            ok &= ((IUPAC_MASK[ch] >> c[20 + i:20 + i + m]) & 1).astype(bool)
        idx = np.nonzero(ok)[0]
        for j in range(5):
            wk = sum(c[idx + 4 * j + b] << (2 * b) for b in range(4))
            gk = sum(gcodes[:, 4 * j + b] << (2 * b) for b in range(4))
            want += int(np.bincount(gk, minlength=256)[wk].sum())
    assert got == want
    assert hit_set(hits) == oracle_set(genome, guides, 3)
