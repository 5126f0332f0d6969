"""GPU finder: both engines reproduce the oracle finder's table and the
reference's known answers."""
import logging

import numpy as np
import pandas as pd
import pytest

import scenarios as S
from beguidedoff_amd import beguidedoff as make_finder
from beguidedoff_amd import finder as finder_module
from beguidedoff_amd import synth
from beguidedoff_amd.types import GuideInput
from oracle.finder_oracle import FinderOracle, editable
from test_host_epilogue import random_case

pytestmark = pytest.mark.gpu
ENGINES = ["plain", "batched"]


def finder(contigs_or_genome, engine, **kw):
    g = contigs_or_genome if isinstance(contigs_or_genome, dict) else synth.genome_from_strings(contigs_or_genome)
    return make_finder(genome=g, engine=engine, **kw)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("seed", range(4))
def test_tables_match_oracle(gpu, engine, seed):
    genome, guides = random_case(seed)
    for kw in ({}, {"return_all": True, "count_score_threshold": 5.0}):
        got = finder(genome, engine, max_mismatches=5).search(guides=guides, verbose=False, **kw)
        want = FinderOracle(genome, max_mismatches=5).search(guides, **kw)
        pd.testing.assert_frame_equal(got, want)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("max_mm", [3, 5])
def test_small_config_table(gpu, engine, max_mm):
    """3 contigs (10 Mbp, 1% N), 8 random NGG guides: at <= 3 mm random
    guides rarely have sites; at <= 5 mm the table has dozens of rows."""
    w = synth.workload("small")
    got = finder(w.genome, engine, max_mismatches=max_mm).search(
        guides=w.guides, return_all=True, count_score_threshold=5.0, verbose=False)
    want = FinderOracle(w.genome, max_mismatches=max_mm).search(
        w.guides, return_all=True, count_score_threshold=5.0)
    pd.testing.assert_frame_equal(got, want)
    if max_mm == 5:
        assert len(got) > 10


@pytest.mark.parametrize("engine", ENGINES)
def test_overflow_rescan_counts_every_site(gpu, engine):
    f = finder(S.repeat_genome(), engine, max_mismatches=4, n_top_offtargets=3)
    df = f.search(guides=S.single_guide(), count_score_threshold=5.0, verbose=False)
    assert len(df) == 3 and df.iloc[0]["n_mismatches"] == 1
    assert df.iloc[0]["n_offtargets_below_threshold"] == S.N_COPIES


@pytest.mark.parametrize("engine", ENGINES)
def test_cap_truncates_and_warns(gpu, engine, monkeypatch, caplog):
    f = finder(S.repeat_genome(), engine, max_mismatches=4, n_top_offtargets=3)
    monkeypatch.setattr(finder_module, "_MAX_RESULTS_PER_GUIDE",
                        finder_module._DEFAULT_RESULTS_PER_GUIDE)
    with caplog.at_level(logging.WARNING):
        df = f.search(guides=S.single_guide(), return_all=True, verbose=False)
    assert len(df) == finder_module._DEFAULT_RESULTS_PER_GUIDE
    msgs = [r.getMessage() for r in caplog.records]
    assert any("Truncated off-target enumeration" in m for m in msgs)
    assert any(str(S.N_COPIES) in m for m in msgs)


@pytest.mark.parametrize("engine", ENGINES)
def test_n_gate_adjacency_and_no_mask(gpu, engine):
    df = finder(S.n_gate_genome(), engine).search(guides=S.poly_guides(), return_all=True, verbose=False)
    for _, grp in df.groupby("guide_rna_sequence"):
        assert len(grp) < 15 and (grp["n_mismatches"] == 0).any()
        assert not ((grp["ot_position"] >= 104) & (grp["ot_position"] <= 603)).any()
    df = finder(S.n_adjacent_genome(), engine).search(
        guides=[GuideInput("polyA", "A" * 20, "NGG", "3prime", 4, 8, "CBE")], return_all=True,
        verbose=False)
    assert len(df) == 2 and (df["ot_strand"] == "-").all()
    z = df[df["n_mismatches"] == 0]
    assert len(z) == 1 and z.iloc[0]["ot_sequence"] == "A" * 20
    g = synth.genome_from_strings(S.no_n_mask_genome())
    for k in ("n_mask", "n_mask_le_fwd", "n_mask_le_rc"):
        g.pop(k)
    df = make_finder(genome=g, engine=engine).search(
        guides=[GuideInput("polyA", "A" * 20, "AAA", "3prime", 4, 8, "CBE")], return_all=True,
        verbose=False)
    assert (df["n_mismatches"] == 0).sum() >= 1


@pytest.mark.parametrize("engine", ENGINES)
def test_auto_chunking_2100_keys(gpu, engine):
    contigs, guides, planted = S.auto_chunk()
    df = finder(contigs, engine).search(guides=guides, verbose=False)
    rows = df[df["guide_rna_sequence"] == planted.guide_seq]
    assert len(rows) == 3 and (rows["n_mismatches"] == 0).all()


@pytest.mark.parametrize("engine", ENGINES)
def test_exotic_pams(gpu, engine):
    contigs, inputs = S.exotic_pam()
    got = finder(contigs, engine).search(guides=inputs, return_all=True, verbose=False)
    want = FinderOracle(synth.genome_from_strings(contigs)).search(inputs, return_all=True)
    pd.testing.assert_frame_equal(got, want)


@pytest.mark.parametrize("engine", ENGINES)
def test_positions_ontarget_and_thresholds(gpu, engine):
    df = finder(S.offtarget_pos31_genome(), engine).search(
        guides=[GuideInput("g1", S.GUIDE, "NGG", "3prime", 4, 8, "CBE")], verbose=False)
    assert df["ot_position"].tolist() == [31] and abs(df.iloc[0]["score"] - 1.2 ** 3) < 1e-9
    ex = GuideInput("g1", S.GUIDE, "NRG", "3prime", 4, 8, "CBE", on_target_chrom="chr1",
                    on_target_pos=37, on_target_strand="-")
    df = finder(S.minus_strand_genome(), engine).search(guides=[ex], verbose=False)
    assert df["ot_position"].tolist() == [83]
    f = finder(S.threshold_genome(), engine)
    for thr, count in ((5.0, 3), (3.0, 2), (2.0, 1), (0.5, 0)):
        df = f.search(guides=[S.threshold_guide()], count_score_threshold=thr, verbose=False)
        assert df["n_offtargets_below_threshold"].iloc[0] == count


def test_tsv_input_and_output(gpu, tmp_path):
    w = synth.workload("small")
    tsv = tmp_path / "g.tsv"
    rows = ["guide_rna_sequence\tbase_editor"] + [f"{g.guide_seq}\tABE8e" for g in w.guides]
    tsv.write_text("\n".join(rows) + "\n")
    out = tmp_path / "o.tsv"
    df = make_finder(genome=w.genome, max_mismatches=5).search(input_file=str(tsv),
                                                             output_file=str(out), verbose=False)
    assert len(df) > 0 and set(df["base_editor"]) == {"ABE8e"}
    back = pd.read_csv(out, sep="\t")
    assert len(back) == len(df) and list(back.columns) == list(df.columns)


@pytest.mark.parametrize("engine", ENGINES)
def test_multimapper_truncation_is_bounded(gpu, engine, monkeypatch, caplog):
    """A poly-A guide over A/T blocks at <= 5 mm with a ceiling of 1,000 sites
    per strand: the table keeps exactly the ceiling per strand, the warning
    carries the exact genome-wide count, and (batched engine) no launch holds
    more than the record limit."""
    from beguidedoff_amd import batched as B
    from test_gpu_kernels import block_genome, oracle_set

    genome = synth.genome_from_strings(block_genome())
    want = oracle_set(genome, [("A" * 20, "NNN")], 5)
    per_strand = [sum(1 for h in want if h[1] == s) for s in (0, 1)]
    assert min(per_strand) > 1000
    monkeypatch.setattr(finder_module, "_MAX_RESULTS_PER_GUIDE", 1000)
    fetched = []
    orig = B.Plan.fetch
    monkeypatch.setattr(B.Plan, "fetch", lambda self, n=None: fetched.append(
        (n if n is not None else self.hit_capacity)) or orig(self, n))
    f = finder(genome, engine, max_mismatches=5, results_per_guide=1000)
    f.max_records = 4096
    guides = [GuideInput("polyA", "A" * 20, "NNN", "3prime", 4, 8, "ABE"),
              GuideInput("other", "ACGTTGCAACGTTGCAACGT", "NGG", "3prime", 4, 8, "ABE")]
    with caplog.at_level(logging.WARNING):
        df = f.search(guides=guides, return_all=True, verbose=False)
    rows = df[df["guide_rna_sequence"] == "A" * 20]
    assert len(rows) == 2000
    assert (rows["ot_strand"] == "+").sum() == 1000
    msgs = [r.getMessage() for r in caplog.records]
    assert any("Truncated off-target enumeration" in m and str(len(want)) in m for m in msgs)
    if engine == "batched":
        assert fetched and max(fetched) <= 4096


@pytest.mark.parametrize("engine", ENGINES)
def test_chromosome_name_tie_break(gpu, engine):
    """Equal scores are ordered by the chromosome name as a string
    (finder.py:552): chr1, chr10, chr10, chr11, ... on twelve contigs."""
    contigs, zero = S.chrom_name_tie()
    f = finder(contigs, engine)
    df = f.search(guides=[S.tie_guide()], verbose=False)
    assert list(zip(df["ot_chromosome"], df["ot_position"])) == zero[:3]
    df = f.search(guides=[S.tie_guide()], return_all=True, verbose=False)
    z = df[df["score"] == 0]
    assert list(zip(z["ot_chromosome"], z["ot_position"])) == zero
    want = FinderOracle(synth.genome_from_strings(contigs)).search([S.tie_guide()], return_all=True)
    pd.testing.assert_frame_equal(df, want)


def test_batched_truncation_at_the_reference_ceiling(gpu, caplog):
    """The reference's ceiling, 2,621,440 sites per strand (finder.py:111): a
    3 Mbp poly-A contig (then a poly-C one) gives a poly-A guide with PAM NNN
    about 3 M '+' sites and no '-' sites.  The batched finder keeps the
    lowest-position 2,621,440, all editable 0-mm sites, so the count below a
    threshold is exactly the ceiling; the warning carries the exact
    genome-wide count (the oracle's); the top rows are the first positions."""
    from test_gpu_kernels import oracle_set

    contigs = [("chrA", "A" * 3_000_000), ("chrB", "C" * 2_000)]
    genome = synth.genome_from_strings(contigs)
    want = oracle_set(genome, [("A" * 20, "NNN")], 4)
    assert all(s == 0 for _, s, _, _ in want) and len(want) > 2_999_000
    f = finder(genome, "batched", max_mismatches=4)
    guide = GuideInput("polyA", "A" * 20, "NNN", "3prime", 4, 8, "ABE")
    with caplog.at_level(logging.WARNING):
        df = f.search(guides=[guide], count_score_threshold=5.0, verbose=False)
    assert df["n_offtargets_below_threshold"].tolist() == [2_621_440] * 3
    assert df["ot_position"].tolist() == [1, 2, 3] and (df["ot_chromosome"] == "chrA").all()
    msgs = [r.getMessage() for r in caplog.records if "Truncated off-target" in r.getMessage()]
    assert len(msgs) == 1 and f" {len(want)} raw" in msgs[0] and "keeping the first 2621440" in msgs[0]


def assembly_case(seed):
    """random_case plus: a second input sharing an id with another sequence,
    an input sharing both id and sequence (a repeated block), an on-target
    input (its own key), and many contigs with names out of string order."""
    genome, guides = random_case(seed)
    g0 = guides[0]
    guides.append(GuideInput(g0.guide_id, guides[1].guide_seq, guides[1].pam, "3prime", 2, 9, "ABE",
                             base_editor="shared"))
    guides.append(GuideInput(guides[2].guide_id, guides[2].guide_seq, guides[2].pam, "3prime",
                             guides[2].edit_window_start, guides[2].edit_window_end,
                             guides[2].editor_type))
    names = [str(n) for n in genome["chrom_names"]]
    ot = GuideInput("ot", guides[3].guide_seq, guides[3].pam, "3prime", 1, 20, "ABE",
                    on_target_chrom=names[1], on_target_pos=5, on_target_strand="-")
    guides.append(ot)
    return genome, guides


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("kw", [{}, {"return_all": True, "count_score_threshold": 5.0},
                                {"count_score_threshold": 3.0}, {"n_top": 1}, {"n_top": 0}])
def test_device_assembly_equals_host_and_oracle(gpu, seed, kw):
    """The device table (bgo_plan_assemble) equals the host assembly of the
    same records and the finder oracle, row for row."""
    kw = dict(kw)
    top = kw.pop("n_top", 3)
    genome, guides = assembly_case(seed)
    dev = finder(genome, "batched", max_mismatches=5, n_top_offtargets=top)
    host = finder(genome, "batched", max_mismatches=5, n_top_offtargets=top)
    host.device_assembly = False
    got = dev.search(guides=guides, verbose=False, **kw)
    ref = host.search(guides=guides, verbose=False, **kw)
    want = FinderOracle(genome, max_mismatches=5, n_top_offtargets=top).search(guides, **kw)
    pd.testing.assert_frame_equal(got, ref)
    pd.testing.assert_frame_equal(got, want)
    if top > 0:
        assert len(got) > 5


def test_device_assembly_names_out_of_order(gpu):
    """Twelve contigs chr1..chr12 and equal scores: the device order follows
    the name as a string (finder.py:552), with a shared id over two keys."""
    contigs, zero = S.chrom_name_tie()
    g = synth.genome_from_strings(contigs)
    inputs = [S.tie_guide(), GuideInput("tie", "ACGTACGTACGTACGTACGT", "NGG", "3prime", 4, 8, "CBE")]
    f = finder(g, "batched")
    df = f.search(guides=inputs, return_all=True, verbose=False)
    host = finder(g, "batched")
    host.device_assembly = False
    pd.testing.assert_frame_equal(df, host.search(guides=inputs, return_all=True, verbose=False))
    z = df[df["score"] == 0]
    assert list(zip(z["ot_chromosome"], z["ot_position"]))[:13] == zero


@pytest.mark.parametrize("seed", range(3))
def test_seed_engine_tables_match_oracle(gpu, seed):
    """The finder's table with the seed-and-verify compare (engine 4, budgets
    <= 4) equals the oracle finder's, assembled on the device and on the
    host."""
    genome, guides = assembly_case(seed)
    kw = {"return_all": True, "count_score_threshold": 5.0}
    want = FinderOracle(genome, max_mismatches=4).search(guides, **kw)
    for device_assembly in (True, False):
        f = finder(genome, "batched", max_mismatches=4, tuning={"engine": 4})
        f.device_assembly = device_assembly
        got = f.search(guides=guides, verbose=False, **kw)
        pd.testing.assert_frame_equal(got, want)
    assert len(want) > 5


def test_searches_in_a_row_reuse_the_plan(gpu):
    """The batched finder keeps its device-table plan between searches: on
    one finder, searches with other guide sets, orders and options each give
    the oracle's table, and close() releases the plan."""
    genome, guides = random_case(1)
    f = finder(genome, "batched", max_mismatches=5)
    oracle = FinderOracle(genome, max_mismatches=5)
    for sub in (guides, guides[: max(1, len(guides) // 3)], guides[::-1], guides):
        for kw in ({}, {"return_all": True, "count_score_threshold": 5.0}):
            got = f.search(guides=sub, verbose=False, **kw)
            pd.testing.assert_frame_equal(got, oracle.search(sub, **kw))
    assert f._table_plan is not None
    f.close()
    assert f._table_plan is None
