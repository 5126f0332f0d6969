"""Unit tests of the host-side semantics modules (mirroring the reference's
tests/test_encoding.py, test_pam.py, test_scoring.py, test_base_editing.py,
test_io.py, test_genome_converter.py)."""
import numpy as np
import pandas as pd
import pytest

from beguidedoff_amd.base_editing import (edit_window_mask, editable_base_in_window,
                                          get_edit_window_from_profile,
                                          get_trinucleotide_contexts, has_editable_base)
from beguidedoff_amd.encoding import (decode_dna, decode_guide_64bit, encode_dna,
                                      encode_guide_64bit, hamming_distance_2bit,
                                      mismatch_positions_2bit, pack_2bit_to_uint32,
                                      reverse_complement_seq, unpack_uint32_to_2bit)
from beguidedoff_amd.genome_converter import convert, load_genome, verify
from beguidedoff_amd.io import load_guides_from_tsv, offtargets_to_dataframe
from beguidedoff_amd.pam import all_possible_pams, encode_pam_mask, group_guides_by_pam, validate_pam
from beguidedoff_amd.scoring import calculate_cctop_score, cctop_table
from beguidedoff_amd.types import GuideInput, OffTarget


def test_codec():
    assert decode_dna(encode_dna("ACGTACGTNNACGT")) == "ACGTACGTAAACGT"
    assert reverse_complement_seq("ACGT") == "ACGT" and reverse_complement_seq("ATCG") == "CGAT"
    e = encode_dna("ACGT" * 10)
    assert np.array_equal(unpack_uint32_to_2bit(pack_2bit_to_uint32(e), 40), e)
    v = encode_guide_64bit(encode_dna("ACGTACGTACGTACGTACGT"))
    assert np.array_equal(decode_guide_64bit(v, 20), encode_dna("ACGTACGTACGTACGTACGT"))
    a = encode_guide_64bit(encode_dna("ACGTACGTACGTACGTACGT"))
    b = encode_guide_64bit(encode_dna("ACGTACGTACGTACGTACGA"))
    c = encode_guide_64bit(encode_dna("TCGTACGTACGTACGTACGT"))
    assert hamming_distance_2bit(a, b) == 1 and hamming_distance_2bit(a, a) == 0
    assert mismatch_positions_2bit(a, b) == [20] and mismatch_positions_2bit(a, c) == [1]
    # word layout: base i at bits 2*(i % 16) of word i // 16
    w = pack_2bit_to_uint32(encode_dna("C" + "A" * 15 + "T"))
    assert w.tolist() == [1, 3]


def test_pam_masks():
    validate_pam("NGG")
    validate_pam("NNGRRT")
    with pytest.raises(ValueError):
        validate_pam("NNNNNNN")
    assert encode_pam_mask("NGG").tolist() == [0b1111, 0b0100, 0b0100, 0, 0, 0]
    pams = all_possible_pams("NGG")
    assert len(pams) == 4 and "AGG" in pams and "GGG" in pams
    gs = [GuideInput("g1", "A" * 20, "NGG", "3prime", 4, 8, "CBE"),
          GuideInput("g2", "A" * 20, "NNGRRT", "3prime", 4, 8, "CBE"),
          GuideInput("g3", "A" * 20, "NGG", "3prime", 4, 8, "ABE")]
    groups = group_guides_by_pam(gs)
    assert len(groups[("NGG", "3prime")]) == 2 and len(groups[("NNGRRT", "3prime")]) == 1


def test_cctop():
    assert calculate_cctop_score([]) == 0.0
    assert abs(calculate_cctop_score([20]) - 1.2 ** 20) < 1e-9
    assert abs(calculate_cctop_score([1, 5, 10]) - (1.2 + 1.2 ** 5 + 1.2 ** 10)) < 1e-9
    assert calculate_cctop_score([10, 15]) == pytest.approx(
        calculate_cctop_score([10]) + calculate_cctop_score([15]))
    t = cctop_table()
    assert t[0] == 1.0 and t[1] == 1.2 and t[20] == 1.2 ** 20
    assert all(t[i] < t[i + 1] for i in range(20))


def test_editing_window():
    s = "ACGTACGTACGTACGTACGT"
    assert 6 in editable_base_in_window(s, "CBE", 4, 8)
    assert {1, 5} <= set(editable_base_in_window(s, "ABE", 1, 5))
    assert editable_base_in_window("G" * 20, "CBE", 1, 10) == []
    assert get_trinucleotide_contexts("ACGTACGT", [3]) == ["CGT"]
    assert has_editable_base(s, "CBE", 4, 8) and not has_editable_base("G" * 20, "CBE", 1, 10)
    with pytest.raises(ValueError):
        editable_base_in_window(s, "XBE", 1, 3)
    # mask = indices inspected, including Python negative indexing at start <= 0
    assert edit_window_mask(4, 8) == 0b11111000
    assert edit_window_mask(0, 2) == (1 << 19) | 0b11
    assert edit_window_mask(18, 25) == 0b111 << 17


PROFILE = ("name\tediting\tgeneral_edit_start\tgeneral_edit_end\thigh_edit_start\thigh_edit_end\t"
           "pam\tcas\tsource\nABE8e\tABE\t2\t13\t3\t8\tNGG\tSpCas9\tx\n"
           "CBE6d\tCBE\t2\t10\t6\t6\tNGG\tSpCas9\tx\nSaABE8e\tABE\t5\t13\t5\t13\tNNGRRT\tSaCas9\tx\n")


def test_io(tmp_path):
    prof = tmp_path / "p.tsv"
    prof.write_text(PROFILE)
    assert get_edit_window_from_profile("ABE8e", str(prof)) == (2, 13, "NGG")
    with pytest.raises(ValueError, match="DoesNotExist"):
        get_edit_window_from_profile("DoesNotExist", str(prof))
    t = tmp_path / "g.tsv"
    t.write_text("chromosome\tposition\tstrand_used\tguide_rna_sequence\tbase_editor\n"
                 "1\t100\t+\tACGTACGTACGTACGTACGT\tABE8e\n2\t200\t-\tGGGGGGGGGGGGGGGGGGGG\tCBE6d\n")
    g = load_guides_from_tsv(str(t), presets_path=str(prof))
    assert [x.guide_id for x in g] == ["row_0", "row_1"]
    assert (g[0].pam, g[0].edit_window_start, g[0].edit_window_end, g[0].editor_type) == ("NGG", 2, 13, "ABE")
    assert (g[0].on_target_chrom, g[0].on_target_pos, g[0].on_target_strand) == ("1", 100, "+")
    assert g[1].editor_type == "CBE" and g[1].base_editor == "CBE6d"
    t2 = tmp_path / "e.tsv"
    t2.write_text("guide_seq\tpam\tedit_window_start\tedit_window_end\teditor_type\n"
                  "ACGTACGTACGTACGTACGT\tNGG\t4\t8\tCBE\n")
    g2 = load_guides_from_tsv(str(t2))
    assert g2[0].editor_type == "CBE" and g2[0].base_editor is None
    t3 = tmp_path / "bad.tsv"
    t3.write_text("foo\tbar\n1\t2\n")
    with pytest.raises(ValueError, match="guide_rna_sequence"):
        load_guides_from_tsv(str(t3), presets_path=str(prof))
    t4 = tmp_path / "p2.tsv"
    t4.write_text("guide_rna_sequence\tbase_editor\nACGTACGTACGTACGTACGT\tABE8e\n")
    with pytest.raises(ValueError, match="presets_path"):
        load_guides_from_tsv(str(t4))
    ot = OffTarget("chr1", 100, "+", "ACGTACGTACGTACGTACGA", "AGG", "NGG", "ACGTACGTACGTACGTACGT",
                   "ABE8e", 1, [20], 1.2, [5], ["TAC"])
    df = offtargets_to_dataframe([ot], n_offtargets_below_threshold=3)
    assert list(df.columns)[:3] == ["guide_rna_sequence", "base_editor", "ot_rank"]
    assert df.iloc[0]["base_editor"] == "ABE8e" and df.iloc[0]["n_offtargets_below_threshold"] == 3


def test_genome_converter(tmp_path):
    fa = tmp_path / "m.fa"
    fa.write_text(">chr1\nACGTACGTACGT\n>chr2\nTGCATGCATGCA\n")
    convert(str(fa), str(tmp_path / "m.npz"), verbose=False)
    d = load_genome(str(tmp_path / "m.npz"))
    assert d["total_bp"] == 24 and list(d["chrom_names"]) == ["chr1", "chr2"] and "genome_rc" in d
    fa2 = tmp_path / "v.fa"
    fa2.write_text(">chr1\n" + "ACGTNNACGT" * 5 + "\n")
    convert(str(fa2), str(tmp_path / "v.npz"), verbose=False)
    assert verify(str(tmp_path / "v.npz"), str(fa2), sample_n=100, verbose=False)


def test_guide_descs_fields():
    """Batched-scan guide descriptors: the first 20 bases 2 bits each (A0 C1 G2
    T3, base i at bits 2i), IUPAC PAM nibbles and length, budget, editor target
    (A for ABE, C for CBE/CGE) and the edit window's bit mask."""
    from beguidedoff_amd.batched import GuideSpec, guide_descs
    d = guide_descs([GuideSpec("ACGT" * 5, "NGG", 3, "ABE", (4, 8)),
                     GuideSpec("T" * 21, "NNGRRT", 5, "CBE", (3, 9)),
                     GuideSpec("ACGT" * 5, "NGG", 0, "CGE", (4, 8))])
    assert d["seq2bit"].tolist() == [0xE4E4E4E4E4, (1 << 40) - 1, 0xE4E4E4E4E4]
    assert d["pam_mask"].tolist() == [[15, 4, 4, 0, 0, 0], [15, 15, 4, 5, 5, 8],
                                      [15, 4, 4, 0, 0, 0]]
    assert d["pam_len"].tolist() == [3, 6, 3]
    assert d["max_mm"].tolist() == [3, 5, 0]
    assert d["editor"].tolist() == [0, 1, 1]
    assert d["edit_mask"].tolist() == [0b11111000, 0b111111100, 0b11111000]
    assert guide_descs([]).size == 0
    with pytest.raises(ValueError, match="max_mm"):
        guide_descs([GuideSpec("A" * 20, "NGG", 21)])
    with pytest.raises(ValueError, match="editor_type"):
        guide_descs([GuideSpec("A" * 20, "NGG", 2, "XBE")])
