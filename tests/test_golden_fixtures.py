"""The reference's test fixtures (tests/fixtures/mini_genome.fa and
test_guides.tsv, kept as data under tests/golden/): loader fields, the
encoded genome, and the searches.  Neither guide has a site: the genome's only
GG follows base 20 directly, so no NGG window fits, and the extensive scan
range of a 32-bp genome is positions 0-5."""
from pathlib import Path

import numpy as np
import pytest

from beguidedoff_amd import beguidedoff as make_finder
from beguidedoff_amd import synth
from beguidedoff_amd.genome_converter import encode_contigs, read_fasta
from beguidedoff_amd.io import load_guides_from_tsv
from oracle.encode_oracle import encode as oracle_encode
from oracle.fast_finder_oracle import FastFinderOracle
from oracle.finder_oracle import FinderOracle

GOLD = Path(__file__).resolve().parent / "golden"


def fixture_guides():
    return load_guides_from_tsv(str(GOLD / "test_guides.tsv"))


def fixture_contigs():
    return list(read_fasta(str(GOLD / "mini_genome.fa")))


def test_fixture_guides_load_like_the_reference():
    rows = [(g.guide_id, g.guide_seq, g.pam, g.pam_position, g.edit_window_start,
             g.edit_window_end, g.editor_type) for g in fixture_guides()]
    # io.py:51-75: ids are row_<i>, the file's guide_id column is not read
    assert rows == [("row_0", "ACGTACGTACGTACGTACGT", "NGG", "3prime", 4, 8, "CBE"),
                    ("row_1", "TGCATGCATGCATGCATGCA", "NGG", "3prime", 4, 8, "ABE")]


def test_fixture_genome_encodes():
    contigs = fixture_contigs()
    assert contigs == [("chr1", "ACGTACGTACGTACGTACGTGGAAAAAAAAAA")]
    d = encode_contigs(contigs)
    # ACGT x 4 | ACGT GG A x 10: A C G T = 0 1 2 3 at bits 2i
    assert d["genome_forward"].tolist() == [0xE4E4E4E4, 0xE4 | 2 << 8 | 2 << 10]
    want = oracle_encode(contigs)
    for key in ("genome_forward", "genome_rc", "n_mask"):
        assert np.array_equal(d[key], want[key])


def test_fixture_searches_find_no_site_in_the_oracles():
    genome = synth.genome_from_strings(fixture_contigs())
    guides = fixture_guides()
    assert FinderOracle(genome, max_mismatches=4).search(guides, return_all=True).empty
    assert FastFinderOracle(genome).search(guides, return_all=True).empty


@pytest.mark.gpu
def test_fixture_searches_find_no_site_on_the_gpu(gpu):
    genome = synth.genome_from_strings(fixture_contigs())
    guides = fixture_guides()
    for engine in ("batched", "plain"):
        f = make_finder(genome=genome, max_mismatches=4, engine=engine)
        assert f.search(guides=guides, return_all=True, verbose=False).empty
    fast = make_finder(mode="fast", genome=genome)
    assert fast.search(guides=guides, return_all=True, verbose=False).empty
