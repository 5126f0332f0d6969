"""The command line (beguidedoff_amd/cli.py) against the reference's
subcommands (beguidedoff/cli.py:75-126): encode-genome and verify on the CPU,
search on the GPU compared with the finder's own TSV."""
import numpy as np
import pandas as pd
import pytest

from beguidedoff_amd import cli, synth
from beguidedoff_amd import beguidedoff as make_finder
from beguidedoff_amd.genome_converter import encode_contigs, load_genome, write_npz
from brute import random_contigs


def _fasta(path, contigs, width=60):
    with open(path, "w") as fh:
        for name, s in contigs:
            fh.write(f">{name} description\n")
            for i in range(0, len(s), width):
                fh.write(s[i:i + width] + "\n")


def test_encode_genome_and_verify(tmp_path):
    rng = np.random.default_rng(4)
    contigs = random_contigs(rng, [1500, 700, 2300], n_frac=0.05)
    contigs = [(f"chr{i + 1}", s) for i, (_, s) in enumerate(contigs)]
    fa, npz = tmp_path / "g.fa", tmp_path / "g.npz"
    _fasta(fa, contigs)
    assert cli.main(["encode-genome", "--fasta", str(fa), "--output", str(npz)]) == 0
    got, want = load_genome(str(npz)), encode_contigs(contigs)
    assert got["total_bp"] == 4500 and list(got["chrom_names"]) == ["chr1", "chr2", "chr3"]
    for key in ("genome_forward", "genome_rc", "n_mask", "chrom_offsets", "chrom_lengths"):
        assert np.array_equal(got[key], want[key]), key
    assert cli.main(["verify", "--encoded", str(npz), "--fasta", str(fa), "--sample-n", "500"]) == 0
    # an artefact of other sequences under the same names fails verification
    other = tmp_path / "o.npz"
    seqs = random_contigs(np.random.default_rng(5), [1500, 700, 2300])
    write_npz(str(other), encode_contigs([(n, s) for (n, _), (_, s) in zip(contigs, seqs)]))
    assert cli.main(["verify", "--encoded", str(other), "--fasta", str(fa), "--sample-n", "500"]) == 1


def test_missing_genome_or_index_exits_with_a_message(tmp_path):
    with pytest.raises(SystemExit, match="--index is required"):
        cli.main(["search", "--mode", "fast", "--guides", "g.tsv"])
    with pytest.raises(SystemExit, match="--genome is required"):
        cli.main(["search", "--guides", "g.tsv"])


def test_build_index_writes_the_encoded_genome(tmp_path):
    contigs = [("chr1", "ACGTNNACGTTTGCA" * 7), ("chr2", "GGGCCCAAATTT" * 5)]
    fa = tmp_path / "g.fa"
    _fasta(fa, contigs)
    assert cli.main(["build-index", "--fasta", str(fa), "--output", str(tmp_path / "i.npz")]) == 0
    got = load_genome(str(tmp_path / "i.npz"))
    want = encode_contigs(contigs)
    for key in ("genome_forward", "genome_rc", "n_mask"):
        assert np.array_equal(got[key], want[key])


@pytest.mark.gpu
def test_search_writes_the_finders_table(gpu, tmp_path):
    w = synth.workload("small")
    npz = tmp_path / "g.npz"
    write_npz(str(npz), w.genome)
    tsv = tmp_path / "g.tsv"
    rows = ["guide_rna_sequence\tbase_editor"] + [f"{g.guide_seq}\tABE8e" for g in w.guides]
    tsv.write_text("\n".join(rows) + "\n")
    out_cli, out_api = tmp_path / "cli.tsv", tmp_path / "api.tsv"
    assert cli.main(["search", "--genome", str(npz), "--guides", str(tsv), "--max-mismatches", "5",
                     "--n-top", "4", "--output", str(out_cli)]) == 0
    df = make_finder(genome_path=str(npz), max_mismatches=5, n_top_offtargets=4).search(
        input_file=str(tsv), output_file=str(out_api), verbose=False)
    assert len(df) > 0
    assert out_cli.read_text() == out_api.read_text()
    back = pd.read_csv(out_cli, sep="\t")
    assert list(back.columns) == list(df.columns) and len(back) == len(df)
    # fast mode over the same artefact as its index
    out_fast, out_fapi = tmp_path / "fast.tsv", tmp_path / "fapi.tsv"
    assert cli.main(["search", "--mode", "fast", "--index", str(npz), "--guides", str(tsv),
                     "--search-cutoff", "5", "--output", str(out_fast)]) == 0
    make_finder(mode="fast", index_path=str(npz)).search(input_file=str(tsv),
                                                          output_file=str(out_fapi), verbose=False)
    assert out_fast.read_text() == out_fapi.read_text()
