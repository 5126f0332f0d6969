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


def test_build_index_writes_the_artefact_and_the_reference_sidecars(tmp_path):
    """fast/fmindex.py:63-128 and csrc/fmindex_binding.cpp:47-56: the index
    path holds the encoded genome (no extension added), ``.meta`` the
    chromosome table with separator-shifted offsets, ``.nmask`` the N runs
    (lower-case n included) in the same coordinates."""
    contigs = [("chr1", "ACGTNNACGTTTGCA" * 7), ("chrM", "nnnACGT"), ("chr2", "GGGCCCAAATTTNN" * 5)]
    fa = tmp_path / "g.fa"
    _fasta(fa, contigs)
    idx = tmp_path / "hg.fmi"
    assert cli.main(["build-index", "--fasta", str(fa), "--output", str(idx)]) == 0
    assert not (tmp_path / "hg.fmi.npz").exists()
    got = load_genome(str(idx))
    want = encode_contigs(contigs)
    for key in ("genome_forward", "genome_rc", "n_mask"):
        assert np.array_equal(got[key], want[key])
    # offsets: 105 bases of chr1, one separator, 7 of chrM, one separator
    assert (tmp_path / "hg.fmi.meta").read_text() == (
        "#beguidedoff-fmindex\t1\nchr1\t105\t0\nchrM\t7\t106\nchr2\t70\t114\n")
    runs = []
    off = 0
    for name, seq in contigs:
        i = 0
        while i < len(seq):
            if seq[i] in "Nn":
                j = i
                while j < len(seq) and seq[j] in "Nn":
                    j += 1
                runs.append(f"{off + i}\t{off + j}")
                i = j
            else:
                i += 1
        off += len(seq) + 1
    assert (tmp_path / "hg.fmi.nmask").read_text() == "".join(r + "\n" for r in runs)
    assert runs[:2] == ["4\t6", "19\t21"] and "106\t109" in runs


def test_loading_a_file_that_is_not_an_artefact_says_so(tmp_path):
    bad = tmp_path / "sdsl.fmi"
    bad.write_bytes(b"\x00\x01 not an npz" * 10)
    with pytest.raises(ValueError, match="not an encoded genome artefact"):
        load_genome(str(bad))


def test_encode_genome_to_a_name_without_npz(tmp_path):
    contigs = [("chr1", "ACGT" * 40)]
    fa = tmp_path / "g.fa"
    _fasta(fa, contigs)
    assert cli.main(["encode-genome", "--fasta", str(fa), "--output", str(tmp_path / "g.bin")]) == 0
    assert load_genome(str(tmp_path / "g.bin.npz"))["total_bp"] == 160


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
    # and over index files written by build-index under the reference's name
    fa, fmi, out_idx = tmp_path / "g.fa", tmp_path / "g.fmi", tmp_path / "idx.tsv"
    names = [str(n) for n in w.genome["chrom_names"]]
    offs, lens = w.genome["chrom_offsets"], w.genome["chrom_lengths"]
    L = int(w.genome["total_bp"])
    chars = np.frombuffer(b"ACGT", dtype=np.uint8)[synth.extract_codes(w.genome["genome_forward"], 0, L)]
    chars[np.unpackbits(w.genome["n_mask"])[:L].astype(bool)] = ord("N")
    text = chars.tobytes().decode("ascii")
    _fasta(fa, [(n, text[o:o + l]) for n, o, l in zip(names, offs.tolist(), lens.tolist())])
    assert cli.main(["build-index", "--fasta", str(fa), "--output", str(fmi)]) == 0
    assert cli.main(["search", "--mode", "fast", "--index", str(fmi), "--guides", str(tsv),
                     "--search-cutoff", "5", "--output", str(out_idx)]) == 0
    assert out_idx.read_text() == out_fast.read_text()
