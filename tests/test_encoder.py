"""Genome artefact encoders against the per-base oracle
(oracle/encode_oracle.py, which restates genome_converter.py:44-100): its
pins on the CPU, the host (numpy) encoder on the CPU, and the GPU encoder
(bgo_encode_genome) on MI355X, bit-exact."""
import ctypes as C

import numpy as np
import pytest

from beguidedoff_amd import _lib, cli
from beguidedoff_amd.genome_converter import encode_contigs, encode_hip, load_genome
from oracle.encode_oracle import encode as oracle_encode

KEYS = ("genome_forward", "genome_rc", "n_mask", "chrom_offsets", "chrom_lengths")


def random_text_contigs(rng, lengths, alphabet="ACGTacgtNnRYKMSWBDHV-*.", p_n_run=0.3):
    out = []
    for i, n in enumerate(lengths):
        s = rng.choice(list("ACGTacgt"), size=n)
        if n:
            k = rng.random(n) < 0.05  # scattered other symbols
            s[k] = rng.choice(list(alphabet), size=int(k.sum()))
            if rng.random() < p_n_run and n > 10:  # one N run
                a = int(rng.integers(0, n - 5))
                s[a:a + int(rng.integers(1, min(400, n - a)))] = "N"
        out.append((f"ctg{i}", "".join(s)))
    return out


def same(a, b):
    for k in KEYS:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    assert int(a["total_bp"]) == int(b["total_bp"])


def test_oracle_pins():
    # reference tests/test_genome_converter.py:13-22: two 12-bp contigs
    d = oracle_encode([("chr1", "ACGTACGTACGT"), ("chr2", "TGCATGCATGCA")])
    assert int(d["total_bp"]) == 24 and d["chrom_offsets"].tolist() == [0, 12]
    # closed forms: A C G T -> 0 1 2 3 at bits 2i; "ACGT" is its own reverse complement
    d = oracle_encode([("c", "ACGT")])
    assert d["genome_forward"].tolist() == [0b11100100] and d["genome_rc"].tolist() == [0b11100100]
    # N is coded A (0) forward and T (3) in the reverse complement; mask bit 7 first
    d = oracle_encode([("c", "ANNNNNNNG")])
    assert d["n_mask"].tolist() == [0x7F, 0x00]
    assert d["genome_forward"].tolist() == [2 << 16]
    assert d["genome_rc"].tolist() == [1 | sum(3 << (2 * i) for i in range(1, 9))]
    # per-contig reverse complement, lower case folded, other symbols -> 0
    d = oracle_encode([("a", "aC"), ("b", "gRt")])
    assert d["genome_forward"].tolist() == [0 | 1 << 2 | 2 << 4 | 0 << 6 | 3 << 8]
    assert d["genome_rc"].tolist() == [2 | 3 << 2 | 0 << 4 | 3 << 6 | 1 << 8]
    # 17 bases: two words, the second zero-padded
    d = oracle_encode([("c", "T" * 17)])
    assert d["genome_forward"].tolist() == [0xFFFFFFFF, 3]


@pytest.mark.parametrize("seed", range(4))
def test_numpy_encoder_equals_oracle(seed):
    rng = np.random.default_rng(seed)
    contigs = random_text_contigs(rng, [int(x) for x in rng.integers(0, 700, size=6)] + [0, 1, 16])
    same(encode_contigs(contigs), oracle_encode(contigs))


def test_encoder_rejects_a_contig_table_that_does_not_tile():
    """Argument checks run before any device call."""
    lib = _lib.load()
    seq = np.frombuffer(b"ACGTACGT", dtype=np.uint8)
    out = np.zeros(1, np.uint32)
    for offs, lens in (([0, 5], [4, 4]), ([0], [7]), ([1], [8])):
        o, l = np.array(offs, np.int64), np.array(lens, np.int64)
        rc = lib.bgo_encode_genome(0, _lib.ptr(seq), 8, _lib.ptr(o), _lib.ptr(l), len(offs),
                                   _lib.ptr(out), None, None, None)
        assert rc == -1
        with pytest.raises(RuntimeError, match="tile|sum"):
            _lib.check(rc)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_gpu_encoder_equals_oracle(gpu, seed):
    rng = np.random.default_rng(100 + seed)
    lengths = [0, 1, 15, 16, 17, 33] + [int(x) for x in rng.integers(0, 900, size=8)]
    rng.shuffle(lengths)
    contigs = random_text_contigs(rng, lengths)
    same(encode_contigs(contigs, engine="hip"), oracle_encode(contigs))


@pytest.mark.gpu
def test_gpu_encoder_equals_numpy_at_scale(gpu):
    """24 contigs, 20 Mbp of mixed-case text with IUPAC symbols and N runs."""
    rng = np.random.default_rng(7)
    lengths = [int(x) for x in rng.integers(1, 1_700_000, size=24)]
    contigs = random_text_contigs(rng, lengths, p_n_run=0.8)
    a, b = encode_contigs(contigs, engine="hip"), encode_contigs(contigs)
    same(a, b)
    assert list(a["chrom_names"]) == list(b["chrom_names"])
    raw = np.frombuffer("".join(s for _, s in contigs).encode(), dtype=np.uint8)
    fwd, rc, nm, ms = encode_hip(raw, a["chrom_offsets"], a["chrom_lengths"],
                                 include_rc=False, include_n_mask=False)
    assert rc is None and nm is None and np.array_equal(fwd, b["genome_forward"]) and ms > 0


@pytest.mark.gpu
def test_cli_encode_genome_engines_agree(gpu, tmp_path):
    rng = np.random.default_rng(3)
    contigs = random_text_contigs(rng, [5000, 1, 12345, 0, 777])
    fa = tmp_path / "g.fa"
    fa.write_text("".join(f">{n}\n{s}\n" for n, s in contigs))
    for eng in ("numpy", "hip"):
        assert cli.main(["encode-genome", "--fasta", str(fa), "--output",
                         str(tmp_path / f"{eng}.npz"), "--engine", eng]) == 0
    a, b = load_genome(str(tmp_path / "numpy.npz")), load_genome(str(tmp_path / "hip.npz"))
    same(a, b)
    same(a, oracle_encode(contigs))
