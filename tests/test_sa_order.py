"""bgo_sa_text_select (host code of libbgo_hip.so, no GPU) against the
brute-force suffix sort of the reference's index text (oracle/sa_order.py):
the anchor-occurrence cap of the reference's fast mode keeps the first
max_hits entries of the anchor's suffix-array interval
(csrc/fmindex_binding.cpp:164-176)."""
import ctypes as C

import numpy as np
import pytest

from beguidedoff_amd import _lib, synth
from oracle.sa_order import capped_occurrences, index_text

CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


class SaText:
    def __init__(self, contigs):
        g = synth.genome_from_strings(contigs)
        self.words = np.ascontiguousarray(g["genome_forward"], dtype=np.uint32)
        self.offs = np.ascontiguousarray(g["chrom_offsets"], dtype=np.int64)
        self.lens = np.ascontiguousarray(g["chrom_lengths"], dtype=np.int64)
        self.total = int(g["total_bp"])
        h = C.c_void_p()
        _lib.check(_lib.load().bgo_sa_text_create(
            _lib.ptr(self.words), self.words.size, self.total, _lib.ptr(self.offs),
            _lib.ptr(self.lens), self.offs.size, C.byref(h)))
        self.h = h

    def long_runs(self):
        n = C.c_int64(0)
        _lib.check(_lib.load().bgo_sa_text_long_runs(self.h, C.byref(n)))
        return n.value

    def select(self, anchor, cap, positions=None):
        pos = (np.arange(self.total, dtype=np.uint32) if positions is None
               else np.ascontiguousarray(positions, dtype=np.uint32))
        codes = np.array([CODE.get(c, 0) for c in anchor.upper()], dtype=np.uint8)
        out = np.zeros(cap, dtype=np.uint32)
        n_occ, n = C.c_int64(0), C.c_int64(0)
        _lib.check(_lib.load().bgo_sa_text_select(self.h, _lib.ptr(pos), pos.size, _lib.ptr(codes),
                                                  codes.size, cap, _lib.ptr(out), C.byref(n_occ),
                                                  C.byref(n)))
        return out[:n.value].tolist(), n_occ.value

    def close(self):
        _lib.load().bgo_sa_text_destroy(self.h)


def test_known_answer_runs_and_separator():
    """Text 1111 2 6 111111 3 0: occurrences of AAA in SA order are text 6,
    7, 0, 8, 1, 9 (longer A-runs first; at 4 A's, ...AC before ...AG), i.e.
    genome positions 5, 6, 0, 7, 1, 8."""
    contigs = [("a", "AAAAC"), ("b", "AAAAAAG")]
    text, offs = index_text(contigs)
    assert list(text) == [1, 1, 1, 1, 2, 6, 1, 1, 1, 1, 1, 1, 3, 0] and offs == [0, 6]
    assert capped_occurrences(contigs, "AAA", 3) == ([5, 6, 0], 6)
    assert capped_occurrences(contigs, "AAA", 0) == ([5, 6, 0, 7, 1, 8], 6)
    t = SaText(contigs)
    for cap, want in ((1, [5]), (3, [0, 5, 6]), (4, [0, 5, 6, 7]), (6, [0, 1, 5, 6, 7, 8])):
        assert t.select("AAA", cap) == (want, 6)
    t.close()


def test_known_answer_terminal_run():
    """A run that ends the text is followed by the terminal 0 (< A): there
    shorter runs come first.  Text 2 1111 0: AA at 1, 2, 3 -> suffixes
    1110 (1), 110 (2), 10 (3); SA order 3, 2, 1."""
    contigs = [("a", "CAAAA")]
    assert capped_occurrences(contigs, "AA", 2) == ([3, 2], 3)
    t = SaText(contigs)
    assert t.select("AA", 2) == ([2, 3], 3)
    t.close()


def random_genome(rng, n_contigs, lo, hi, long_runs=True):
    contigs = []
    for k in range(n_contigs):
        parts = []
        for _ in range(int(rng.integers(2, 6))):
            parts.append("".join(rng.choice(list("ACGT"), size=int(rng.integers(lo, hi)),
                                            p=[0.295, 0.205, 0.205, 0.295])))
            if long_runs:
                kind = int(rng.integers(0, 5))
                n = int(rng.integers(10, 200))
                parts.append(["N" * n, "A" * n, "T" * n, "AC" * (n // 2), "ACGTTGCA" * (n // 8 + 1)][kind])
        contigs.append((f"c{k}", "".join(parts)))
    return contigs


@pytest.mark.parametrize("seed", range(8))
def test_select_equals_brute_force(seed):
    rng = np.random.default_rng(seed)
    contigs = random_genome(rng, int(rng.integers(1, 5)), 30, 400)
    full = "".join(s for _, s in contigs).upper().replace("N", "A")
    t = SaText(contigs)
    assert t.long_runs() >= 0
    anchors = ["A" * 12, "A" * 10, "A" * 14, "ACACACACACAC", "ACGTTGCAACGT", "TTTTTTTTTTTT"]
    for _ in range(6):  # substrings of the genome, some near contig ends
        m = int(rng.choice([10, 12, 14]))
        i = int(rng.integers(0, len(full) - m))
        anchors.append(full[i:i + m])
    for a in anchors:
        _, total = capped_occurrences(contigs, a, 0)
        for cap in sorted({1, 2, 3, max(1, total // 3), max(1, total - 1), total + 5}):
            want, n_occ = capped_occurrences(contigs, a, cap)
            got, got_occ = t.select(a, cap)
            assert got_occ == n_occ, (a, cap)
            assert got == sorted(want), (a, cap)
    t.close()


def test_long_runs_table_and_candidate_subsets():
    """Runs of >= 64 equal bases (N blocks read as A, homopolymers) are in the
    run table; the selection only considers the positions it is given."""
    contigs = [("a", "CG" * 40 + "N" * 300 + "CATG" + "A" * 70 + "G"), ("b", "T" * 90 + "A" * 65)]
    t = SaText(contigs)
    assert t.long_runs() == 4
    for cap in (1, 5, 50, 200, 400):
        want, n = capped_occurrences(contigs, "A" * 12, cap)
        assert t.select("A" * 12, cap) == (sorted(want), n)
    sub = np.arange(0, 400, 3, dtype=np.uint32)
    got, n_sub = t.select("A" * 12, 10, positions=sub)
    _, n_all = capped_occurrences(contigs, "A" * 12, 0)
    assert 0 < n_sub < n_all and len(got) == 10 and set(got) <= set(sub.tolist())
    t.close()
