"""Full-size properties (3,088,269,832 bp hg38-like genome, both strands).

At this size the oracle cannot rescan everything, so the tests check what
holds independently of size:
* every planted near-copy is found at its position, strand and mismatch count;
* the sites at <= k mismatches are exactly the sites at <= k + 1 with n_mm <= k;
* the three compare engines and the plain drop-in kernel give the same set;
* oracle rescans of position slices at the high end of the 32-bit range and
  across a contig junction match the batched scan of the same slices;
* with 8192 guides the VALU and matrix-core engines agree.
"""
import numpy as np
import pytest

from beguidedoff_amd import synth
from beguidedoff_amd.batched import DeviceGenome, GuideSpec, scan
from test_gpu_kernels import hit_set, oracle_set, plain_set

pytestmark = pytest.mark.gpu


def specs_of(guides, max_mm):
    return [GuideSpec(g.guide_seq, g.pam, max_mm, g.editor_type,
                      (g.edit_window_start, g.edit_window_end)) for g in guides]


@pytest.fixture(scope="module")
def mixed(gpu):
    w = synth.workload("hg38_mixed")
    g = w.genome
    dg = DeviceGenome(g, g["n_mask_le_fwd"], g["n_mask_le_rc"])
    yield w, dg
    dg.close()


@pytest.fixture(scope="module")
def planted(gpu):
    w = synth.workload("hg38_ngg")
    g = w.genome
    dg = DeviceGenome(g, g["n_mask_le_fwd"], g["n_mask_le_rc"])
    yield w, dg
    dg.close()


def test_planted_sites_found_with_their_mismatch_counts(planted):
    w, dg = planted
    g = w.genome
    assert g["total_bp"] == 3_088_269_832
    r = scan(dg, specs_of(w.guides, 4), tuning={"engine": 3})
    got = hit_set(r.hits)
    offsets, lengths = g["chrom_offsets"], g["chrom_lengths"]
    win = 23
    missing = []
    for p in g["plants"]:
        if p.strand == "+":
            key = (p.guide, 0, p.pos, p.n_mm)
        else:  # the reverse-complement array is reversed per contig
            k = int(np.searchsorted(offsets, p.pos, side="right") - 1)
            off, clen = int(offsets[k]), int(lengths[k])
            key = (p.guide, 1, off + clen - (p.pos - off) - win, p.n_mm)
        if key not in got:
            missing.append((p, key))
    assert len(g["plants"]) == 128 * 40
    assert not missing, missing[:3]


def test_plain_kernel_equals_batched_at_full_size(planted):
    w, dg = planted
    guides = [(x.guide_seq, x.pam) for x in w.guides]
    want = hit_set(scan(dg, specs_of(w.guides, 4), tuning={"engine": 1}).hits)
    assert plain_set(w.genome, guides, 4) == want


def test_budget_threshold_is_exact(mixed):
    w, dg = mixed
    s4 = hit_set(scan(dg, specs_of(w.guides, 4), tuning={"engine": 3}).hits)
    s3 = hit_set(scan(dg, specs_of(w.guides, 3), tuning={"engine": 3}).hits)
    assert s3 == {h for h in s4 if h[3] <= 3}
    assert len(s4) > len(s3) > 0


def test_engines_agree_at_full_size(mixed):
    w, dg = mixed
    specs = specs_of(w.guides, 4)
    sets = [hit_set(scan(dg, specs, tuning={"engine": e}).hits) for e in (1, 2, 3)]
    assert sets[0] == sets[1] == sets[2]
    assert len(sets[0]) > 100_000


@pytest.mark.parametrize("where", ["end", "junction"])
def test_oracle_slice_matches(mixed, where):
    w, dg = mixed
    g = w.genome
    guides = [(x.guide_seq, x.pam) for x in w.guides]
    span = 1_000_000
    if where == "end":
        b = int(g["total_bp"]) - 26 - span
    else:  # straddle the start of contig 12
        b = int(g["chrom_offsets"][12]) - span // 2
    e = b + span
    want = oracle_set(g, guides, 4, pos=(b, e))
    r = scan(dg, specs_of(w.guides, 4), pos_begin=b, pos_end=e, tuning={"engine": 3})
    assert hit_set(r.hits) == want
    assert len(want) > 0


def test_8192_guides_engines_agree(mixed):
    w, dg = mixed
    seqs = synth.random_guides(np.random.default_rng(0), 8192)
    specs = [GuideSpec(s, "NGG", 4) for s in seqs]
    sets = [hit_set(scan(dg, specs, tuning={"engine": e}).hits) for e in (1, 3)]
    assert sets[0] == sets[1]
    assert len(sets[0]) > 10_000


def test_small_sets_direct_compare_equals_buffer(mixed, planted):
    """Small PAM families compare their windows in place (VALU engine): with
    6 guides per family the dense PAMs (NG, NNNRRT) do and the sparse ones
    (NGG, NNGRRT) use the window buffer.  The sites equal those of the
    all-buffer path and of fp4, and a direct scan of 4 planted guides finds
    every one of their plants."""
    w, dg = mixed
    by_pam = {}
    for x in w.guides:
        by_pam.setdefault(x.pam, []).append(x)
    specs = specs_of([x for xs in by_pam.values() for x in xs[:6]], 5)
    r = scan(dg, specs, tuning={"engine": 1})
    assert r.tuning["direct_max"] == 0 and all(n == 6 for n, _ in r.families)
    direct = hit_set(r.hits)
    assert direct == hit_set(scan(dg, specs, tuning={"engine": 1, "direct_max": -1}).hits)
    assert direct == hit_set(scan(dg, specs, tuning={"engine": 3}).hits)
    assert len(direct) > 1000
    w, dg = planted
    sub = hit_set(scan(dg, specs_of(w.guides[:4], 4), tuning={"engine": 1}).hits)
    full = hit_set(scan(dg, specs_of(w.guides, 4), tuning={"engine": 3}).hits)
    assert sub == {h for h in full if h[0] < 4} and len(sub) >= 4 * 40
