"""Full-size properties (3,088,269,832 bp hg38-like genome, both strands).

At this size the oracle cannot rescan everything for every guide, so the tests
combine:
* every planted near-copy is found at its position, strand and mismatch count
  (128 NGG guides at <= 4 mm; every 8th mixed guide at <= 5 mm);
* the sites at <= k mismatches are exactly the sites at <= k + 1 with n_mm <= k;
* the three compare engines and the plain drop-in kernel give the same set;
* a whole-genome oracle scan of a 10-guide subset (two guides of each PAM
  family and two A-rich guides, <= 5 mm) equals the plain kernel and every
  batched engine;
* oracle rescans of position slices (genome start and end, N-block edges,
  every second contig junction) match the sites of the full scan in them;
* with 8192 guides the VALU, int8, fp4 and seed engines agree, and an oracle
  slice matches;
* serving on the resident genome: Poisson arrivals of config-3 requests, each
  request's records equal a scan of it alone and the oracle on a slice.
"""
import sys
import time

import numpy as np
import pytest

from beguidedoff_amd import synth
from beguidedoff_amd.batched import DeviceGenome, GuideSpec, scan
from test_gpu_kernels import hit_set, oracle_set, plain_set

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[fullsize] {msg}", file=sys.stderr, flush=True)


def specs_of(guides, max_mm):
    return [GuideSpec(g.guide_seq, g.pam, max_mm, g.editor_type,
                      (g.edit_window_start, g.edit_window_end)) for g in guides]


def packed(hits):
    """Sorted uint64 keys of (guide, strand, pos, n_mm): set equality of large
    hit lists without Python tuples."""
    k = ((hits["guide"].astype(np.uint64) << np.uint64(38))
         | (hits["strand"].astype(np.uint64) << np.uint64(37))
         | (hits["pos"].astype(np.uint64) << np.uint64(5)) | hits["n_mm"].astype(np.uint64))
    k.sort()
    return k


def in_range(hits, b, e):
    return hits[(hits["pos"] >= b) & (hits["pos"] < e)]


def _device(w):
    g = w.genome
    return DeviceGenome(g, g["n_mask_le_fwd"], g["n_mask_le_rc"])


@pytest.fixture(scope="module")
def mixed(gpu):
    t = time.perf_counter()
    w = synth.workload("hg38_mixed")
    dg = _device(w)
    say(f"hg38_mixed resident in {time.perf_counter() - t:.1f} s")
    yield w, dg
    dg.close()


@pytest.fixture(scope="module")
def mixed_5mm(mixed):
    """Config 3 at its stated <= 5 mm: the full fp4 scan."""
    w, dg = mixed
    r = scan(dg, specs_of(w.guides, 5), tuning={"engine": 3})
    say(f"config 3 <= 5 mm: {r.n_hits} sites, kernel {r.kernel_ms:.2f} ms")
    return r


@pytest.fixture(scope="module")
def planted(gpu):
    w = synth.workload("hg38_ngg")
    dg = _device(w)
    yield w, dg
    dg.close()


def _missing_plants(g, plants, guides, hits):
    got = hit_set(hits)
    offsets, lengths = g["chrom_offsets"], g["chrom_lengths"]
    missing = []
    for p in plants:
        win = 20 + len(guides[p.guide].pam)
        if p.strand == "+":
            key = (p.guide, 0, p.pos, p.n_mm)
        else:  # the reverse-complement array is reversed per contig
            k = int(np.searchsorted(offsets, p.pos, side="right") - 1)
            off, clen = int(offsets[k]), int(lengths[k])
            key = (p.guide, 1, off + clen - (p.pos - off) - win, p.n_mm)
        if key not in got:
            missing.append((p, key))
    return missing


def test_planted_sites_found_with_their_mismatch_counts(planted):
    w, dg = planted
    g = w.genome
    assert g["total_bp"] == 3_088_269_832
    r = scan(dg, specs_of(w.guides, 4), tuning={"engine": 3})
    assert len(g["plants"]) == 128 * 40
    missing = _missing_plants(g, g["plants"], w.guides, r.hits)
    assert not missing, missing[:3]


def test_plain_kernel_equals_batched_at_full_size(planted):
    w, dg = planted
    guides = [(x.guide_seq, x.pam) for x in w.guides]
    want = hit_set(scan(dg, specs_of(w.guides, 4), tuning={"engine": 1}).hits)
    assert plain_set(w.genome, guides, 4) == want


def test_budget_threshold_is_exact(mixed, mixed_5mm):
    w, dg = mixed
    s5 = mixed_5mm.hits
    s4 = scan(dg, specs_of(w.guides, 4), tuning={"engine": 3}).hits
    s3 = scan(dg, specs_of(w.guides, 3), tuning={"engine": 3}).hits
    assert np.array_equal(packed(s4), packed(s5[s5["n_mm"] <= 4]))
    assert np.array_equal(packed(s3), packed(s4[s4["n_mm"] <= 3]))
    assert s5.size > s4.size > s3.size > 0


def test_engines_agree_at_full_size(mixed):
    w, dg = mixed
    specs = specs_of(w.guides, 4)
    sets = [packed(scan(dg, specs, tuning={"engine": e}).hits) for e in (1, 2, 3, 4)]
    assert all(np.array_equal(sets[0], x) for x in sets[1:])
    assert sets[0].size > 100_000


def test_config3_5mm_engines_agree(mixed, mixed_5mm):
    """Config 3 at <= 5 mm (about 1.55 M sites, the LDS record stage and fp4
    biases up to +5): VALU and int8 equal fp4 in every record field."""
    w, dg = mixed
    specs = specs_of(w.guides, 5)
    ref = mixed_5mm.hits
    assert ref.size > 1_000_000

    fields = ("guide", "strand", "pos", "n_mm", "x0", "x1", "mm_mask", "editable", "score")

    def full(h):
        h = h[np.lexsort((h["pos"], h["strand"], h["guide"]))]
        return [h[f] for f in fields]

    want = full(ref)
    for e in (1, 2):
        got = full(scan(dg, specs, tuning={"engine": e}).hits)
        for f, a, b in zip(fields, got, want):
            assert np.array_equal(a, b), (e, f)


def _slices(g, span):
    """Position ranges [b, e): genome start and end, both edges of the first
    N block and of one in contig 12, and every second contig junction."""
    L = int(g["total_bp"])
    offs = [int(x) for x in g["chrom_offsets"]]
    n_s, n_e = g["n_runs"]
    mid = int(np.searchsorted(n_s, offs[12]))
    out = [("start", 0, span), ("end", L - 26 - span, L - 26)]
    for j, tag in ((0, "first"), (mid, "contig12")):
        for edge, x in (("begin", int(n_s[j])), ("end", int(n_e[j]))):
            out.append((f"N-{tag}-{edge}", max(0, x - span // 2), x + span // 2))
    for k in range(2, len(offs), 2):
        out.append((f"junction{k}", offs[k] - span // 2, offs[k] + span // 2))
    return out


def test_config3_5mm_oracle_slices(mixed, mixed_5mm):
    """The oracle's sites in each slice equal the full scan's sites there, on
    both strand arrays (coordinates of the scanned array)."""
    w, dg = mixed
    g = w.genome
    guides = [(x.guide_seq, x.pam) for x in w.guides]
    hits = mixed_5mm.hits
    total = 0
    for name, b, e in _slices(g, 1_000_000):
        want = oracle_set(g, guides, 5, pos=(b, e))
        got = hit_set(in_range(hits, b, e))
        assert got == want, name
        total += len(want)
    say(f"oracle slices: {total} sites")
    assert total > 5000


def test_config3_5mm_whole_genome_oracle_subset(mixed):
    """Two guides of each PAM family and two A-rich guides, <= 5 mm, over the
    whole genome: the oracle equals the plain kernel and the VALU, int8 and fp4
    engines; at <= 4 mm, the oracle's sites with n_mm <= 4 equal every engine's,
    the seed engine's included."""
    w, dg = mixed
    g = w.genome
    pick = []
    for pam in ("NGG", "NNNRRT", "NNGRRT", "NG"):
        pick += [x for x in w.guides if x.pam == pam][:2]
    guides = [(x.guide_seq, x.pam) for x in pick] + [("A" * 20, "NG"), ("A" * 21, "NNNRRT")]
    t = time.perf_counter()
    want = oracle_set(g, guides, 5)
    say(f"whole-genome oracle, 10 guides: {len(want)} sites in {time.perf_counter() - t:.1f} s")
    assert len(want) > 20_000
    specs = [GuideSpec(s, p, 5) for s, p in guides]
    for e in (1, 2, 3):
        assert hit_set(scan(dg, specs, tuning={"engine": e}).hits) == want, e
    assert plain_set(g, guides, 5, rpg=200_000) == want
    want4 = {h for h in want if h[3] <= 4}
    specs4 = [GuideSpec(s, p, 4) for s, p in guides]
    for e in (1, 3, 4):
        assert hit_set(scan(dg, specs4, tuning={"engine": e}).hits) == want4, e


def test_config3_5mm_plants_found(gpu):
    w = synth.workload("hg38_mixed_planted")
    g = w.genome
    dg = _device(w)
    try:
        assert len(g["plants"]) == 128 * 8
        r = scan(dg, specs_of(w.guides, 5), tuning={"engine": 3})
        missing = _missing_plants(g, g["plants"], w.guides, r.hits)
        assert not missing, missing[:3]
        assert {p.n_mm for p in g["plants"]} == set(range(6))
    finally:
        dg.close()


def test_8192_guides_engines_agree(mixed):
    """Config 4: 8192 NGG guides at <= 4 mm; the VALU compare, the int8
    one-hot MFMA contraction, fp4 and the seed engine (two 4,096-guide
    tables) give the same sites, which equal the oracle on a slice across a
    contig junction; the cost model picks the seed engine."""
    w, dg = mixed
    g = w.genome
    seqs = synth.random_guides(np.random.default_rng(0), 8192)
    specs = [GuideSpec(s, "NGG", 4) for s in seqs]
    res = [scan(dg, specs, tuning={"engine": e}).hits for e in (1, 2, 3, 4)]
    sets = [packed(h) for h in res]
    assert all(np.array_equal(sets[0], x) for x in sets[1:])
    from beguidedoff_amd.batched import Plan
    p = Plan(dg, specs, hit_capacity=1 << 10)
    assert p.tuning()["engine"] == 4
    p.close()
    assert sets[0].size > 500_000
    b = int(g["chrom_offsets"][5]) - 250_000
    e = b + 500_000
    want = oracle_set(g, [(s, "NGG") for s in seqs], 4, pos=(b, e))
    assert hit_set(in_range(res[1], b, e)) == want and len(want) > 20


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


def test_serving_on_the_resident_genome(mixed):
    """Config 5: Poisson arrivals (32 req/s) of requests of log-uniform size
    1-500 drawn from the config-3 mix at <= 4 mm, served with continuous
    batching over 8 segments.  Each request's records (every field) equal a
    scan of that request alone on the VALU engine, and its sites in a slice
    across a segment boundary equal the oracle's."""
    from beguidedoff_amd.serving import ScanServer, latency_stats
    from test_gpu_serving import as_set

    w, dg = mixed
    g = w.genome
    rng = np.random.default_rng(5)
    reqs = []
    for _ in range(24):
        n = int(np.exp(rng.uniform(0, np.log(500))))
        idx = rng.choice(len(w.guides), size=n, replace=False)
        reqs.append(specs_of([w.guides[int(j)] for j in idx], 4))
    with ScanServer(dg, n_segments=8, hit_budget=1 << 24, pipeline=2) as srv:
        handles = []
        for specs in reqs:
            time.sleep(rng.exponential(1 / 32))
            handles.append(srv.submit(specs))
        got = [h.result(timeout=120) for h in handles]
        stats = latency_stats(handles)
        bounds = srv.bounds
    say(f"served {stats}")
    overlapped = sum(1 for i, a in enumerate(handles) for b2 in handles[i + 1:]
                     if b2.admitted_at < a.done_at)
    assert overlapped > 0  # requests joined while others were on the GPU
    b = bounds[3][0] - 500_000
    e = b + 1_000_000
    n_sites = 0
    for specs, h in zip(reqs, got):
        alone = scan(dg, specs, tuning={"engine": 1}).hits
        assert as_set(h, specs) == as_set(alone, specs)
        want = oracle_set(g, [(s.seq, s.pam) for s in specs], 4, pos=(b, e))
        assert hit_set(in_range(h, b, e)) == want
        n_sites += h.size
    assert n_sites > 10_000
