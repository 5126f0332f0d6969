"""Serving: requests admitted mid-scan, results identical alone and in the mix."""
import time

import numpy as np
import pytest

from beguidedoff_amd import synth
from beguidedoff_amd.batched import DeviceGenome, GuideSpec, scan
from beguidedoff_amd.serving import ScanServer, latency_stats
from test_gpu_kernels import hit_set, oracle_set

pytestmark = pytest.mark.gpu

KEYS = ["strand", "pos", "n_mm", "mm_mask", "editable", "x0", "x1"]


def as_set(h, specs):
    return {(specs[int(r["guide"])].seq, specs[int(r["guide"])].pam, int(r["guide"]))
            + tuple(int(r[k]) for k in KEYS) + (float(r["score"]),) for r in h}


@pytest.fixture(scope="module")
def served_genome(gpu):
    rng = np.random.default_rng(4)
    contigs = [("a", 3_000_000), ("b", 2_000_000)]
    names, offs, lens = synth.contig_layout(contigs)
    runs = synth.place_n_runs(rng, offs, lens, 0.02, 1000, 20_000)
    pool = synth.mixed_guides(rng, 64, 32, 16, 16)
    plants = synth.plan_plants(rng, offs, lens, *runs, [(g.guide_seq, g.pam) for g in pool], 3, 4)
    g = synth.make_genome(contigs, seed=9, n_runs=runs, plants=plants)
    dg = DeviceGenome(g, g["n_mask_le_fwd"], g["n_mask_le_rc"])
    yield dg, pool, g
    dg.close()


@pytest.mark.parametrize("solo", [None, 20])
def test_results_same_alone_and_in_mix(served_genome, solo):
    """Each request's records equal a scan of it alone (every field) and its
    sites equal the oracle's (per-guide budgets); also with requests of 20
    guides or more in launches of their own."""
    dg, pool, g = served_genome
    rng = np.random.default_rng(1)
    reqs = []
    for i in range(12):
        n = int(rng.integers(1, 40))
        idx = rng.choice(len(pool), size=n, replace=False)
        specs = [GuideSpec(pool[j].guide_seq, pool[j].pam, int(rng.integers(2, 6)),
                           pool[j].editor_type, (pool[j].edit_window_start, pool[j].edit_window_end))
                 for j in idx]
        reqs.append(specs)
    assert any(len(r) >= 20 for r in reqs) and any(len(r) < 20 for r in reqs)
    with ScanServer(dg, n_segments=5, hit_budget=1 << 20, pipeline=2, solo_guides=solo) as srv:
        handles = []
        for specs in reqs:
            handles.append(srv.submit(specs))
            time.sleep(0.002)            # arrive while earlier requests are on the GPU
        got = [h.result(timeout=60) for h in handles]
        stats = latency_stats(handles)
        assert srv.launches >= 5
    assert stats["n"] == len(reqs)
    total = 0
    for specs, h in zip(reqs, got):
        alone = scan(dg, specs).hits
        assert as_set(h, specs) == as_set(alone, specs)
        want = oracle_set(g, [(s.seq, s.pam) for s in specs], None,
                          mm_per=[s.max_mm for s in specs])
        assert hit_set(h) == want
        total += alone.size
    assert total > 0


def test_admission_respects_budget_and_reruns(served_genome):
    dg, pool, g = served_genome
    specs = [GuideSpec(g.guide_seq, g.pam, 5) for g in pool[:20]]
    with ScanServer(dg, n_segments=3, hit_budget=2048, pipeline=1, min_reserve=2048) as srv:
        a = srv.submit(specs)
        b = srv.submit(specs[:5])
        ra, rb = a.result(60), b.result(60)
        # b could only be admitted after a released its reservation
        assert b.admitted_at >= a.done_at
    assert as_set(ra, specs) == as_set(scan(dg, specs).hits, specs)
    assert as_set(rb, specs[:5]) == as_set(scan(dg, specs[:5]).hits, specs[:5])
    assert hit_set(ra) == oracle_set(g, [(s.seq, s.pam) for s in specs], 5)
