"""batched.scan_capped's host logic (group splits, per-strand range walks,
truncation to the lowest positions, exact counts, the record limit) on the
CPU: the device plan is replaced by a stand-in that serves a fixed site list
with the kernel's capacity rule (every site counted, the first `capacity`
stored)."""
import numpy as np
import pytest

from beguidedoff_amd import batched as B


class FakeGenome:
    total_bp = 100_000


def make_sites(seed=0):
    """{guide index: [(strand, pos)]}: guide 2 is a heavy multi-mapper."""
    rng = np.random.default_rng(seed)
    sites = {}
    for g, n in enumerate([5, 40, 9000, 0, 300, 1200, 17]):
        sites[g] = [(int(rng.integers(0, 2)), int(p))
                    for p in rng.choice(FakeGenome.total_bp - 26, size=n, replace=False)]
    return sites


def fake_plan_factory(sites, fetched, launches):
    class FakePlan:
        def __init__(self, genome, specs, tuning=None, hit_capacity=1 << 20):
            self.specs = list(specs)
            self.hit_capacity = int(hit_capacity)
            self._last = None

        def launch(self, b=0, e=None, strands=3):
            e = FakeGenome.total_bp if e is None else e
            e = min(e, FakeGenome.total_bp - 26)
            recs = []
            for i, sp in enumerate(self.specs):
                for s, p in sites[sp.max_mm]:   # max_mm carries the guide's key
                    if strands & (1 << s) and b <= p < e:
                        recs.append((i, p, s))
            rng = np.random.default_rng(len(recs))
            order = rng.permutation(len(recs))  # atomic order: arbitrary
            self._last = [recs[k] for k in order]
            launches.append((b, e, strands, len(self.specs)))

        def wait(self):
            return len(self._last), 0.1

        def set_hit_capacity(self, cap):
            self.hit_capacity = int(cap)

        def fetch(self, n=None):
            m = min(len(self._last), self.hit_capacity, n if n is not None else 1 << 62)
            out = np.zeros(m, dtype=B.HIT_DTYPE)
            for k, (i, p, s) in enumerate(self._last[:m]):
                out[k]["guide"], out[k]["pos"], out[k]["strand"] = i, p, s
            fetched.append(m)
            return out

        def close(self):
            pass

    return FakePlan


@pytest.mark.parametrize("cap,limit", [(10**9, 1 << 24), (100, 1024), (7, 1024), (1000, 1500)])
def test_scan_capped_host_logic(monkeypatch, cap, limit):
    sites = make_sites()
    fetched, launches = [], []
    monkeypatch.setattr(B, "Plan", fake_plan_factory(sites, fetched, launches))
    specs = [B.GuideSpec("A" * 20, "NGG", g) for g in sites]  # max_mm = key
    r = B.scan_capped(FakeGenome(), specs, cap, max_records=limit)
    assert max(fetched) <= max(limit, 1024)
    want_counts = np.zeros((len(specs), 2), np.int64)
    want = set()
    for g, lst in sites.items():
        for s in (0, 1):
            ps = sorted(p for ss, p in lst if ss == s)
            want_counts[g, s] = len(ps)
            want |= {(g, s, p) for p in ps[:cap]}
    assert np.array_equal(r.counts, want_counts)
    got = set(zip(r.hits["guide"].tolist(), r.hits["strand"].tolist(), r.hits["pos"].tolist()))
    assert got == want and len(got) == r.hits.size
    assert r.truncated.tolist() == [bool((want_counts[g] > cap).any()) for g in range(len(specs))]
    if limit < 9000:  # the heavy guide was scanned alone, per strand, over ranges
        assert any(n == 1 and st in (1, 2) and e - b < FakeGenome.total_bp
                   for b, e, st, n in launches)
