"""The drop-in driver's upload cache (beguidedoff/cuda_search.py:321-362):
eviction drops the cache's references only, so a strand that a caller still
holds is never released under it.  Runs without a GPU: the device strand is
replaced by a stand-in that records when it is released."""
import gc
import threading

import numpy as np

from beguidedoff_amd import cuda_search


class FakeStrand:
    released: list = []

    def __init__(self, genome_packed, total_bp, n_mask_packed=None):
        self.total_bp = int(total_bp)
        self.tag = int(genome_packed[0])
        self.closed = False

    def close(self):
        if not self.closed:
            self.closed = True
            FakeStrand.released.append(self.tag)

    def __del__(self):
        self.close()


def test_eviction_never_releases_a_strand_in_use(monkeypatch):
    monkeypatch.setattr(cuda_search, "DeviceStrand", FakeStrand)
    monkeypatch.setattr(cuda_search, "_CACHE", {})
    FakeStrand.released = []
    arrays = [np.full(8, i, dtype=np.uint32) for i in range(10)]
    held = cuda_search._device_strand(arrays[0], 100, None)   # a call in flight holds it
    assert cuda_search._device_strand(arrays[0], 100, None) is held
    for a in arrays[1:]:  # fills the cache past its limit twice
        cuda_search._device_strand(a, 100, None)
    gc.collect()
    assert not held.closed and 0 not in FakeStrand.released
    assert len(cuda_search._CACHE) <= cuda_search._CACHE_LIMIT
    # the strands evicted with no other holder were released
    assert set(FakeStrand.released) >= {1, 2, 3, 4}
    del held
    gc.collect()
    assert 0 in FakeStrand.released


def test_clear_device_cache_leaves_held_strands_alive(monkeypatch):
    monkeypatch.setattr(cuda_search, "DeviceStrand", FakeStrand)
    monkeypatch.setattr(cuda_search, "_CACHE", {})
    FakeStrand.released = []
    a, b = np.full(4, 7, dtype=np.uint32), np.full(4, 8, dtype=np.uint32)
    held = cuda_search._device_strand(a, 50, None)
    cuda_search._device_strand(b, 50, None)
    cuda_search.clear_device_cache()
    gc.collect()
    assert FakeStrand.released == [8] and not held.closed
    del held
    gc.collect()
    assert FakeStrand.released == [8, 7]


def test_concurrent_callers_keep_their_strands(monkeypatch):
    """Threads that take strands from the cache while others evict it never
    see their own strand released before they drop it."""
    monkeypatch.setattr(cuda_search, "DeviceStrand", FakeStrand)
    monkeypatch.setattr(cuda_search, "_CACHE", {})
    FakeStrand.released = []
    arrays = [np.full(4, i, dtype=np.uint32) for i in range(32)]
    bad = []

    def worker(k):
        rng = np.random.default_rng(k)
        for _ in range(200):
            s = cuda_search._device_strand(arrays[int(rng.integers(0, 32))], 10, None)
            for _ in range(3):  # the "scan": the strand must stay open
                if s.closed:
                    bad.append(s.tag)
            del s

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(8)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not bad
