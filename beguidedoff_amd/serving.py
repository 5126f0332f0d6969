"""Continuous-batching scan server over a genome resident in HBM.

The genome's scan range [0, total_bp - 26) is cut into `n_segments` position
ranges (both strands each).  The server launches one segment after another,
cyclically, each launch covering the guides of every active request.  A
request is admitted at a segment boundary — as soon as hit-buffer memory
allows — and is part of the next `n_segments` launches, i.e. exactly one full
wrap of the genome starting wherever the scan currently is; then its sites
are complete and it retires, releasing its reservation.  Requests that arrive
while others are on the GPU therefore join the running scan instead of
waiting for it to finish.

Launches alternate between `pipeline` plans with their own HIP streams, so a
segment can be enqueued while the previous one still runs and the host-side
dispatch of finished segments overlaps device work.

Segment count: more segments shorten the wait to join but add a launch's
start-up and tail to every request's pass (about 0.015 ms to start a
persistent grid, plus the last work items).  On a 3.09 Gbp genome with
1-500-guide requests, same arrivals: p50 / p99 at 2 and 32 req/s were
4.05 / 11.48 and 4.90 / 17.77 ms with 4 segments, 4.27 / 11.09 and
4.79 / 16.53 ms with 8, 5.65 / 11.62 and 6.49 / 14.70 ms with 16.

Large requests (`solo_guides`, off by default): requests of at least that
many guides get launches of their own, taken in turn with the shared launch
of the smaller requests, so a small request's launches never carry a large
request's guides.  Each launch group keeps its own segment cursor.  With
300 on the same 3.09 Gbp workload p99 did not fall (10.53 ms either way at
2 req/s) and rose at 32 req/s (12.34 -> 15.92 ms): overlapping large
requests in separate launches pay the per-window work twice.

Admission: each request reserves an estimated number of hit records (from
its PAMs and mismatch budgets); requests are admitted in arrival order while
the sum of active reservations stays within `hit_budget`.  A segment whose
sites exceed its buffer is re-run with an exact-size buffer, so results never
depend on the estimate.  Per-guide results are independent of the other
guides in a launch, so a request's sites are the same alone and in any mix.
"""
from __future__ import annotations

import collections
import math
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from .batched import HIT_DTYPE, DeviceGenome, GuideSpec, Plan, guide_descs
from .pam import IUPAC_MASK


def _pam_pass_rate(pam: str) -> float:
    rate = 1.0
    for c in pam.upper():
        rate *= bin(IUPAC_MASK[c]).count("1") / 4.0
    return rate


def _p_le_k(k: int, n: int = 20, p_mm: float = 0.75) -> float:
    return sum(math.comb(n, j) * p_mm ** j * (1 - p_mm) ** (n - j) for j in range(0, k + 1))


def estimate_hits(specs: Sequence[GuideSpec], total_bp: int) -> float:
    """Expected sites of a guide set on a uniform random genome (both strands)."""
    per: Dict[tuple, float] = {}
    total = 0.0
    for s in specs:
        k = (s.pam, s.max_mm)
        v = per.get(k)
        if v is None:
            v = per[k] = 2.0 * total_bp * _pam_pass_rate(s.pam) * _p_le_k(s.max_mm)
        total += v
    return total


@dataclass
class Request:
    rid: int
    specs: List[GuideSpec]
    arrival: float
    reserve: int
    descs: np.ndarray = None
    launched: int = 0
    retired: int = 0
    admitted_at: float = 0.0
    seg: int = 0  # next segment of a request scanned in launches of its own
    done_at: float = 0.0
    parts: List[np.ndarray] = field(default_factory=list)
    done: threading.Event = field(default_factory=threading.Event)
    error: Optional[BaseException] = None

    def result(self, timeout: Optional[float] = None) -> np.ndarray:
        """Site records of this request (``guide`` indexes its own guide list)."""
        if not self.done.wait(timeout):
            raise TimeoutError(f"request {self.rid} not finished")
        if self.error is not None:
            raise RuntimeError(f"request {self.rid} failed") from self.error
        return np.concatenate(self.parts) if self.parts else np.zeros(0, dtype=HIT_DTYPE)

    @property
    def latency(self) -> float:
        return self.done_at - self.arrival


class ScanServer:
    def __init__(self, genome: DeviceGenome, n_segments: int = 8, hit_budget: int = 1 << 24,
                 pipeline: int = 2, tuning: Optional[dict] = None, min_reserve: int = 1024,
                 solo_guides: Optional[int] = None):
        if n_segments < 1 or pipeline < 1:
            raise ValueError("n_segments and pipeline must be >= 1")
        if solo_guides is not None and solo_guides < 1:
            raise ValueError("solo_guides must be >= 1")
        self.solo_guides = solo_guides
        self._turn = 0
        self.genome = genome
        self.n_segments = n_segments
        self.hit_budget = int(hit_budget)
        self.min_reserve = int(min_reserve)
        end = max(0, genome.total_bp - 26)
        self.bounds = [(end * k // n_segments, end * (k + 1) // n_segments)
                       for k in range(n_segments)]
        # 8 MB of hit records per plan to start; buffers grow at least 2x, since
        # each reallocation frees device memory, which waits for the other
        # plans' running segments
        self._plans = [Plan(genome, [], tuning=tuning, hit_capacity=1 << 18)
                       for _ in range(pipeline)]
        self._plan_set = {id(p): () for p in self._plans}  # request ids a plan's tables hold
        self._free = collections.deque(self._plans)
        self._queue: collections.deque = collections.deque()
        self._active: List[Request] = []
        self._reserved = 0
        self._cv = threading.Condition()
        self._stop = False
        self._rid = 0
        self._seg = 0
        self.launches = 0
        self.reruns = 0
        self.busy_s = 0.0
        self.completed: List[Request] = []
        self._thread = threading.Thread(target=self._run, name="bgo-scan-server", daemon=True)
        self._thread.start()

    # ------------------------------------------------------------ client

    def submit(self, specs: Sequence[GuideSpec]) -> Request:
        specs = list(specs)
        arrival = time.perf_counter()
        descs = guide_descs(specs)  # validates the guides on the caller's thread
        reserve = max(self.min_reserve, int(4 * estimate_hits(specs, self.genome.total_bp)) + 64)
        with self._cv:
            if self._stop:
                raise RuntimeError("server stopped")
            req = Request(self._rid, specs, arrival, min(reserve, self.hit_budget), descs)
            self._rid += 1
            self._queue.append(req)
            self._cv.notify_all()
        return req

    def close(self, timeout: Optional[float] = None) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        self._thread.join(timeout)
        for p in self._plans:
            p.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ loop

    def _admit(self) -> None:
        while self._queue and (not self._active or
                               self._reserved + self._queue[0].reserve <= self.hit_budget):
            req = self._queue.popleft()
            req.admitted_at = time.perf_counter()
            self._reserved += req.reserve
            self._active.append(req)

    def _run(self) -> None:
        inflight: collections.deque = collections.deque()
        try:
            while True:
                with self._cv:
                    self._admit()
                    launchable = [r for r in self._active if r.launched < self.n_segments]
                    if not launchable and not inflight:
                        if self._stop and not self._queue:
                            return
                        self._cv.wait(0.05)
                        continue
                if launchable and self._free:
                    inflight.append(self._launch(*self._next_group(launchable)))
                    continue
                self._retire(inflight.popleft())
        except BaseException as exc:  # fail every pending request loudly
            with self._cv:
                for r in list(self._active) + list(self._queue):
                    r.error = exc
                    r.done.set()
                self._active.clear()
                self._queue.clear()
            raise

    def _next_group(self, launchable: List[Request]):
        """The requests of the next launch and whether they are a solo group."""
        if self.solo_guides is None:
            return launchable, False
        shared = [r for r in launchable if len(r.specs) < self.solo_guides]
        groups = ([(shared, False)] if shared else []) + \
            [([r], True) for r in launchable if len(r.specs) >= self.solo_guides]
        self._turn += 1
        return groups[self._turn % len(groups)]

    def _launch(self, reqs: List[Request], solo: bool = False):
        plan = self._free.popleft()
        specs: List[GuideSpec] = []
        mapping = []
        for r in reqs:
            mapping.append((r, len(specs), len(r.specs)))
            specs.extend(r.specs)
            r.launched += 1
        if solo:  # a large request alone walks its own segments
            seg = reqs[0].seg
            reqs[0].seg = (seg + 1) % self.n_segments
        else:
            seg = self._seg
            self._seg = (self._seg + 1) % self.n_segments
        cap = max(1 << 12, int(2 * sum(r.reserve for r in reqs) / self.n_segments))
        if cap > plan.hit_capacity:
            plan.set_hit_capacity(max(cap, 2 * plan.hit_capacity))
        key = tuple(r.rid for r in reqs)
        if self._plan_set[id(plan)] != key:  # the active set changed: new guide tables
            plan.set_guides(specs, np.concatenate([r.descs for r in reqs]))
            self._plan_set[id(plan)] = key
        b, e = self.bounds[seg]
        plan.launch(b, e, 3)
        self.launches += 1
        return plan, seg, mapping

    def _retire(self, item) -> None:
        plan, seg, mapping = item
        n, ms = plan.wait()
        if n > plan.hit_capacity:  # exact re-run of this segment
            plan.set_hit_capacity(int(n * 1.25) + 1024)
            b, e = self.bounds[seg]
            plan.launch(b, e, 3)
            n, ms = plan.wait()
            self.reruns += 1
        self.busy_s += ms / 1e3
        hits = plan.fetch_view(n)
        if len(mapping) == 1:  # one request: its records as they are
            hits = hits.copy()  # the plan's buffer may be refilled from here on
            g = None
        else:
            order = np.argsort(hits["guide"], kind="stable")
            hits = hits[order]  # a copy, as above
            g = hits["guide"]
        self._free.append(plan)
        now = time.perf_counter()
        finished = []
        for req, off, cnt in mapping:
            if g is None:
                part = hits
            else:
                lo, hi = np.searchsorted(g, [off, off + cnt])
                part = hits[lo:hi].copy() if hi > lo else hits[:0]
            if part.size:
                part["guide"] -= off
                req.parts.append(part)
            req.retired += 1
            if req.retired == self.n_segments:
                req.done_at = now
                finished.append(req)
        if finished:
            with self._cv:
                for req in finished:
                    self._active.remove(req)
                    self._reserved -= req.reserve
                    self.completed.append(req)
                    req.done.set()
                self._cv.notify_all()


def latency_stats(reqs: Sequence[Request]) -> Dict[str, float]:
    lat = np.array([r.latency for r in reqs if r.done_at > 0]) * 1e3
    if lat.size == 0:
        return {"n": 0}
    return {"n": int(lat.size), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "max_ms": float(lat.max())}
