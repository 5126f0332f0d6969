"""Autotuner of the batched kernel's launch shape.

Tunables (bgo_tuning): compare engine (VALU xor/popcount, int8 or fp4 MFMA,
seed-and-verify when every budget is <= 4),
positions per work item (genome tile length), waves per block, resident
blocks per CU, and strand interleave of the work items.
Each candidate runs the real kernel on a slice of the resident genome with
the caller's guide set (both strands) and is timed with HIP events; the
fastest is kept.  Results are cached per (device, PAM-family mix, guide
count bucket, budget) in a JSON file (``$BGO_TUNE_CACHE`` or
``~/.cache/beguidedoff_amd/tune.json``).
"""
from __future__ import annotations

import itertools
import json
import os
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

from . import _lib
from .batched import DeviceGenome, GuideSpec, Plan

DEFAULT_SPACE = {
    "engine": [1, 2, 3, 4],
    "chunk_positions": [16384, 32768, 131072],
    "waves_per_block": [2, 4],
    "blocks_per_cu": [0],
    "interleave": [0, 1],
}


def _cache_path() -> Path:
    env = os.environ.get("BGO_TUNE_CACHE")
    return Path(env) if env else Path.home() / ".cache" / "beguidedoff_amd" / "tune.json"


def tuning_key(specs: Sequence[GuideSpec]) -> str:
    fams: Dict[str, int] = {}
    for s in specs:
        fams[s.pam.upper()] = fams.get(s.pam.upper(), 0) + 1
    info = _lib.device_info(0)
    bucket = 1 << max(0, (len(specs) - 1).bit_length())
    mm = max((s.max_mm for s in specs), default=0)
    mix = ",".join(f"{p}:{(1 << max(0, (n - 1).bit_length()))}" for p, n in sorted(fams.items()))
    return f"{info['arch']}|cu{info['cu_count']}|n{bucket}|mm{mm}|{mix}"


def _load_cache() -> dict:
    p = _cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def _store_cache(cache: dict) -> None:
    p = _cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(cache, indent=1, sort_keys=True))
        os.replace(tmp, p)
    except OSError:
        pass


def autotune(genome: DeviceGenome, specs: Sequence[GuideSpec], sample_positions: int = 200_000_000,
             space: Optional[dict] = None, repeats: int = 2, use_cache: bool = True
             ) -> Tuple[dict, List[Tuple[dict, float]]]:
    """Return (best tuning, [(candidate, kernel ms)]) for this guide set."""
    key = tuning_key(specs)
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            return dict(hit["tuning"]), []
    space = dict(space or DEFAULT_SPACE)
    if "engine" in space and max((s.max_mm for s in specs), default=0) > 4:
        space["engine"] = [e for e in space["engine"] if e != 4]  # seed: budgets <= 4 only
    names = list(space)
    L = genome.total_bp
    begin = max(0, L // 2 - sample_positions // 2)
    end = min(L, begin + sample_positions)
    results = []
    plan = None
    for values in itertools.product(*(space[n] for n in names)):
        cand = dict(zip(names, values))
        try:
            plan = Plan(genome, specs, tuning=cand, hit_capacity=1 << 20)
            best = float("inf")
            for _ in range(repeats + 1):           # first run warms up
                plan.launch(begin, end, 3)
                n, ms = plan.wait()
                best = min(best, ms)
            results.append((plan.tuning(), best))
        finally:
            if plan is not None:
                plan.close()
                plan = None
    results.sort(key=lambda r: r[1])
    best = results[0][0]
    if use_cache:
        cache = _load_cache()
        cache[key] = {"tuning": best, "ms": results[0][1], "sample_positions": end - begin}
        _store_cache(cache)
    return best, results
