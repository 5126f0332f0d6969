#!/usr/bin/env python3
"""Benchmark of the extensive-mode off-target scan on one MI355X.

Workload (default): the 3.09 Gbp hg38-like synthetic genome (24 contigs,
41% GC, ~5% N) and 1024 mixed guides (512 NGG/20 nt, 256 NNNRRT/21 nt,
128 NNGRRT/21 nt, 128 NG/20 nt; half ABE 4-8, half CBE 3-9) at <= 4 mm.

A step is one full scan of both strands for every guide by the batched
persistent kernel (hipGraph replay) including the copy of all site records
to the host; the genome and guide table are resident in HBM before timing.
The plain drop-in kernel (the reference's decomposition) is timed on the
same workload, and the oracle (CPU restatement) on a bounded sample of it.
Beside them, the reference's CPU fast mode (exact anchors at the CCTOP cutoff
5, restated in C over its own index) runs on the whole genome for the 20-nt
guides, and its keep/remove verdicts are checked against the GPU sites.

Prints one JSON line.  guide*Gbp/s = guides x single-strand genome Gbp /
seconds, both strands always scanned.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "guide·Gbp/s at ≤4 mm, batched vs plain kernel; p50/p99 request latency"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def device_peaks(info: dict) -> dict:
    """VALU peak in 32-bit lane-ops/s: MI355X's published FP32 vector peak
    (157.3 TFLOP/s = 128 FMA lanes/clk/CU x 256 CU x 2.4 GHz) counts 128
    lanes/clk/CU; measured on gfx950, v_xor/v_add/v_bitop3 with VGPR
    operands issue at up to ~98 and v_bcnt/v_min/v_bfi/SGPR-operand forms at
    ~61 lanes/clk/CU."""
    clk = info["clock_khz"] * 1e3 if info["clock_khz"] > 0 else 2.4e9
    return {"valu_lane_ops": info["cu_count"] * 128 * clk, "clock_hz": clk,
            "hbm_bytes_per_s": 8.0e12}


def algorithmic_ops(L_pos: int, n_strands: int, families) -> tuple:
    """(ops of the bit-plane formulation, ops of a per-unit cost model of a
    plain INT32 compare: 3 per position and family, 11 per (guide, window)
    pair) for one launch; families = [(guides, windows)] where windows counts
    PAM-passing windows over both strands."""
    ops = n_strands * L_pos * 2                       # window + PAM extraction
    ops += sum(n_strands * L_pos * 4 for _ in families)  # 3 bit-selects + compare
    ops += sum(4 * g * w for g, w in families)        # xor, bitop3, bcnt, test
    model = sum(n_strands * L_pos * 3 + 11 * g * w for g, w in families)
    return ops, model


ENGINE_NAMES = {1: "batched VALU", 2: "batched int8 MFMA", 3: "batched fp4 MFMA",
                4: "batched seed-and-verify (VALU)"}


def seed_ops(L_pos: int, n_strands: int, families, verified) -> int:
    """32-bit lane-ops of the seed-and-verify formulation for one launch: the
    window and PAM extraction and PAM tests of the bit-plane scan, per
    (window, 4-base block) the block's key and the window's other 16 bases
    in both halves of a word per plane (10 ops), and per pair of verified
    guide-table entries (the table holds two guides per word) one xor per
    plane and their or, then per entry its half, popcount and compare (9 ops
    per pair); `verified` counts the entries on the device."""
    ops = n_strands * L_pos * 2 + sum(n_strands * L_pos * 4 for _ in families)
    ops += sum(5 * 10 * w for _, w in families)
    ops += 9 * int(sum(verified)) // 2
    return ops


def mfma_ops(families) -> int:
    """Matrix-core ops of the MFMA formulation for one launch: each (guide,
    PAM-passing window) pair is a 60-long dot product over the tetrahedral
    code (3 elements per base x 20 bases), i.e. 60 multiply-adds = 120 ops
    (int8 ops for the int8 engine, fp4 flops for the fp4 engine)."""
    return sum(120 * g * w for g, w in families)


# HBM traffic of one launch of the dominant kernel on the default workload,
# from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this benchmark's own
# step loop, each scaled by the same counter's reading for a copy kernel of
# known size.  Keyed by (workload, engine, max_mm): (bytes, the profile that
# holds the counter runs and their commands).  None when not profiled.
TRAFFIC_BYTES = {("hg38_mixed", 3, 4): (1_601_989_374,
                                        "profiles/batched_scan_kernel_fp4_hg38_mixed_1024g_4mm.md")}

# Busy fractions of the units of the dominant kernel from rocprofv3 --pmc
# (VALU: 4 cycles per SQ_INSTS_VALU instruction; MFMA: SQ_VALU_MFMA_BUSY_CYCLES;
# both together: SQ_VALU_MFMA_COEXEC_CYCLES; all over GRBM_GUI_ACTIVE
# SIMD-cycles) on the same workload, same profile as the traffic.
UNIT_BUSY = {("hg38_mixed", 3, 4): {"valu": 0.681, "mfma": 0.517, "valu_and_mfma": 0.232}}

# The seed engine on 8192 NGG guides (two 4,096-guide tables, two launches
# per step): VALU busy from SQ_INSTS_VALU over GRBM_GUI_ACTIVE SIMD-cycles.
SEED_BUSY = {("hg38_8k", 4, 4): ({"valu": 0.849},
                                 "profiles/seed_scan_kernel_hg38_8k_4mm.md")}


def site_keys(guide, strand, pos, n_mm) -> np.ndarray:
    """(guide, strand, position, N-corrected mismatches) of each site packed
    in one uint64, for set comparisons and a hash of the site set."""
    return ((np.asarray(guide, np.uint64) << np.uint64(38))
            | (np.asarray(strand, np.uint64) << np.uint64(37))
            | (np.asarray(pos, np.uint64) << np.uint64(5)) | np.asarray(n_mm, np.uint64))


def roofline_block(bound, achieved, peak, unit, workload, engine, max_mm, alg, families,
                   model_ops, L, kms, peaks) -> dict:
    """The dominant kernel against its peak, plus what binds it and how the
    per-unit cost models of a plain compare price the same launch."""
    traffic = TRAFFIC_BYTES.get((workload, engine, max_mm))
    busy = UNIT_BUSY.get((workload, engine, max_mm))
    pairs = int(sum(gg * ww for gg, ww in families))
    secs = kms / 1e3
    valu_peak = peaks["valu_lane_ops"]
    out = {"bound": bound, "achieved": round(achieved, 3), "peak": round(peak, 3), "unit": unit,
           "frac": round(achieved / peak, 4),
           "traffic": traffic[0] if traffic else None,
           "traffic_source": traffic[1] if traffic else None,
           "kernel": "seed_scan_kernel" if engine == 4 else "batched_scan_kernel",
           "ops_per_launch": int(alg),
           "pairs_per_launch": pairs,
           "unit_model_ops_per_launch": int(model_ops),
           "hbm_gbps_planes": round(2 * L / 4 / secs / 1e9, 1)}
    if engine >= 2 and busy:
        top = max(("valu", "mfma"), key=busy.get)
        both = busy.get("valu_and_mfma")
        out["binding_unit"] = (
            f"{top.upper()} {busy[top]:.0%} busy, none saturated: the VALU and MFMA phases "
            "overlap little" + (f" (both busy in {both:.0%} of SIMD-cycles, neither in "
                                f"{1 - busy['valu'] - busy['mfma'] + both:.0%})" if both else "")
            + "; the scan and append phases are chains of dependent VALU/LDS steps at 4 "
            "waves per SIMD, and a wave's compare waits for its products before testing them"
            if busy[top] < 0.8 else f"{top.upper()}")
        out["unit_busy"] = dict(busy, source=traffic[1] if traffic else None)
    seed_busy = SEED_BUSY.get((workload, engine, max_mm))
    if engine == 4 and seed_busy:
        out["binding_unit"] = (
            f"VALU {seed_busy[0]['valu']:.0%} busy; the counted lane-ops are 68% of the issued "
            "VALU lane-slots: a wave walks its 64 windows' guide lists for as long as the "
            "longest of them")
        out["unit_busy"] = dict(seed_busy[0], source=seed_busy[1])
    # Per-unit cost models of the same launch: 11 INT32 VALU ops per
    # (guide, window) pair + 3 per window and family, against the INT32 VALU
    # peak; 160 int8 ops per pair against the dense int8 MFMA peak.  The
    # bit-plane and MFMA formulations do less work per pair than the first
    # model assumes, so a frac above 1 there means the model no longer bounds
    # this engine, not that a peak was exceeded.
    out["unit_models"] = {
        "int32_valu": {"ops": int(model_ops), "peak_tops": round(valu_peak / 2 / 1e12, 1),
                       "frac": round(model_ops / secs / (valu_peak / 2), 3)},
        "int8_onehot_mfma": {"ops": 160 * pairs, "peak_tops": 5000.0,
                             "frac": round(160 * pairs / secs / 5.0e15, 3)}}
    return out


def copy_probe(mbytes: int) -> None:
    """Calibration for the HBM byte counters: a device copy of `mbytes` MiB
    (reads and writes exactly that many bytes per launch), three launches.
    Runs without the scan library in the process."""
    import torch

    n = mbytes << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy probe: 3 x {n} bytes")


def genome_text(g: dict) -> np.ndarray:
    """ASCII bytes of a generated genome: A C G T from the forward words, N
    where the artefact's N mask is set (N runs are coded 0 in the words)."""
    L = int(g["total_bp"])
    lut = np.zeros(256, dtype=np.uint32)
    for b in range(256):
        lut[b] = sum(ord("ACGT"[(b >> (2 * k)) & 3]) << (8 * k) for k in range(4))
    words = np.ascontiguousarray(g["genome_forward"], dtype=np.uint32)
    text = np.empty(words.size * 16, dtype=np.uint8)
    step = 1 << 24  # words per chunk
    for a in range(0, words.size, step):
        chunk = words[a:a + step].view(np.uint8)
        text[a * 16:(a + chunk.size // 4) * 16] = np.take(lut, chunk).view(np.uint8)
    text = text[:L]
    nm = g["n_mask"]
    bstep = 1 << 27  # mask bytes per chunk
    for a in range(0, nm.size, bstep):
        bits = np.unpackbits(nm[a:a + bstep])[: max(0, L - 8 * a)].view(bool)
        text[8 * a:8 * a + bits.size][bits] = ord("N")
    return text


def encoder_probe(workload: str) -> dict:
    """The GPU genome encoder (bgo_encode_genome) on the text of a generated
    genome; its outputs must equal the generator's own arrays bit for bit."""
    from beguidedoff_amd import synth
    from beguidedoff_amd.genome_converter import encode_hip

    g = synth.workload(workload).genome
    t0 = time.perf_counter()
    text = genome_text(g)
    log(f"encoder probe: {text.size} bytes of text in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    fwd, rc, nm, kernel_ms = encode_hip(text, g["chrom_offsets"], g["chrom_lengths"])
    wall = time.perf_counter() - t0
    exact = bool(np.array_equal(fwd, g["genome_forward"]) and np.array_equal(rc, g["genome_rc"])
                 and np.array_equal(nm, g["n_mask"]))
    n = int(text.size)
    algo = n + 2 * 4 * ((n + 15) // 16) + (n + 7) // 8  # read text; write fwd, rc, mask
    return {"metric": "genome_encoder", "workload": workload, "bases": n,
            "kernel_ms": round(kernel_ms, 3), "algorithmic_bytes": algo,
            "achieved_gbps": round(algo / kernel_ms / 1e6, 1), "peak_gbps": 8000.0,
            "frac": round(algo / kernel_ms / 1e6 / 8000.0, 3),
            "call_s_with_pcie": round(wall, 3), "outputs_equal_generator": exact}


def fast_probe(workload: str, cutoff: float = 5.0, reps: int = 5) -> dict:
    """The GPU fast mode's anchor index (bgo_fast_index_*) on a generated
    genome: build time and bytes, then `reps` queries of every 20-nt guide of
    the workload (kernel time from HIP events, and the call end to end)."""
    from beguidedoff_amd.fast import FastIndexGPU, anchor_table
    from beguidedoff_amd.batched import DeviceGenome
    from beguidedoff_amd.finder import _ScanGuide
    from beguidedoff_amd import _lib, synth

    _lib.init(0)
    w = synth.workload(workload)
    g = w.genome
    keys = list(dict.fromkeys(_ScanGuide(x.guide_seq, x.pam) for x in w.guides
                              if len(x.guide_seq) == 20))
    dg = DeviceGenome(g, g["n_mask_le_fwd"], g["n_mask_le_rc"])
    ix = FastIndexGPU(dg, g["chrom_offsets"], g["chrom_lengths"])
    log(f"fast probe: index of {int(g['total_bp'])} positions built in {ix.build_s:.3f} s")
    anchors = anchor_table(keys, cutoff)
    kms, walls, n = [], [], 0
    for _ in range(reps):
        t0 = time.perf_counter()
        hits, ms = ix.query(anchor_table(keys, cutoff), cutoff)
        walls.append(time.perf_counter() - t0)
        kms.append(ms)
        n = int(hits.size)
    out = {"metric": "fast_index", "workload": workload, "positions": int(g["total_bp"]),
           "guides": len(keys), "anchors": int(anchors.size), "cutoff": cutoff,
           "index_build_s": round(ix.build_s, 4), "index_bytes": ix.nbytes,
           "query_kernel_ms": round(float(np.median(kms)), 4),
           "query_call_ms": round(1e3 * float(np.median(walls)), 3), "candidates": n}
    ix.close()
    dg.close()
    return out


def cpu_baseline(workload, max_mm: int, target_s: float = 12.0) -> dict:
    from oracle.scan import scan_strand
    from beguidedoff_amd.encoding import encode_dna, encode_guide_64bit
    from beguidedoff_amd.pam import encode_pam_mask

    g = workload.genome
    guides = workload.guides
    vals = np.array([encode_guide_64bit(encode_dna(x.guide_seq)) & ((1 << 64) - 1) for x in guides],
                    dtype=np.uint64)
    masks = np.stack([encode_pam_mask(x.pam) for x in guides])
    lens = np.array([len(x.pam) for x in guides], np.uint8)
    tp = np.ones(len(guides), np.uint8)
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 64))
    start = int(g["chrom_offsets"][0]) + 1_000_000

    def run(n_pos):
        t = time.perf_counter()
        for key, nm, enc in (("genome_forward", "n_mask_le_fwd", 0), ("genome_rc", "n_mask_le_rc", 3)):
            scan_strand(g[key], g["total_bp"], vals, masks, lens, tp, max_mm, g.get(nm), enc,
                        pos_begin=start, pos_end=start + n_pos, threads=threads)
        return time.perf_counter() - t

    probe = 20_000
    dt = run(probe)
    n_pos = int(min(max(probe, probe * target_s / max(dt, 1e-6)), 50_000_000))
    dt = run(n_pos)
    value = len(guides) * (n_pos / 1e9) / dt
    return {"value": value, "unit": "guide·Gbp/s", "cores": threads, "kind": "port",
            "sample": f"oracle scan of all {len(guides)} guides, both strands, {n_pos} positions "
                      f"({dt:.1f} s, {threads} threads)"}


def fast_mode_leg(workload, hits: np.ndarray, threads: int, cutoff: float = 5.0) -> dict:
    """The reference's CPU fast mode (exact cold-region anchors at the CCTOP
    cutoff), restated in C over its own index of the forward genome: index
    build, query time single-threaded and on all threads, and its keep/remove
    verdict per 20-nt guide against this run's GPU sites (remove = an
    editable, contig-internal site with score <= cutoff)."""
    from oracle.fastmode import FastIndex, any_editable_row

    g = workload.genome
    sel = [i for i, x in enumerate(workload.guides) if len(x.guide_seq) == 20]
    guides = [(workload.guides[i].guide_seq, workload.guides[i].pam) for i in sel]
    ix = FastIndex(g, threads=threads)
    build_s = ix.build_s
    t0 = time.perf_counter()
    res = ix.search(guides, cutoff, threads=1)
    q1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = ix.search(guides, cutoff, threads=threads)
    qn = time.perf_counter() - t0
    index_bytes = ix.nbytes
    ix.close()

    offs = np.asarray(g["chrom_offsets"], dtype=np.int64)
    lens = np.asarray(g["chrom_lengths"], dtype=np.int64)
    nm = g.get("n_mask_le_fwd")

    def fwd_window(strand, pos, win):
        k = int(np.searchsorted(offs, pos, side="right") - 1)
        off, ln = int(offs[k]), int(lens[k])
        if pos + win > off + ln:
            return None  # junction window (scanned by the extensive mode only)
        return pos if strand == 0 else off + ln - (pos - off) - win

    def has_n(left, win):
        if nm is None:
            return False
        return any((nm[i >> 3] >> (i & 7)) & 1 for i in range(left, left + win))

    remove_ext, n_ext = {}, {}
    gsel = set(sel)
    keep = hits[(hits["editable"] == 1) & (hits["score"] <= cutoff)]
    for h in keep:
        gi = int(h["guide"])
        if gi not in gsel:
            continue
        win = 20 + len(workload.guides[gi].pam)
        left = fwd_window(int(h["strand"]), int(h["pos"]), win)
        if left is None:
            continue
        remove_ext[gi] = True
        n_ext[gi] = n_ext.get(gi, False) or has_n(left, win)
    agree = disagree = disagree_n = overflow = 0
    n_remove_fast = 0
    for gi, r in zip(sel, res):
        x = workload.guides[gi]
        rf = any_editable_row(r, x.editor_type, x.edit_window_start, x.edit_window_end)
        n_remove_fast += rf
        overflow += r.overflow
        if rf == remove_ext.get(gi, False):
            agree += 1
        else:
            disagree += 1
            win = 20 + len(x.pam)
            if n_ext.get(gi, False) or any(has_n(int(l), win) for l in r.leftmost.tolist()):
                disagree_n += 1
    # the same search on the GPU (beguidedoff_amd.fast), both candidate engines:
    # candidate sets per guide against the C restatement's
    from beguidedoff_amd.fast import beguidedofffast
    from beguidedoff_amd.types import GuideInput

    gbp = g["total_bp"] / 1e9
    ff = beguidedofffast(genome=g, search_cutoff=cutoff, engine="index")
    inputs = [GuideInput(f"g{i}", s, p, "3prime", 1, 20, "ABE") for i, (s, p) in enumerate(guides)]
    keys, _ = ff._ext._deduplicate_guides(inputs)

    def cand_sets(fh):
        out = {}
        for kk, st, pos, mm, sc in zip(fh.key.tolist(), fh.strand.tolist(), fh.pos.tolist(),
                                       fh.mm_out.tolist(), fh.score.tolist()):
            win = 20 + len(keys[kk].pam)
            if st == 1:
                c = int(np.searchsorted(offs, pos, side="right") - 1)
                pos = int(offs[c] + lens[c] - (pos - offs[c]) - win)
            out.setdefault((keys[kk].guide_seq, keys[kk].pam), set()).add((pos, st, mm, sc))
        return out

    def n_equal(sets):
        return int(sum(sets.get(key, set()) == set(zip(r.leftmost.tolist(), r.strand.tolist(),
                                                         r.mm.tolist(), r.score.tolist()))
                       for key, r in zip(guides, res)))

    gpu_fast = {"unit": "guide·Gbp/s", "guides": len(guides)}
    for engine, candidates in (("index", ff._candidates_index), ("scan", ff._candidates_scan)):
        candidates(keys, cutoff)  # first launch
        t0 = time.perf_counter()
        fh = candidates(keys, cutoff)
        gpu_s = time.perf_counter() - t0
        leg = {"value": round(len(guides) * gbp / gpu_s, 1), "query_s": round(gpu_s, 5),
               "candidates": int(fh.key.size),
               "guides_with_equal_candidate_sets": n_equal(cand_sets(fh))}
        if engine == "index":
            leg.update({"query_kernel_ms": round(ff.last_kernel_ms, 3),
                        "index_build_s": round(ff._index.build_s, 3),
                        "index_bytes": ff._index.nbytes})
        gpu_fast[engine] = leg
    gpu_fast["value"] = gpu_fast["index"]["value"]
    ff.close()
    ff._ext.close()

    log(f"GPU fast mode: {gpu_fast}")
    return {"value": round(len(guides) * gbp / qn, 1), "value_1t": round(len(guides) * gbp / q1, 1),
            "unit": "guide·Gbp/s", "cores": threads, "kind": "port",
            "sample": f"all {len(guides)} 20-nt guides, whole genome, cutoff {cutoff}",
            "algorithm": "the reference's anchored search over a direct-address 10-mer position "
                         "index (oracle/bgo_fast.c), not its sdsl FM-index locate: faster per "
                         "query than the algorithm it stands for, whose published time is "
                         "1.266 s for 100 guides on real hg38 (about 245 guide·Gbp/s, "
                         "benchmarks/REPORT.md:333); anchor cap in position order",
            "query_s": round(qn, 4), "query_s_1t": round(q1, 4), "index_build_s": round(build_s, 2),
            "index_bytes": index_bytes,
            "verdicts": {"guides": len(guides), "agree": agree, "disagree": disagree,
                         "disagree_at_n_windows": disagree_n, "remove_fast": n_remove_fast,
                         "remove_extensive": len(remove_ext), "anchor_cap_overflow": overflow},
            "gpu": gpu_fast}


def _request(rng, pool, max_mm):
    """One request: a log-uniform number (1-500) of distinct guides from `pool`."""
    from beguidedoff_amd.batched import GuideSpec

    n = int(np.exp(rng.uniform(0, np.log(500))))
    idx = rng.choice(len(pool), size=n, replace=False)
    return [GuideSpec(pool[j].guide_seq, pool[j].pam, max_mm, pool[j].editor_type,
                      (pool[j].edit_window_start, pool[j].edit_window_end)) for j in idx]


def serving_leg(dg, pool, max_mm: int, rates, duration: float, min_requests: int = 0,
                seed: int = 0) -> dict:
    """Poisson arrivals of guide-set requests (log-uniform size 1-500 drawn
    from `pool`) at each rate for max(duration, min_requests / rate) seconds;
    p50/p99 latency from arrival to complete site records, offered
    guide*Gbp/s and the GPU's busy fraction."""
    from beguidedoff_amd.serving import ScanServer, latency_stats

    rng = np.random.default_rng(seed)
    gbp = dg.total_bp / 1e9
    out = {}
    for rate in rates:
        span = max(duration, min_requests / rate)
        with ScanServer(dg, n_segments=8, hit_budget=1 << 24, pipeline=2) as srv:
            t0 = time.perf_counter()
            handles = []
            t_next = t0
            t_log = t0 + 30.0
            while True:
                t_next += rng.exponential(1.0 / rate)
                if t_next - t0 > span:
                    break
                delay = t_next - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                handles.append(srv.submit(_request(rng, pool, max_mm)))
                if time.perf_counter() > t_log:
                    done = sum(1 for h in handles if h.done.is_set())
                    log(f"serving {rate} req/s: {time.perf_counter() - t0:.0f} s, "
                        f"{len(handles)} submitted, {done} done")
                    t_log += 30.0
            for h in handles:
                h.result(timeout=600)
            wall = time.perf_counter() - t0
            st = latency_stats(handles)
            guides = sum(len(h.specs) for h in handles)
            st.update({"rate_req_s": rate, "requests": len(handles), "guides": guides,
                       "duration_s": round(wall, 2), "offered_s": round(span, 1),
                       "guide_gbp_s": round(guides * gbp / wall, 1),
                       "launches": srv.launches, "reruns": srv.reruns,
                       "gpu_busy_frac": round(srv.busy_s / wall, 3)})
            out[str(rate)] = {k: (round(v, 2) if isinstance(v, float) else v) for k, v in st.items()}
            log(f"serving {rate} req/s: {out[str(rate)]}")
    return out


def serving_capacity(dg, pool, max_mm: int, duration: float, outstanding: int = 32,
                     seed: int = 1) -> dict:
    """Closed loop: `outstanding` requests (same size distribution) are kept in
    the server for `duration` seconds, each replaced as it completes.  The
    completed requests per second and guide*Gbp/s are the server's capacity
    at that concurrency; latency then includes queueing."""
    from beguidedoff_amd.serving import ScanServer, latency_stats

    rng = np.random.default_rng(seed)
    gbp = dg.total_bp / 1e9
    with ScanServer(dg, n_segments=8, hit_budget=1 << 24, pipeline=2) as srv:
        live = [srv.submit(_request(rng, pool, max_mm)) for _ in range(outstanding)]
        done = []
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < duration:
            live[0].done.wait(1.0)
            still = []
            for h in live:
                if h.done.is_set():
                    h.result(0)
                    done.append(h)
                    still.append(srv.submit(_request(rng, pool, max_mm)))
                else:
                    still.append(h)
            live = still
        wall = time.perf_counter() - t0
        busy = srv.busy_s
        for h in live:
            h.result(timeout=600)
    st = latency_stats(done)
    guides = sum(len(h.specs) for h in done)
    st.update({"outstanding": outstanding, "requests": len(done),
               "req_s": round(len(done) / wall, 1), "guide_gbp_s": round(guides * gbp / wall, 1),
               "gpu_busy_frac": round(min(1.0, busy / wall), 3), "duration_s": round(wall, 1)})
    out = {k: (round(v, 2) if isinstance(v, float) else v) for k, v in st.items()}
    log(f"serving capacity: {out}")
    return out


def finder_leg(w, max_mm: int, reps: int = 3) -> dict:
    """End to end through the drop-in API: beguidedoff(mode='extensive')
    .search(guides=...) for every guide of the workload (dedup, both strands,
    scoring, editability, top 3 per guide, the 12-column table), median of
    `reps` calls after one warm call (the genome is resident after the
    first).  The reference publishes search() times for 100 and 500 guides
    at <= 4 mm, top 3 (benchmarks/REPORT.md:331-335; RTX 3060, real hg38)."""
    from beguidedoff_amd import beguidedoff

    f = beguidedoff(mode="extensive", genome=w.genome, max_mismatches=max_mm, n_top_offtargets=3)
    t0 = time.perf_counter()
    df = f.search(guides=w.guides, verbose=False)
    first = time.perf_counter() - t0
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        df = f.search(guides=w.guides, verbose=False)
        times.append(time.perf_counter() - t0)
    f.close()
    out = {"guides": len(w.guides), "rows": int(len(df)), "search_s": round(float(np.median(times)), 4),
           "first_call_s": round(first, 3),
           "reference_search_s": {"100 guides": 15.76, "500 guides": 180.42}}
    log(f"finder: {out}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="hg38_mixed")
    ap.add_argument("--max-mm", type=int, default=4)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tuning", default="auto",
                    help="'auto' (autotune on a genome slice), '' (library default) or a JSON dict")
    ap.add_argument("--eager", action="store_true", help="launch without the hipGraph")
    ap.add_argument("--rates", default="2,8,32", help="serving arrival rates (req/s); '' skips")
    ap.add_argument("--serving-duration", type=float, default=15.0,
                    help="seconds of Poisson arrivals per rate (at least)")
    ap.add_argument("--serving-requests", type=int, default=120,
                    help="arrivals per rate at least (the window grows to requests / rate)")
    ap.add_argument("--capacity-s", type=float, default=10.0,
                    help="seconds of the closed-loop serving capacity run (0 skips)")
    ap.add_argument("--no-finder", action="store_true",
                    help="skip the end-to-end finder search (API call, table of the top 3)")
    ap.add_argument("--serving-only", action="store_true",
                    help="only the serving legs (long runs of config 5)")
    ap.add_argument("--no-engines", action="store_true", help="skip the VALU/MFMA comparison")
    ap.add_argument("--guides", type=int, default=0,
                    help="scan a fixed random subset of this many of the workload's guides "
                         "(small request sizes; 0 = all); skips the CPU legs")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device copy of this many MiB (HBM counter calibration; "
                         "a process of its own: torch carries its own HIP runtime)")
    ap.add_argument("--fast-probe", action="store_true",
                    help="build the fast mode's anchor index on the workload's genome and time "
                         "queries of its 20-nt guides, then exit")
    ap.add_argument("--encoder-probe", action="store_true",
                    help="only time the GPU genome encoder on the workload's genome text")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "the scan runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        print("bench.py: only --gpus 1 is supported (one MI355X; how the scan would shard is a design note)", file=sys.stderr)
        return 2

    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return 0
    if args.encoder_probe:
        print(json.dumps(encoder_probe(args.workload)), flush=True)
        return 0
    if args.fast_probe:
        print(json.dumps(fast_probe(args.workload)), flush=True)
        return 0

    from beguidedoff_amd import _lib, synth
    from beguidedoff_amd.batched import DeviceGenome, GuideSpec, Plan, default_capacity

    _lib.init(0)
    info = _lib.device_info(0)
    peaks = device_peaks(info)
    log(f"device {info}")

    t0 = time.perf_counter()
    w = synth.workload(args.workload, scale=args.scale)
    g = w.genome
    L = int(g["total_bp"])
    log(f"workload {w.name}: {L} bp, {len(w.guides)} guides, generated in "
        f"{time.perf_counter() - t0:.1f} s")

    specs = [GuideSpec(x.guide_seq, x.pam, args.max_mm, x.editor_type,
                       (x.edit_window_start, x.edit_window_end)) for x in w.guides]
    if args.guides > 0:
        keep = np.sort(np.random.default_rng(1).permutation(len(specs))[:args.guides])
        specs = [specs[i] for i in keep]
        args.no_cpu = True  # the CPU legs time the whole workload's guide set
        log(f"guide subset: {len(specs)} of {len(w.guides)} guides")
    t0 = time.perf_counter()
    dg = DeviceGenome(g, g["n_mask_le_fwd"], g["n_mask_le_rc"])
    log(f"genome resident in HBM ({dg.device_bytes() / 1e9:.2f} GB) in {time.perf_counter() - t0:.1f} s")
    if args.serving_only:
        rates = [float(r) for r in args.rates.split(",") if r]
        out = {"metric": "serving latency (config 5)", "workload": w.name, "max_mismatches": args.max_mm,
               "serving": serving_leg(dg, w.guides, args.max_mm, rates, args.serving_duration,
                                      args.serving_requests)}
        if args.capacity_s > 0:
            out["serving_capacity"] = serving_capacity(dg, w.guides, args.max_mm, args.capacity_s)
        dg.close()
        print(json.dumps(out), flush=True)
        return 0
    tune_trials = None
    if args.tuning == "auto":
        from beguidedoff_amd.autotune import autotune

        t0 = time.perf_counter()
        tuning, trials = autotune(dg, specs, use_cache=False)
        tune_trials = [{"tuning": t, "sample_ms": round(ms, 3)} for t, ms in trials[:4]]
        log(f"autotune ({len(trials)} candidates, {time.perf_counter() - t0:.1f} s): {tuning}")
    else:
        tuning = json.loads(args.tuning) if args.tuning else None
    plan = Plan(dg, specs, tuning=tuning, hit_capacity=default_capacity(len(specs)))

    def step():
        if args.eager:
            plan.launch()
        else:
            plan.graph_launch()
        n, ms = plan.wait()
        if n > plan.hit_capacity:
            raise RuntimeError(f"hit buffer too small ({n} > {plan.hit_capacity})")
        hits = plan.fetch_view(n)  # all site records, DMA'd into page-locked host memory
        return n, ms, hits

    for i in range(args.warmup):
        n, ms, _ = step()
        log(f"warmup {i}: {n} sites, kernel {ms:.2f} ms")
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        n, ms, hits = step()
        kernel_ms.append(ms)
    wall = time.perf_counter() - t0
    hits = hits.copy()  # the last step's records, kept past the plan's next fetch
    ms_per_step = wall * 1e3 / args.steps
    gbp = L / 1e9
    value = len(specs) * gbp / (ms_per_step / 1e3)
    kms = float(np.mean(kernel_ms))
    families = plan.family_stats()
    L_pos = max(0, L - 26)
    ops, model_ops = algorithmic_ops(L_pos, 2, families)
    engine = plan.tuning()["engine"]
    if engine == 2:  # int8 MFMA compare: priced against the dense int8 MFMA peak
        bound, unit, dtype = "mfma", "TOP/s", "int (exact); int8 MFMA operands"
        alg = mfma_ops(families)
        peak = 5.0e3
    elif engine == 3:  # fp4 MFMA compare: priced against the dense fp4 MFMA peak
        bound, unit, dtype = "mfma", "TFLOP/s", "int (exact); fp4 e2m1 MFMA operands"
        alg = mfma_ops(families)
        peak = 10.1e3
    elif engine == 4:  # seed-and-verify: its own lane-ops against the VALU issue peak
        bound, unit, dtype = "valu", "Tops/s", "u32 (exact)"
        alg = seed_ops(L_pos, 2, families, plan.seed_stats())
        peak = peaks["valu_lane_ops"] / 1e12
    else:            # VALU compare: 32-bit lane-ops against the VALU issue peak
        bound, unit, dtype = "valu", "Tops/s", "u32 (exact)"
        alg = ops
        peak = peaks["valu_lane_ops"] / 1e12
    achieved = alg / (kms / 1e3) / 1e12
    log(f"batched: {ms_per_step:.2f} ms/step (kernel {kms:.2f} ms), {value:.1f} guide·Gbp/s, "
        f"{n} sites, families {families}")

    # both compare engines on the same workload (same tiling otherwise)
    engines = {}
    legs = ((1, "valu"), (2, "mfma_i8"), (3, "mfma_fp4"))
    if args.max_mm <= 4:
        legs += ((4, "seed"),)
    for eng, name in () if args.no_engines else legs:
        t_eng = dict(plan.tuning())
        t_eng["engine"] = eng
        if eng == 4 or engine == 4:
            # the seed engine runs one 16-wave block per CU; the others keep the
            # library's occupancy and block shape
            t_eng.update(blocks_per_cu=0, waves_per_block=4)
        p2 = Plan(dg, specs, tuning=t_eng, hit_capacity=default_capacity(len(specs)))
        times = []
        for _ in range(3):
            p2.launch()
            n2, ms2 = p2.wait()
            times.append(ms2)
        engines[name] = {"kernel_ms": round(min(times[1:]), 3), "sites": int(n2)}
        p2.close()
    log(f"engines: {engines}")

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "guide·Gbp/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": w.name, "description": w.description, "genome_bp": L,
                   "guides": len(specs), "max_mismatches": args.max_mm,
                   "pams": sorted({s.pam for s in specs}),
                   "engine": ENGINE_NAMES[engine],
                   "graph": not args.eager, "tuning": plan.tuning()},
        "sites": int(n),
        "engines": engines,
        "autotune": tune_trials,
        "kernel_ms": round(kms, 3),
        "roofline": roofline_block(bound, achieved, peak, unit, w.name, engine, args.max_mm,
                                   alg, families, model_ops, L, kms, peaks),
    }

    # the batched site set, as a hash of its sorted (guide, strand, pos, n_mm) keys
    bk_keys = np.sort(site_keys(hits["guide"], hits["strand"], hits["pos"], hits["n_mm"]))
    out["sites_sha256"] = hashlib.sha256(bk_keys.tobytes()).hexdigest()[:16]
    if not args.no_plain:
        from beguidedoff_amd.cuda_search import last_kernel_ms, search_genome_cuda
        from beguidedoff_amd.encoding import encode_dna, encode_guide_64bit
        from beguidedoff_amd.pam import encode_pam_mask

        vals = np.array([encode_guide_64bit(encode_dna(s.seq)) & ((1 << 64) - 1) for s in specs],
                        dtype=np.uint64)
        masks = np.stack([encode_pam_mask(s.pam) for s in specs])
        lens = np.array([len(s.pam) for s in specs], np.uint8)
        tp = np.ones(len(specs), np.uint8)
        plain_ms, plain_wall, plain_sites = [], [], 0
        for rep in range(args.plain_steps):
            kt, tw, sites = 0.0, time.perf_counter(), 0
            plain_keys = []
            for s_idx, (key, nm, enc) in enumerate((("genome_forward", "n_mask_le_fwd", 0),
                                                    ("genome_rc", "n_mask_le_rc", 3))):
                res, ovf, cnt = search_genome_cuda(g[key], L, vals, masks, lens, tp, args.max_mm,
                                                   return_overflow=True, n_mask_packed=g[nm],
                                                   n_encoded=enc)
                kt += last_kernel_ms()
                sites += sum(cnt.values())
                if any(ovf.values()):
                    raise RuntimeError("plain kernel result buffer overflowed")
                for gi, lst in res.items():
                    if lst:
                        a = np.array(lst, dtype=np.int64)
                        plain_keys.append(site_keys(np.full(len(lst), gi), np.full(len(lst), s_idx),
                                                    a[:, 0], a[:, 1]))
            plain_wall.append(time.perf_counter() - tw)
            plain_ms.append(kt)
            plain_sites = sites
            log(f"plain rep {rep}: kernels {kt:.1f} ms, wall {plain_wall[-1]:.1f} s, {sites} sites")
        pk = float(np.mean(plain_ms))
        out["plain"] = {"value": round(len(specs) * gbp / (pk / 1e3), 2), "kernel_ms": round(pk, 2),
                        "wall_s": round(float(np.mean(plain_wall)), 2), "sites": int(plain_sites),
                        "steps": args.plain_steps}
        out["speedup_vs_plain"] = round(pk / kms, 2)
        pk_keys = np.sort(np.concatenate(plain_keys)) if plain_keys else np.zeros(0, np.uint64)
        out["sites_equal_plain"] = bool(np.array_equal(pk_keys, bk_keys))

    if args.rates:
        rates = [float(r) for r in args.rates.split(",") if r]
        out["serving"] = serving_leg(dg, w.guides, args.max_mm, rates, args.serving_duration,
                                     args.serving_requests)
    if args.capacity_s > 0:
        out["serving_capacity"] = serving_capacity(dg, w.guides, args.max_mm, args.capacity_s)

    if not args.no_finder and args.guides == 0:
        out["finder"] = finder_leg(w, args.max_mm)

    if not args.no_cpu:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(w, args.max_mm)
        log(f"cpu baseline {out['cpu_baseline']} in {time.perf_counter() - t0:.1f} s")
        t0 = time.perf_counter()
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        out["cpu_fast_mode"] = fast_mode_leg(w, hits, max(1, min(threads, 64)))
        log(f"cpu fast mode {out['cpu_fast_mode']} in {time.perf_counter() - t0:.1f} s")

    plan.close()
    dg.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
