"""bgo_scan_batched_noalloc through the C ABI: the same records as
bgo_scan_batched on two calls in a row (the second with fewer guides, so a
stale table or hit buffer of the first would show), with a capped output,
and the size limits rejected with an error."""
import ctypes as C
import dataclasses
import itertools

import numpy as np
import pytest

from beguidedoff_amd import _lib, synth
from beguidedoff_amd.batched import HIT_DTYPE, DeviceGenome, GuideSpec, guide_descs
from beguidedoff_amd.scoring import cctop_table
from test_gpu_kernels import oracle_set
from test_oracle_pins import build_case

pytestmark = pytest.mark.gpu

FIELDS = ("guide", "strand", "pos", "x0", "x1", "mm_mask", "n_mm", "editable", "score")


def _call(name, dg, specs, cap, tuning=None):
    lib = _lib.load()
    descs = guide_descs(specs)
    table = np.array(cctop_table(), dtype=np.float64)
    out = np.zeros(max(cap, 1), dtype=HIT_DTYPE)
    n = C.c_int64(-1)
    t = None
    if tuning:
        t = _lib.Tuning()
        t.engine = tuning
    _lib.check(getattr(lib, name)(dg.handle, _lib.ptr(descs), len(specs), _lib.ptr(table),
                                  C.byref(t) if t is not None else None, _lib.ptr(out), cap,
                                  C.byref(n)))
    return out[:min(int(n.value), cap)], int(n.value)


def _sorted(h):
    h = h[np.lexsort((h["pos"], h["strand"], h["guide"]))]
    return [h[f] for f in FIELDS]


@pytest.fixture(scope="module")
def case(gpu):
    contigs, guides, _ = build_case(61, length=(60_000, 90_000), n_guides=6)
    rng = np.random.default_rng(61)
    seqs = ["".join(rng.choice(list("ACGT"), size=int(rng.choice([20, 21])))) for _ in range(70)]
    pams = ["NGG", "NG", "NNNRRT", "NNGRRT", "TTTV"]
    specs = [GuideSpec(s, pams[i % len(pams)], int(rng.integers(3, 7)), ["ABE", "CBE"][i % 2],
                       (int(rng.integers(1, 5)), int(rng.integers(6, 12))))
             for i, s in enumerate(seqs)]
    specs += [GuideSpec(s, p, 5) for s, p in guides]
    genome = synth.genome_from_strings(contigs)
    dg = DeviceGenome(genome, genome["n_mask_le_fwd"], genome["n_mask_le_rc"])
    yield dg, specs, genome
    dg.close()


@pytest.mark.parametrize("engine", [0, 1, 3, 4])
def test_noalloc_equals_batched_twice(case, engine):
    dg, specs, genome = case
    if engine == 4:  # the seed engine takes budgets <= 4 only; planted guides first
        specs = ([dataclasses.replace(s, max_mm=4) for s in specs[-6:]] +
                 [s for s in specs[:-6] if s.max_mm <= 4])
    for sub in (specs, specs[3:17]):  # the second call is smaller
        want, n_want = _call("bgo_scan_batched", dg, sub, 1 << 16, engine)
        got, n_got = _call("bgo_scan_batched_noalloc", dg, sub, 1 << 16, engine)
        assert n_got == n_want == got.size > 0
        for f, a, b in zip(FIELDS, _sorted(got), _sorted(want)):
            assert np.array_equal(a, b), f
        sites = set(zip(got["guide"].tolist(), got["strand"].tolist(), got["pos"].tolist(),
                        got["n_mm"].tolist()))
        assert sites == oracle_set(genome, [(s.seq, s.pam) for s in sub], None,
                                   mm_per=[s.max_mm for s in sub])


def test_noalloc_capped_output_and_limits(case):
    dg, specs, _ = case
    full, n = _call("bgo_scan_batched_noalloc", dg, specs, 1 << 16)
    part, n2 = _call("bgo_scan_batched_noalloc", dg, specs, 5)
    assert n2 == n > 5 and part.size == 5
    key = lambda h: set(zip(h["guide"].tolist(), h["strand"].tolist(), h["pos"].tolist()))
    assert key(part) <= key(full)
    with pytest.raises(RuntimeError, match="out_capacity"):
        _call("bgo_scan_batched_noalloc", dg, specs, 262_145)
    many = [GuideSpec(specs[0].seq, "NGG", 3)] * 4097
    with pytest.raises(RuntimeError, match="4097 guides"):
        _call("bgo_scan_batched_noalloc", dg, many, 16)
    pams = ["".join(p) for p in itertools.product("ACGT", repeat=4)]
    fam = [GuideSpec(specs[0].seq, p, 3) for p in pams[:65]]
    with pytest.raises(RuntimeError, match="65 PAM families"):
        _call("bgo_scan_batched_noalloc", dg, fam, 16)
    # the array is still good after the rejected calls
    again, n3 = _call("bgo_scan_batched_noalloc", dg, specs, 1 << 16)
    assert n3 == n and key(again) == key(full)
