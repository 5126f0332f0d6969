"""The C ABI library loads on a CPU-only host, exports every entry point
declared in include/bgo.h, and maps failures to RuntimeError."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from beguidedoff_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "bgo.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bgo_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib._SIGNATURES)


def test_struct_sizes_match_header():
    assert _lib.GUIDE_DESC_DTYPE.itemsize == 24
    assert _lib.HIT_DTYPE.itemsize == 32
    assert C.sizeof(_lib.Tuning) == 32


def test_version_and_device_count():
    lib = _lib.load()
    assert lib.bgo_abi_version() == 1
    assert _lib.device_count() >= 0


def test_injected_failure_maps_to_runtime_error():
    lib = _lib.load()
    _lib.check(lib.bgo_debug_fail_next(1))
    words = np.zeros(4, np.uint32)
    h = C.c_void_p()
    rc = lib.bgo_strand_create(_lib.ptr(words), 4, None, 0, 64, C.byref(h))
    assert rc == -4
    with pytest.raises(RuntimeError, match="injected"):
        _lib.check(rc)


def test_bad_arguments_are_rejected_before_device_work():
    lib = _lib.load()
    h = C.c_void_p()
    with pytest.raises(RuntimeError, match="bad arguments"):
        _lib.check(lib.bgo_strand_create(None, 0, None, 0, 0, C.byref(h)))
    with pytest.raises(RuntimeError, match="Too many guides"):
        _lib.check(lib.bgo_scan_plain(C.c_void_p(1), None, None, None, None, 2049, 4, 100, 10,
                                      0, None, None, None))


def test_no_device_means_runtime_error():
    if _lib.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError):
        _lib.init(0)
    from beguidedoff_amd import beguidedoff
    with pytest.raises(RuntimeError, match="requires an MI355X GPU"):
        beguidedoff(genome_path="unused.npz")
