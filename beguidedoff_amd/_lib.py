"""ctypes binding of libbgo_hip.so (the C ABI declared in include/bgo.h).

Loading fails loudly: there is no CPU fallback for the scan.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libbgo_hip.so"

GUIDE_DESC_DTYPE = np.dtype([
    ("seq2bit", "<u8"), ("pam_mask", "u1", (6,)), ("pam_len", "u1"), ("max_mm", "u1"),
    ("edit_mask", "<u4"), ("editor", "u1"), ("pad", "u1", (3,)),
], align=True)
HIT_DTYPE = np.dtype([
    ("guide", "<u4"), ("pos", "<u4"), ("x0", "<u4"), ("x1", "<u4"), ("mm_mask", "<u4"),
    ("strand", "u1"), ("n_mm", "u1"), ("editable", "u1"), ("pad", "u1"), ("score", "<f8"),
], align=True)
ASM_KEY_DTYPE = np.dtype([
    ("g0", "<u4"), ("g1", "<u4"), ("ot_local", "<i8"), ("ot_contig", "<i4"), ("ot_strand", "<i4"),
    ("first_input", "<u4"), ("n_inputs", "<u4"), ("pam_len", "<u4"), ("pad", "<u4"),
], align=True)
ASM_INPUT_DTYPE = np.dtype([("id", "<u4"), ("window", "<u4"), ("target", "<u4"), ("order", "<u4")],
                           align=True)
FAST_ANCHOR_DTYPE = np.dtype([
    ("key_lo", "<u4"), ("key_hi", "<u4"), ("a0", "<u4"), ("a1", "<u4"), ("g0", "<u4"),
    ("g1", "<u4"), ("n_code", "<u4"), ("pam_nibs", "<u4"), ("guide", "<u4"), ("strand", "u1"),
    ("pam_len", "u1"), ("cold_first", "u1"), ("pad", "u1"),
], align=True)
ASM_ROW_DTYPE = np.dtype([("hit", HIT_DTYPE), ("input", "<u4"), ("rank", "<u4"),
                          ("contig", "<u4"), ("local", "<u4")], align=True)
assert GUIDE_DESC_DTYPE.itemsize == 24 and HIT_DTYPE.itemsize == 32
assert ASM_KEY_DTYPE.itemsize == 40 and ASM_INPUT_DTYPE.itemsize == 16
assert ASM_ROW_DTYPE.itemsize == 48
assert FAST_ANCHOR_DTYPE.itemsize == 40


class Tuning(C.Structure):
    _fields_ = [("chunk_positions", C.c_int32), ("waves_per_block", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("engine", C.c_int32),
                ("guide_unroll", C.c_int32), ("reserved", C.c_int32 * 3)]

    def as_dict(self) -> dict:
        d = {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "reserved"}
        d["interleave"] = int(self.reserved[0])
        d["direct_max"] = int(self.reserved[1])
        return d


# Every symbol of include/bgo.h with its ctypes signature.
_P = C.c_void_p
_SIGNATURES = {
    "bgo_abi_version": (C.c_int, []),
    "bgo_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "bgo_init": (C.c_int, [C.c_int32]),
    "bgo_last_error": (C.c_char_p, []),
    "bgo_device_info": (C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int64), C.c_char_p, C.c_int32]),
    "bgo_debug_fail_next": (C.c_int, [C.c_int32]),
    "bgo_strand_create": (C.c_int, [_P, C.c_int64, _P, C.c_int64, C.c_int64, C.POINTER(_P)]),
    "bgo_strand_destroy": (C.c_int, [_P]),
    "bgo_scan_plain": (C.c_int, [_P, _P, _P, _P, _P, C.c_int32, C.c_int32, C.c_int64,
                                 C.c_int64, C.c_int32, _P, _P, _P]),
    "bgo_scan_plain_last_ms": (C.c_int, [C.POINTER(C.c_float)]),
    "bgo_genome_create": (C.c_int, [_P, _P, C.c_int64, _P, _P, C.c_int64, C.c_int64,
                                    C.POINTER(_P)]),
    "bgo_genome_destroy": (C.c_int, [_P]),
    "bgo_genome_device_bytes": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "bgo_plan_create": (C.c_int, [_P, _P, C.c_int32, _P, C.POINTER(Tuning), C.c_int64,
                                  C.POINTER(_P)]),
    "bgo_plan_destroy": (C.c_int, [_P]),
    "bgo_plan_set_guides": (C.c_int, [_P, _P, C.c_int32]),
    "bgo_plan_set_hit_capacity": (C.c_int, [_P, C.c_int64]),
    "bgo_plan_set_end_guard": (C.c_int, [_P, C.c_int32]),
    "bgo_plan_launch": (C.c_int, [_P, C.c_int64, C.c_int64, C.c_int32]),
    "bgo_plan_wait": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_float)]),
    "bgo_plan_fetch": (C.c_int, [_P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "bgo_plan_fetch_pinned": (C.c_int, [_P, C.c_int64, C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_int64)]),
    "bgo_plan_graph_launch": (C.c_int, [_P, C.c_int32]),
    "bgo_plan_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "bgo_plan_family_stats": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), _P, _P]),
    "bgo_plan_seed_stats": (C.c_int, [_P, C.c_int32, _P]),
    "bgo_scan_batched": (C.c_int, [_P, _P, C.c_int32, _P, C.POINTER(Tuning), _P, C.c_int64,
                                   C.POINTER(C.c_int64)]),
    "bgo_plan_assemble": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int32, C.c_int32, _P, _P, _P,
                                    C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                    C.c_int64, _P, C.c_int64, C.POINTER(C.c_int64), _P,
                                    C.POINTER(C.c_int64)]),
    "bgo_scan_batched_noalloc": (C.c_int, [_P, _P, C.c_int32, _P, C.POINTER(Tuning), _P,
                                           C.c_int64, C.POINTER(C.c_int64)]),
    "bgo_fast_index_offsets": (C.c_int, [_P, _P]),
    "bgo_fast_index_positions": (C.c_int, [_P, C.c_int64, C.c_int64, _P]),
    "bgo_sa_text_create": (C.c_int, [_P, C.c_int64, C.c_int64, _P, _P, C.c_int32,
                                     C.POINTER(_P)]),
    "bgo_sa_text_destroy": (C.c_int, [_P]),
    "bgo_sa_text_long_runs": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "bgo_sa_text_select": (C.c_int, [_P, _P, C.c_int64, _P, C.c_int32, C.c_int64, _P,
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "bgo_encode_genome": (C.c_int, [C.c_int32, _P, C.c_int64, _P, _P, C.c_int32, _P, _P, _P,
                                    C.POINTER(C.c_float)]),
    "bgo_fast_index_create": (C.c_int, [_P, _P, _P, C.c_int32, C.POINTER(_P)]),
    "bgo_fast_index_destroy": (C.c_int, [_P]),
    "bgo_fast_index_bytes": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "bgo_fast_index_query": (C.c_int, [_P, _P, C.c_int32, C.c_double, _P, _P, C.c_int64,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_float)]),
}

_lock = threading.Lock()
_lib = None


class LibraryUnavailable(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load libbgo_hip.so once; raise LibraryUnavailable when it is missing."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise LibraryUnavailable(
                f"{LIB_PATH.name} is not built; run `python -m beguidedoff_amd.build` "
                "(the extensive scan has no CPU fallback)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int) -> None:
    """Map a non-zero status to RuntimeError carrying bgo_last_error()."""
    if rc != 0:
        msg = load().bgo_last_error().decode("utf-8", "replace")
        raise RuntimeError(msg or f"libbgo_hip call failed with status {rc}")


def ptr(arr: np.ndarray) -> C.c_void_p:
    return C.c_void_p(arr.ctypes.data) if arr is not None and arr.size else C.c_void_p(0)


def device_count() -> int:
    n = C.c_int32(0)
    check(load().bgo_device_count(C.byref(n)))
    return int(n.value)


def device_info(device: int = 0) -> dict:
    cu, clk, mem = C.c_int32(0), C.c_int32(0), C.c_int64(0)
    arch = C.create_string_buffer(64)
    check(load().bgo_device_info(int(device), C.byref(cu), C.byref(clk), C.byref(mem), arch, 64))
    return {"cu_count": int(cu.value), "clock_khz": int(clk.value), "hbm_bytes": int(mem.value),
            "arch": arch.value.decode()}


_inited: set = set()


def init(device: int = 0) -> None:
    """Select `device` for this thread (idempotent per process+device)."""
    check(load().bgo_init(int(device)))
    _inited.add(int(device))
