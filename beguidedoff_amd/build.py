"""Build the native libraries of beguidedoff_amd in-tree.

* ``libbgo_hip.so`` — the gfx950 kernels and the C ABI of ``include/bgo.h``
  (hipcc, cross-compiles without a GPU).
* ``libbgo_host.so`` — host-only C helpers (synthetic genome generator).

Run ``python -m beguidedoff_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

HIP_SOURCES = ["bgo_lib.hip", "bgo_plain.hip", "bgo_batched.hip", "bgo_encode.hip",
               "bgo_fastidx.hip", "bgo_sa_order.hip", "bgo_assemble.hip"]
HOST_SOURCES = ["bgo_synth.c"]

HIP_LIB = PKG / "libbgo_hip.so"
HOST_LIB = PKG / "libbgo_host.so"

# -packed-fp32-ops is disabled for every device compile: packed FP32 forms
# are avoided on this hardware.  The host pass warns that it does not know
# the feature; that warning is expected.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA results in VGPRs (not AGPRs): the epilogue reads them with plain VALU ops
    "-mllvm", "-amdgpu-mfma-vgpr-form",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def _run(cmd, verbose: bool) -> None:
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        err = "\n".join(l for l in proc.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"build failed ({cmd[0]}):\n{err}")


def build_hip(verbose: bool = False) -> Path:
    """Compile libbgo_hip.so for gfx950 (atomic replace of the output)."""
    tmp = HIP_LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", *[str(CSRC / s) for s in HIP_SOURCES],
           "-o", str(tmp)]
    _run(cmd, verbose)
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_host(verbose: bool = False) -> Path:
    """Compile libbgo_host.so with gcc (OpenMP)."""
    tmp = HOST_LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fPIC", "-shared", "-fopenmp", f"-I{INCLUDE}",
           *[str(CSRC / s) for s in HOST_SOURCES], "-o", str(tmp)]
    _run(cmd, verbose)
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


def build_all(verbose: bool = False) -> None:
    build_host(verbose)
    build_hip(verbose)


if __name__ == "__main__":
    build_all(verbose="-q" not in sys.argv)
