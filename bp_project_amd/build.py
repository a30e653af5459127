"""Build libbpmi.so (HIP, gfx950) in-tree with hipcc.

    python -m bp_project_amd.build          # or bp_project_amd.build.build()

All translation units in ``csrc/`` are compiled for gfx950 only and linked
into ``bp_project_amd/libbpmi.so``.  Packed FP32 instructions are disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``); FP64 FMA has the
same issue rate as unpacked FP32 on CDNA4, which the kernels exploit.
``-ffp-contract=on`` fuses multiply-adds only within one expression, so the
same source rounds identically in every kernel that inlines it (batched and
per-node FFBP stay bitwise equal) and the double-float error-free
transformations are not fused across statements.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
LIB = HERE / "libbpmi.so"
ARCH = os.environ.get("BPMI_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; set HIPCC")


def sources():
    return sorted(CSRC.glob("*.hip"))


def flags():
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=on",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", f"-I{INCLUDE}", f"-I{CSRC}"]


def libraries():
    """hipFFT (rocFFT) for the range-compression FFTs."""
    rocm = Path(hipcc()).resolve().parent.parent
    return [f"-L{rocm / 'lib'}", "-lhipfft", f"-Wl,-rpath,{rocm / 'lib'}"]


def build(verbose=False):
    """Compile every HIP source into libbpmi.so (always rebuilds)."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *flags(), *map(str, sources()), *libraries(), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with status {proc.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
