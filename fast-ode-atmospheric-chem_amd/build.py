"""Compile the HIP library for gfx950 in-tree (no JIT cache): the .so lands
next to this file and travels with the source tree."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRCS = [HERE / "csrc" / "fastode_hip.hip", HERE / "csrc" / "batched.hip", HERE / "csrc" / "adjoint.hip"]
LIB = HERE / "libfastode_hip.so"
ARCH = os.environ.get("FODE_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # The controller and stage arithmetic reproduce NumPy's separate
    # multiply and add roundings; fused multiply-adds are written as fma().
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False) -> Path:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, *FLAGS, "-I", str(HERE.parent / "include"), *map(str, SRCS), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
