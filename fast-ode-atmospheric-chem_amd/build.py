"""Compile the HIP library for gfx950 in-tree (no JIT cache): the .so lands
next to this file and travels with the source tree."""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRCS = [HERE / "csrc" / f for f in ("fastode_hip.hip", "batched.hip", "batched_unstaged.hip", "batched_occ2.hip",
                                    "adjoint.hip")]
LIB = HERE / "libfastode_hip.so"
ARCH = os.environ.get("FODE_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    f"--offload-arch={ARCH}",
    # The controller and stage arithmetic reproduce NumPy's separate
    # multiply and add roundings; fused multiply-adds are written as fma().
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False, phase_timers=False) -> Path:
    """Each source to an object in parallel (the batched kernels dominate),
    then one shared library.  phase_timers: the same code with the batched
    kernel's phase timers (-DFODE_PHASE_TIMERS), as libfastode_hip_timed.so."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    lib = LIB.with_name("libfastode_hip_timed.so") if phase_timers else LIB
    extra = ["-DFODE_PHASE_TIMERS"] if phase_timers else []
    with tempfile.TemporaryDirectory(prefix="fode_build_") as tmpdir:
        objs, procs = [], []
        for src in SRCS:
            obj = Path(tmpdir) / (src.stem + ".o")
            cmd = [hipcc, *FLAGS, *extra, "-I", str(HERE.parent / "include"), "-c", str(src), "-o", str(obj)]
            if verbose:
                print(" ".join(cmd), file=sys.stderr)
            objs.append(obj)
            procs.append((cmd, subprocess.Popen(cmd)))
        failed = [cmd for cmd, pr in procs if pr.wait() != 0]
        if failed:
            raise subprocess.CalledProcessError(1, failed[0])
        tmp = lib.with_suffix(".so.tmp")
        cmd = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
        if verbose:
            print(" ".join(cmd), file=sys.stderr)
        subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    return lib


PROBES = [HERE / "csrc" / "probes" / "core_factor_probe.hip"]


def build_probes(verbose=False) -> list:
    """The measurement probes (standalone executables next to their sources,
    not part of the library): core_factor_probe times the core factor, six
    restructurings of it checked bit for bit against it, its throughput at
    one to four workgroups per CU, and the latency floors of a pivot round."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out = []
    for src in PROBES:
        exe = src.with_suffix("")
        cmd = [hipcc, *FLAGS, "-I", str(HERE.parent / "include"), str(src), "-o", str(exe)]
        if verbose:
            print(" ".join(cmd), file=sys.stderr)
        subprocess.run(cmd, check=True)
        out.append(exe)
    return out


if __name__ == "__main__":
    if "--probes" in sys.argv[1:]:
        print(*build_probes(verbose=True), sep="\n")
    else:
        print(build(verbose=True, phase_timers="--phase-timers" in sys.argv[1:]))
