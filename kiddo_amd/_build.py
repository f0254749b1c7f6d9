"""Compiles the engine into kiddo_amd/libkiddo_amd.so for gfx950 (in-tree, no JIT cache).

    python -m kiddo_amd._build
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libkiddo_amd.so")
ARCH = os.environ.get("KIDDO_AMD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Kiddo's scalar path rounds every product and sum separately: never fuse.
    "-ffp-contract=off",
    "-fno-fast-math",
    # keep packed-FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


SOURCES = ["api.hip", "k_knn.hip", "k_knn_f64.hip", "k_families.hip", "k_metric_manhattan.hip", "k_metric_chebyshev.hip",
           "k_metric_minkowski.hip", "k_build.hip", "k_forest.hip", "server.hip", "host_build.cpp"]
OBJ_DIR = os.path.join(HERE, "_build_obj")


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def _newest_input(src: str) -> float:
    """Latest modification time of a source, the headers every unit includes, and this
    file's flags."""
    deps = [src, __file__, os.path.join(HERE, "..", "include", "kiddo_amd.h")]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    return max(os.path.getmtime(d) for d in deps if os.path.exists(d))


def build(verbose: bool = False, force: bool = True) -> str:
    """Compiles every translation unit in parallel, then links the shared library.
    force=False recompiles only the units whose object is older than its inputs."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ_DIR, exist_ok=True)
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
        objs.append(obj)
        if not force and os.path.exists(obj) and os.path.getmtime(obj) > _newest_input(src):
            continue
        cmd = [hipcc, *HIP_FLAGS, "-c", "-o", obj, src]
        if src.endswith(".cpp"):
            cmd = [hipcc, "-O3", "-std=c++17", "-fPIC", "-Wall", "-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        # each compiler writes to its own file: with pipes, a compiler whose output is not
        # being read blocks once the pipe fills, and the parallel build runs one at a time
        log = tempfile.TemporaryFile(mode="w+")
        procs.append((src, log, subprocess.Popen(cmd, stdout=log, stderr=subprocess.STDOUT)))
    failed = []
    for src, log, p in procs:
        p.wait()
        if p.returncode != 0:
            failed.append(src)
            log.seek(0)
            sys.stderr.write(log.read())
        log.close()
    if failed:
        raise RuntimeError(f"hipcc failed on {failed}")
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-pthread", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link of libkiddo_amd.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, force="--changed" not in sys.argv))
