"""Build the engine's native library in-tree: libes_sieve.so for gfx950 (hipcc).

Run as a script or call build().  The library is what the engine loads through ctypes; it is
rebuilt on every call so a source change can never be shadowed by a stale binary.  Each HIP
source compiles to its own object (in parallel, each with its own device code object), then
the objects link into one shared library.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libes_sieve.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in (
    "es_sieve.hip", "es_sieve_k0.hip", "es_sieve_k1.hip", "es_sieve_k2.hip", "es_sieve_k3.hip",
    "es_filter.hip", "es_mlp.hip", "es_mitm.hip", "es_calib.hip")]
ARCH = os.environ.get("ES_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    f"--offload-arch={ARCH}",
    # packed-FP32 VALU forms are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed: {' '.join(cmd[-3:])}")
    return res


def build(verbose: bool = False, jobs: int | None = None) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    with tempfile.TemporaryDirectory(prefix="es_build_") as tmpd:
        objs = [os.path.join(tmpd, os.path.basename(s) + ".o") for s in SOURCES]
        cmds = [[hipcc, *HIPCC_FLAGS, "-c", s, "-o", o] for s, o in zip(SOURCES, objs)]
        n = jobs or min(len(cmds), max(1, (os.cpu_count() or 1)))
        with ThreadPoolExecutor(max_workers=n) as ex:
            list(ex.map(lambda c: _run(c, verbose), cmds))
        tmp = LIB + ".tmp"
        _run([hipcc, "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp], verbose)
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
