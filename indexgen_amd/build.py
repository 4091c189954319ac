"""Builds libigc.so in-tree for gfx950 with hipcc.

    python -m indexgen_amd.build        # or: from indexgen_amd.build import build; build()

The library has no CPU fallback: it is the HIP kernels plus their host-side
orchestration, the greedy independent set and the output writers.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libigc.so")
CLI = os.path.join(HERE, "igc_cli")

# device code and its host orchestration (hipcc)
SOURCES = ["igc_runtime.hip", "igc_batch.hip", "igc_kmeans.hip", "igc_refine.hip", "igc_server.hip", "igc_gen.hip",
           "igc_noalloc.hip"]
# plain host C++ (greedy, output writers): g++, no HIP
HOST_SOURCES = ["igc_host.cpp", "igc_greedy.cpp"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-fvisibility=hidden",
    # packed FP32 (v_pk_*_f32) is disabled for every kernel built here
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: libigc.so cannot be built")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False) -> str:
    """Compiles every translation unit (in parallel, one process each) and
    links libigc.so, then the CLI; always rebuilds."""
    from concurrent.futures import ThreadPoolExecutor
    inc = ["-I", os.path.join(ROOT, "include")]
    jobs = []
    for h in HOST_SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(h)[0] + ".host.o")
        jobs.append((obj, ["g++", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-pthread", *inc, "-c", "-o", obj,
                           os.path.join(CSRC, h)]))
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        jobs.append((obj, [hipcc(), *HIPCC_FLAGS, *inc, "-c", "-o", obj, os.path.join(CSRC, src)]))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(workers) as ex:
        for f in [ex.submit(_run, cmd, verbose) for _, cmd in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc(), *HIPCC_FLAGS, "-shared", "-o", tmp, *[o for o, _ in jobs], "-lpthread"], verbose)
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "igc_cli.cpp")
    if os.path.exists(cli_src):
        _run(["g++", "-O2", "-std=c++17", *inc, "-o", CLI + ".tmp", cli_src, "-L", HERE, "-ligc",
              "-Wl,-rpath,$ORIGIN", "-lpthread"], verbose)
        os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
