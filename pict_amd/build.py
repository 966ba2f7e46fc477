"""Build libpict_hip.so (gfx950) in-tree with hipcc.

    python -m pict_amd.build          # or: python pict_amd/build.py

The library is written to pict_amd/lib/libpict_hip.so.  Every HIP source is
compiled with packed-FP32 instructions disabled.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libpict_hip.so")
SOURCES = ["pict_stages.hip", "pict_solve.hip", "pict_metrics.hip", "pict_grad.hip", "pict_step.hip", "pict_batch.hip", "pict_matvec.hip",
           "pict_noalloc.hip", "pict_sgs.hip", "pict_ilu.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PICT_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src, verbose):
    path = os.path.join(CSRC, src)
    obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", path, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose=False, jobs=None):
    """Compile every source (in parallel: `jobs` processes, default up to 6) and link the library."""
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or max(1, min(len(srcs), 6, os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(lambda src: _compile(src, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for obj in objs:
        os.remove(obj)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
