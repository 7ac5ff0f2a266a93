"""Builds libalgotom_hip.so in-tree (algotom_amd/_lib/) with hipcc for gfx950.

    python -m algotom_amd.build
"""

import os
import subprocess
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
SRC_DIR = os.path.join(_HERE, "csrc")
LIB_DIR = os.path.join(_HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libalgotom_hip.so")
SOURCES = ("capi.hip", "ramp_filter.hip", "backproject.hip",
           "forward_project.hip", "conversion.hip", "filter2d.hip",
           "vertical.hip", "metrics.hip", "dfi.hip")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # the reference kernels never contract a*b+c into an FMA; explicit
    # fma() is used where a fused operation is intended
    "-ffp-contract=off",
    # no packed-FP32 instructions anywhere
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand)
                     or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def _compile(src, obj, extra_flags, verbose):
    cmd = ([hipcc()] + [f for f in HIPCC_FLAGS if f != "-shared"]
           + list(extra_flags) + ["-c", os.path.join(SRC_DIR, src), "-o",
                                  obj])
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose=False, extra_flags=(), jobs=None):
    """Compiles every source to an object in parallel (one hipcc per file),
    then links the shared library; raises with hipcc's output on failure."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIB_DIR, exist_ok=True)
    obj_dir = os.path.join(LIB_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, s.replace(".hip", ".o")) for s in SOURCES]
    jobs = jobs or min(len(SOURCES), max(1, os.cpu_count() or 1), 16)
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda so: _compile(so[0], so[1], extra_flags,
                                                  verbose),
                              zip(SOURCES, objs)))
    for src, res in zip(SOURCES, results):
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("hipcc failed compiling %s" % src)
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + [
        "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libalgotom_hip.so")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
