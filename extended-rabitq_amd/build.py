"""Compile the native libraries in-tree.

  lib/liberq.so          HIP search library for gfx950 (C-ABI in include/erq.h)
  lib/liberq_builder.so  host index quantizer (g++ + OpenMP)

Run directly (``python extended-rabitq_amd/build.py``) or through
``__graft_entry__.build()``.  Every call recompiles.  The HIP sources are compiled to
objects in parallel (each translation unit's kernels are self-contained), then linked.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "_build")
INCLUDE = os.path.join(REPO, "include")

HIP_SOURCES = ["erq_capi.hip", "erq_search.hip", "erq_gemm.hip", "erq_probe.hip", "erq_batch.hip",
               "erq_mlp.hip", "erq_knn.hip", "erq_build.hip", "erq_nonha.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-munsafe-fp-atomics",
    "-ffp-contract=off",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _hipcc():
    return os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")


def build_builder():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "liberq_builder.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-march=x86-64-v2",
          "-o", out + ".tmp", os.path.join(CSRC, "builder.cpp")])
    os.replace(out + ".tmp", out)
    return out


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(LIB, "liberq.so")
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJ, s.replace(".hip", ".o")) for s in srcs]

    def compile_one(pair):
        src, obj = pair
        _run([_hipcc()] + HIPCC_FLAGS + ["-I", INCLUDE, "-I", CSRC, "-c", "-o", obj,
                                         os.path.join(CSRC, src)])

    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", "0") or 0) or os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    _run([_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    oracle = os.path.join(REPO, "oracle")
    _run(["make", "-C", oracle, "-B"])


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    what = set(argv) or {"builder", "hip", "oracle"}
    if "builder" in what:
        build_builder()
    if "oracle" in what:
        build_oracle()
    if "hip" in what:
        build_hip()


if __name__ == "__main__":
    main()
