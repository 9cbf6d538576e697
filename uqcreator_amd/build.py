"""Build recipe of the native library: hipcc for gfx950, in-tree output.

    python -m uqcreator_amd.build        # builds uqcreator_amd/_build/libuqk.so

Each .hip translation unit is compiled to an object (in parallel), then linked
into one shared library that exports the C ABI of include/uqk.h.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libuqk.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off (they have given silent wrong results)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in VGPRs (not AGPRs): the accumulator registers are
    # reused by the other phases, which keeps the closure kernels at 2+ waves/SIMD
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    # the closure's factorisation is one fully unrolled loop nest over
    # register-resident tiles; the default limit would leave it rolled (and
    # the tiles in scratch memory)
    "-mllvm", "-pragma-unroll-threshold=500000",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


# extra flags per source file
EXTRA_FLAGS = {
    # the one-wavefront closure is a chain of dependent steps (pivot
    # broadcasts, Newton steps); LLVM's iterative ILP scheduler orders it
    # better than the default occupancy-driven one (about 3% more throughput
    # for the mixed closure, the same for FP64); closure_wave_reuse.hip keeps
    # the default one, which spills less there
    "closure_wave.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"],
    "closure_wave_lag.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"],
}


def sources():
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    if os.path.isdir(HOST):
        srcs += sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    return srcs


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    cmd = [HIPCC] + HIP_FLAGS + EXTRA_FLAGS.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "c++"] + [f for f in HIP_FLAGS if not f.startswith("--offload")] + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
