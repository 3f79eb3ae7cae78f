"""Build libdiskann_mi355x.so (HIP, gfx950) in-tree.

    python -m diskann_amd.build        # or: python diskann_amd/build.py

Flags: -ffp-contract=off keeps every reduction in the reference's exact order (only
explicit fmaf fuses); packed-fp32 instructions are disabled; f32 denormals are kept.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdiskann_mi355x.so")
SOURCES = ["api.hip", "search.hip", "pq.hip", "build.hip", "exact.hip", "storage.hip", "kmeans.hip",
           "prescore.hip", "sq.hip", "sph.hip", "minmax.hip", "partition.hip"]
ARCH = os.environ.get("DKM_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return [
        "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
        "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
        "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result",
    ]


def build(verbose=False):
    cc = hipcc()
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = [os.path.join(bdir, src.replace(".hip", ".o")) for src in SOURCES]
    cmds = [[cc] + flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
            for src, obj in zip(SOURCES, objs)]
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(subprocess.check_call, cmd) for cmd in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
