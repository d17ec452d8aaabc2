"""Build step of the package: compiles the HIP library for gfx950 in-tree."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsphg.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wno-unused-value",
    # packed FP32 instructions stay off: some of their forms have given
    # intermittent wrong results on gfx950
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the scheduler that groups memory clauses: the stage kernels wait on
    # partner loads, and grouping them was measured about 1% faster for the
    # FP64 step (same registers and occupancy; instruction order only)
    "-mllvm", "-amdgpu-sched-strategy=max-memory-clause",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build_hip(verbose=False):
    src = os.path.join(CSRC, "sphg_api.hip")
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + [src, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for sphg_api.hip")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_hip(verbose=True))
