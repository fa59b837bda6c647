"""Builds libcp_hip.so (all HIP kernels + host runtime) in-tree for gfx950.

Usage: python -m constraintphysics_amd.build
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcp_hip.so")
SOURCES = ["k_bodies.hip", "k_broadphase.hip", "k_narrowphase.hip", "k_contacts.hip", "k_solver.hip", "k_joints.hip", "world.hip",
           "shapes.cpp"]

# FP64 parity needs IEEE results in the reference's operation order: no FMA
# contraction anywhere.  Packed-FP32 instructions are disabled for all device code.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, "_obj", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmds.append([hipcc] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # one compiler process per translation unit, a few at a time
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "4")), os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
        for f in [ex.submit(subprocess.run, cmd, check=True) for cmd in cmds]:
            f.result()
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
