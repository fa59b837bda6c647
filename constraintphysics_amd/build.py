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
SOURCES = ["k_bodies.hip", "k_broadphase.hip", "k_narrowphase.hip", "k_contacts.hip", "k_solver.hip", "k_joints.hip", "k_raycast.hip", "k_islands.hip", "world.hip",
           "shapes.cpp"]

# FP64 parity needs IEEE results in the reference's operation order: no FMA
# contraction anywhere.  Packed-FP32 instructions are disabled for all device code.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _stale(src, obj):
    """True when obj is missing or older than its source or any header in csrc/."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose=False, incremental=False):
    """Compiles every translation unit (incremental=True: only those older
    than their source or a header) and links libcp_hip.so."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, "_obj", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        if not incremental or _stale(os.path.join(CSRC, src), obj):
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
    build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv)
