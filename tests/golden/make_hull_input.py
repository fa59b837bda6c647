"""Writes the irregular-hull input of the `hulls` fixture: hulls of 8-32
seeded points on ellipsoids with semi-axes U[0.3, 0.8] (the mixed-primitive
library's generator, constraintphysics_amd.scenes.random_hull), as text the
reference harness reads (count; per hull "n_pts n_faces", the points with 17
significant digits, each face as its vertex count and indices).

Usage: python make_hull_input.py OUT [COUNT] [SEED]
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from constraintphysics_amd.scenes import random_hull  # noqa: E402


def main():
    out = sys.argv[1]
    count = int(sys.argv[2]) if len(sys.argv) > 2 else 24
    seed = int(sys.argv[3]) if len(sys.argv) > 3 else 20240
    rng = np.random.default_rng(seed)
    lines = [str(count)]
    for _ in range(count):
        pts, faces = random_hull(rng)
        lines.append(f"{len(pts)} {len(faces)}")
        lines += [" ".join(repr(float(x)) for x in p) for p in pts]
        lines += [" ".join(str(int(x)) for x in [len(f)] + list(f)) for f in faces]
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
