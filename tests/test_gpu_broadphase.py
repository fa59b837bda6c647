"""Broadphase sweeps with long intervals, against brute force.

The sweep of a sorted position runs to the first box whose min.x exceeds its
max.x, found by binary search and cut into 64-position chunks.  Scenes here
hold boxes whose x-intervals span many others (a slab under everything, long
rods across the scene, many overlapping boxes on one x), several scenes side
by side (pairs never cross scenes), and boxes that share exact bounds (closed
intervals: touching boxes are a pair).  The pair set must equal brute-force
closed-interval overlap within each scene.
"""
import numpy as np
import pytest

import constraintphysics_amd as cpa
from constraintphysics_amd import scenes

pytestmark = pytest.mark.gpu


def _brute(aabb, scene):
    lo, hi = aabb[:, :3], aabb[:, 3:]
    n = len(aabb)
    out = set()
    for i in range(n):
        m = np.all(lo[i + 1:] <= hi[i], axis=1) & np.all(hi[i + 1:] >= lo[i], axis=1) & (scene[i + 1:] == scene[i])
        for j in np.nonzero(m)[0]:  # body i + 1 + j; ids are indices + 1
            out.add((i + 1, i + 2 + int(j)))
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_long_intervals_match_brute_force(seed):
    rng = np.random.default_rng(seed)
    w = cpa.World(dt=1.0 / 120.0, pgs_mode=0)
    b = scenes.SceneBuilder(w)
    slab, mps = b.box_shape(60.0, 0.5, 60.0, 1.0)
    rod, mpr = b.box_shape(40.0, 0.3, 0.3, 1.0)
    cube, mpc = b.box_shape(0.5, 0.5, 0.5, 1.0)
    mat = (1.0, 0.0, 0.5, 0.5)
    scene_of = []
    for s in range(3):
        b.add(slab, mps, 1, mat, s, np.array([0.0, -0.4, 0.0]))
        scene_of.append(s)
        for r in range(6):
            b.add(rod, mpr, 0, mat, s, np.array([rng.uniform(-5, 5), rng.uniform(-1, 4), rng.uniform(-20, 20)]))
            scene_of.append(s)
        for c in range(300):
            x = 3.0 if c % 4 == 0 else rng.uniform(-25, 25)  # many boxes on one x
            b.add(cube, mpc, 0, mat, s, np.array([x, rng.uniform(-0.2, 3), rng.uniform(-6, 6)]))
            scene_of.append(s)
        # two cubes whose faces touch exactly
        b.add(cube, mpc, 0, mat, s, np.array([30.0, 2.0, 0.0]))
        b.add(cube, mpc, 0, mat, s, np.array([30.5, 2.0, 0.0]))
        scene_of += [s, s]
    b.finish()
    got = {tuple(p) for p in w.broadphase().tolist()}
    want = _brute(w.aabbs(), np.asarray(scene_of))
    print(f"seed {seed}: {len(got)} pairs (brute force {len(want)})", flush=True)
    extra, missing = sorted(got - want), sorted(want - got)
    assert not extra and not missing, (len(extra), extra[:8], len(missing), missing[:8])
    assert len(want) > 200
