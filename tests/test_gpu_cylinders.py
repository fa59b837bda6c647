"""Cylinder and irregular-hull narrowphase on the GPU against the oracle, bit
for bit, on seeded random pairs of every order the dispatch distinguishes
(CollisionDetect.cpp:57-152): hull-cylinder, cylinder-hull (normal negated,
MagicID halves swapped), sphere-cylinder, cylinder-sphere, cylinder-cylinder
and hull-hull (8-32-point hulls of the mixed-primitive library and boxes:
the Gauss-arc pair search, its compaction in batches, the wave-parallel
contact-area search and the clipping), including near-parallel and
near-perpendicular axes (cap faces, rims and barrels)."""
import math

import numpy as np
import pytest

import constraintphysics_amd as cpa
from constraintphysics_amd.scenes import BOX_FACES, box_points, hull_library, mass_props_row
from fixtures import Oracle, world_from_description

pytestmark = pytest.mark.gpu

MAT = np.array([1.0, 0.2, 0.5, 0.6])


def _q(rng, snap):
    if snap:  # axis-aligned or near it: exercises cap-face and parallel-barrel paths
        ax = np.eye(3)[rng.integers(0, 3)]
        th = rng.choice([0.0, math.pi / 2, math.pi]) + rng.normal(scale=1e-3)
    else:
        ax, th = rng.normal(size=3), rng.uniform(0, 2 * math.pi)
    ax = ax / np.linalg.norm(ax)
    return np.concatenate([[math.cos(th / 2)], ax * math.sin(th / 2)])


def _body(kind, rng, lib):
    if kind == "cyl":
        r, h = rng.uniform(0.2, 0.5), rng.uniform(0.5, 1.5)
        m, t = cpa.cylinder_mass_properties(r, h)
        return cpa.cylinder_build(r, h), mass_props_row(m, t), math.hypot(r, h / 2)
    if kind == "sphere":
        r = rng.uniform(0.3, 0.7)
        m = 4 / 3 * math.pi * r ** 3
        return dict(sphere=np.array([0, 0, 0, r])), mass_props_row(m, np.eye(3) * 0.4 * m * r * r), r
    if rng.uniform() < 0.5:  # box: face contacts against caps and barrels
        dims = rng.uniform(0.5, 1.2, 3)
        h = cpa.polyhedron_body(box_points(*dims), BOX_FACES, 1.0)
        return h, h["mass_props"], 0.5 * float(np.linalg.norm(dims))
    i = int(rng.integers(0, len(lib.items)))
    return lib.items[i]["body"], lib.mp[i], lib.r[i]


KINDS = [("hull", "cyl"), ("cyl", "hull"), ("sphere", "cyl"), ("cyl", "sphere"), ("cyl", "cyl"), ("hull", "hull")]


def _scene(seed, n_pairs=72):
    rng = np.random.default_rng(seed)
    lib = hull_library(20240, 32)
    kinds = KINDS
    hulls, mps, state, pairs = [], [], [], []
    for p in range(n_pairs):
        ka, kb = kinds[p % len(kinds)]
        snap = p % 3 == 0
        sa, ma, ra = _body(ka, rng, lib)
        sb, mb, rb = _body(kb, rng, lib)
        base = np.array([12.0 * p, 5.0, 0.0])
        d = rng.normal(size=3)
        if snap:
            d = np.eye(3)[rng.integers(0, 3)] + rng.normal(scale=0.05, size=3)
        d *= rng.uniform(0.3, 1.05) * (ra + rb) / np.linalg.norm(d)
        for sh, mp, pos in ((sa, ma, base), (sb, mb, base + d)):
            hulls.append(sh)
            mps.append(mp)
            state.append(np.concatenate([pos, _q(rng, snap), [0, 0, 0], [0, 0, 0]]))
        pairs.append((len(hulls) - 1, len(hulls)))
    n = len(hulls)
    desc = dict(hulls=hulls, movement=np.zeros(n, int), mass_props=np.array(mps), material=np.tile(MAT, (n, 1)), state=np.array(state))
    return desc, np.array(pairs)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cylinder_pairs_bit_exact(seed):
    desc, pairs = _scene(seed)
    o = Oracle.from_description(desc, dt=1 / 120, gravity_y=-6.0, vel_iters=10, pos_iters=1, substeps=8)
    w = world_from_description(desc, pgs_mode=1)
    counts, heads, pts, mg = w.narrowphase(pairs)
    hit = 0
    hits = {k: 0 for k in KINDS}
    for i, (a, b) in enumerate(pairs):
        head, opts, omg = o.narrowphase(int(a), int(b))
        assert counts[i] == len(opts), (seed, i, counts[i], len(opts))
        if len(opts) == 0:
            continue
        hit += 1
        hits[KINDS[i % len(KINDS)]] += 1
        assert np.array_equal(heads[i], head), (seed, i)
        assert np.array_equal(pts[i, :len(opts)], opts), (seed, i)
        assert np.array_equal(mg[i, :len(opts)], omg), (seed, i)
    print(f"seed {seed}: contacts per kind {hits}", flush=True)
    assert hit >= 24
    assert hits[("hull", "hull")] >= 5


def test_cylinder_pile_full_step_matches_oracle():
    """A full reference-order step of the cylinder fixture's last recorded state."""
    from fixtures import golden, world_from_fixture
    fx = golden("cyl4")
    k = 149
    w = world_from_fixture(fx, pgs_mode=1)
    w.set_state(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"])
    w.set_contacts(*fx.cache(k, "pre"))
    w.set_node_order(fx[f"s{k}_node_order"])
    w.step(1)
    err = np.max(np.abs(w.state() - fx[f"s{k}_post"]) / np.maximum(1.0, np.abs(fx[f"s{k}_post"])))
    assert err <= 1e-12, err
