"""Holonomic chains on a heightfield (the C2 configuration) on the GPU.

A small instance (6 chains of 16 cylinder links joined by ball-sockets and one
hinge each, on a 40 x 40 heightfield) is stepped through the fall onto the
terrain and compared with the oracle step by step (reference-order PGS, the
holonomic block solve, mesh contacts).  The full-size instance (512 chains on
256 x 256 cells) is checked through properties that hold at any size: finite
state, joints holding their attach points together, no link below the
terrain, every chain its own holonomic system of degree 47.
"""
import numpy as np
import pytest

import constraintphysics_amd as cpa
from constraintphysics_amd import scenes
from fixtures import Oracle, world_from_description

pytestmark = pytest.mark.gpu

CFG = dict(dt=1.0 / 120.0, gravity=(0.0, -6.0, 0.0), vel_iters=10, pos_iters=1, substeps=8, pgs_mode=1, hol_iters=1)


def _systems(w):
    sy = w.holonomic_systems()
    off = np.concatenate([[0], np.cumsum([len(s[0]) for s in sy])]).astype(np.int64)
    og = np.concatenate([s[0] for s in sy]).astype(np.int64)
    return off, og


def test_small_chains_on_terrain_match_oracle():
    desc = scenes.c2_description(n_chains=6, links=16, grid=40, cell=0.5, amp=1.0, seed=7)
    n = len(desc["hulls"])
    no = cpa.reference_node_order(n)
    probe = world_from_description(desc, **CFG)
    probe.set_node_order(no)
    probe.step(1)
    off, og = _systems(probe)
    probe.close()
    assert len(off) == 7 and all(np.diff(off) == 15)
    w = world_from_description(desc, **CFG)
    w.set_holonomic_systems(off, og)
    w.set_node_order(no)
    o = Oracle.from_description(desc, dt=CFG["dt"], gravity_y=-6.0, vel_iters=10, pos_iters=1, substeps=8, hol_iters=1)
    o.load_joints(np.zeros((len(desc["joints"][0]), 12)), off, og)
    worst, touched = 0.0, False
    for k in range(150):
        w.step(1)
        o.step(no)
        a, b = w.state(), o.state()
        worst = max(worst, float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)))))
        touched = touched or len(w.contacts()[0]) > 0
    print(f"150 steps, worst state rel {worst:.2e}, contacts at the end {len(w.contacts()[0])}", flush=True)
    assert touched
    assert worst <= 1e-12


def test_small_chains_island_colored_replayed_by_oracle():
    """The coloured-island solve through the fall: each step equals the
    oracle's replay of the step's recorded row order."""
    desc = scenes.c2_description(n_chains=6, links=16, grid=40, cell=0.5, amp=1.0, seed=7)
    n = len(desc["hulls"])
    no = cpa.reference_node_order(n)
    cfg = dict(CFG, pgs_mode=2)
    probe = world_from_description(desc, **cfg)
    probe.set_node_order(no)
    probe.step(1)
    off, og = _systems(probe)
    probe.close()
    w = world_from_description(desc, **cfg)
    w.set_holonomic_systems(off, og)
    w.set_node_order(no)
    o = Oracle.from_description(desc, dt=CFG["dt"], gravity_y=-6.0, vel_iters=10, pos_iters=1, substeps=8, hol_iters=1)
    o.load_joints(np.zeros((len(desc["joints"][0]), 12)), off, og)
    worst = 0.0
    for k in range(150):
        w.step(1)
        o.step_scheduled(*w.last_islands())
        a, b = w.state(), o.state()
        worst = max(worst, float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)))))
    print(f"150 steps coloured islands, worst state rel {worst:.2e}, contacts {len(w.contacts()[0])}", flush=True)
    assert len(w.contacts()[0]) > 0
    assert worst <= 1e-12


def test_small_chains_island_colored_split_contacts():
    """Coloured islands with the split precision (FP32 contact rows; joint
    rows, the holonomic block solve and the deltas FP64) through the fall:
    each step, from the GPU's own state, against the oracle's FP64 replay of
    the recorded row order.  Bars: velocities within 1e-4 and poses within
    1e-6 relative (scale max(1, |x|))."""
    desc = scenes.c2_description(n_chains=6, links=16, grid=40, cell=0.5, amp=1.0, seed=7)
    n = len(desc["hulls"])
    no = cpa.reference_node_order(n)
    cfg = dict(CFG, pgs_mode=2, precision=1)
    probe = world_from_description(desc, **cfg)
    probe.set_node_order(no)
    probe.step(1)
    off, og = _systems(probe)
    probe.close()
    w = world_from_description(desc, **cfg)
    w.set_holonomic_systems(off, og)
    w.set_node_order(no)
    o = Oracle.from_description(desc, dt=CFG["dt"], gravity_y=-6.0, vel_iters=10, pos_iters=1, substeps=8, hol_iters=1)
    rel = lambda a, b: float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)))) if a.size else 0.0
    wv = wp = 0.0
    for k in range(120):
        pre = (w.state(), w.inertia(), w.contacts(), w.joint_state())
        w.step(1)
        if k < 60 and k % 10:  # check every tenth step through the free fall, every step once on the terrain
            continue
        o.load(pre[0], pre[1], pre[2])
        o.load_joints(pre[3], off, og)
        o.step_scheduled(*w.last_islands())
        a, b = w.state(), o.state()
        wp = max(wp, rel(a[:, :7], b[:, :7]))
        wv = max(wv, rel(a[:, 7:], b[:, 7:]))
    print(f"split contacts in coloured islands: worst pose rel {wp:.2e}, velocity rel {wv:.2e}, contacts {len(w.contacts()[0])}",
          flush=True)
    assert len(w.contacts()[0]) > 0
    assert wv <= 1e-4 and wp <= 1e-6


def _attach_error(state, joints):
    info, geo = joints
    err = 0.0
    for (kind, i1, i2), g in zip(info, geo):
        p = []
        for idx, tp in ((i1 - 1, g[0:3]), (i2 - 1, g[3:6])):
            r, qi, qj, qk = state[idx, 3:7]
            R = np.array([[1 - 2 * (qj * qj + qk * qk), 2 * (qi * qj - qk * r), 2 * (qi * qk + qj * r)],
                          [2 * (qi * qj + qk * r), 1 - 2 * (qi * qi + qk * qk), 2 * (qj * qk - qi * r)],
                          [2 * (qi * qk - qj * r), 2 * (qj * qk + qi * r), 1 - 2 * (qi * qi + qj * qj)]])
            p.append(state[idx, :3] + R @ tp)
        err = max(err, float(np.linalg.norm(p[0] - p[1])))
    return err


@pytest.mark.parametrize("mode", [1, 2])
def test_full_size_chains_on_terrain_properties(mode):
    desc = scenes.c2_description()
    w = world_from_description(desc, **dict(CFG, pgs_mode=mode))
    for k in range(120):
        w.step(1)
    st = w.state()
    sy = w.holonomic_systems()
    types = desc["joints"][0][:, 0]
    assert len(sy) == 512 and all(len(s[0]) == 15 for s in sy)
    assert all(int(types[s[0]].sum()) == 47 for s in sy)  # 14 ball-sockets and one hinge per chain
    assert np.all(np.isfinite(st))
    err = _attach_error(st, desc["joints"])
    pts = desc["hulls"][0]["mesh"]
    print(f"full size: {len(w.contacts()[0])} contacts, max attach error {err:.3e}, min link y {st[1:, 1].min():.3f}", flush=True)
    assert err < 0.05
    assert st[1:, 1].min() > pts[:, 1].min() - 0.2
    assert len(w.contacts()[0]) > 1000
