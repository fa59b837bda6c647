"""Snapshot/resume, trajectory recording and serving evict/readmit through
the C ABI.

A snapshot (state, world inertia, contact cache, joint state; the file format
of cp_world_save_snapshot) resumes a world bit for bit: the steps after a
load equal the steps the saved world took.  A recorded trajectory holds, per
step, every body's pose in the reference's recording layout (7 doubles per
body, ImageDemo.h:98-133).  A scene evicted from a serving world and admitted
again later — into another body range — continues exactly as the same scene
stepped alone without interruption.
"""
import os

import numpy as np
import pytest

import constraintphysics_amd as cpa
from constraintphysics_amd import scenes, serving
from fixtures import golden, world_from_fixture

pytestmark = pytest.mark.gpu

CFG = dict(dt=1.0 / 120.0, gravity=(0.0, -6.0, 0.0), vel_iters=10, pos_iters=1, substeps=8, pgs_mode=0)


def _batch(precision=1):
    sizes = scenes.c3_sizes(10, 61, lo=8, hi=400)
    w = cpa.World(precision=precision, **CFG)
    scenes.build_batched(w, sizes, 61, hull_lib_size=32)
    return w, sizes


@pytest.mark.parametrize("precision", [1, 0])
def test_snapshot_file_resumes_bit_for_bit(tmp_path, precision):
    w, sizes = _batch(precision)
    w.step(60)
    path = str(tmp_path / "snap.cps")
    w.save_snapshot(path)
    w.step(25)
    a, ca = w.state(), w.contacts()
    w.close()
    v, _ = _batch(precision)
    v.load_snapshot(path)
    v.step(25)
    b, cb = v.state(), v.contacts()
    assert np.array_equal(a, b)
    assert all(np.array_equal(x, y) for x, y in zip(ca, cb))
    # a world of other bodies refuses the file; a corrupted file is refused
    u = cpa.World(**CFG)
    scenes.build_batched(u, sizes[:3], 61, hull_lib_size=32)
    with pytest.raises(cpa.CpError):
        u.load_snapshot(path)
    u.close()
    raw = bytearray(open(path, "rb").read())
    raw[100] ^= 0xFF
    bad = str(tmp_path / "bad.cps")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(cpa.CpError):
        v.load_snapshot(bad)
    v.close()


def test_snapshot_resumes_joints_and_holonomic_systems(tmp_path):
    """The jointed-chain fixture in the reference-order mode (joint impulses
    and hinge bases are part of the snapshot)."""
    fx = golden("chain3")
    k = fx.recorded[-1]
    w = world_from_fixture(fx, pgs_mode=1)
    w.set_state(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"])
    w.set_contacts(*fx.cache(k, "pre"))
    w.set_joint_state(fx[f"s{k}_joint_pre"])
    w.step(5)
    path = str(tmp_path / "chain.cps")
    w.save_snapshot(path)
    w.step(10)
    a, ja = w.state(), w.joint_state()
    v = world_from_fixture(fx, pgs_mode=1)
    v.load_snapshot(path)
    v.step(10)
    assert np.array_equal(a, v.state()) and np.array_equal(ja, v.joint_state())


def test_trajectory_records_every_step(tmp_path):
    w, _ = _batch()
    w.step(20)
    path = str(tmp_path / "run.traj")
    w.trajectory_open(path)
    states = []
    for _ in range(6):
        w.step(1)
        states.append(w.state())
    w.trajectory_close()
    w.step(1)  # not recorded
    n = len(states[0])
    assert cpa.trajectory_frames(path, n) == 6
    for f, st in enumerate(states):
        assert np.array_equal(cpa.trajectory_read(path, n, f), st[:, :7])
    w.close()


@pytest.mark.parametrize("precision", [1, 0])
def test_evicted_scene_resumes_exactly(precision):
    seed = 17
    reqs, sizes = serving.poisson_requests(3, rate=1e6, seed=seed, max_steps=5, lo=150, hi=400)
    w = cpa.World(precision=precision, **CFG)
    w.serving_init(3000)
    src = serving.SceneSource(w, sizes, seed, hull_lib_size=32)
    pay = [src.payload(i) for i in range(3)]
    ha = w.admit_scene(*pay[0][:5])
    hb = w.admit_scene(*pay[1][:5])
    w.step(40)
    snap = w.scene_snapshot(ha)  # evict scene 0 mid-flight
    assert len(snap[2][0]) > 0
    w.retire_scene(ha)
    hc = w.admit_scene(*pay[2][:5])  # another scene takes the freed range
    w.step(7)
    ha2 = w.admit_scene_resume(*pay[0][:4], snap)
    assert ha2 is not None
    w.step(30)
    resumed = w.scene_state(ha2)
    for h in (hb, hc, ha2):
        w.retire_scene(h)
    w.close()
    a = cpa.World(precision=precision, **CFG)
    scenes.build_batched(a, sizes, seed, scene_indices=[0], hull_lib_size=32)
    a.step(70)
    alone = a.state()
    a.close()
    assert np.array_equal(resumed, alone)
