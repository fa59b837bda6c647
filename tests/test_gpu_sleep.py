"""Sleeping on the GPU (cp_config.sleeping) against the compiled reference.

The `sleep` fixture is a reference run with sleeping on for 420 steps: a
cube pile, a lone cube and a sliding cube fall asleep, a sphere dropped on a
sleeping column of three cubes wakes it, and a waking action on a pile cube
before step 250 wakes the pile (RigidBody::alertWakingAction).  In the
reference-order mode the device path must reproduce every step's sleep
state (asleep, sleep_ready_counter, non_sleepy_tick_count, history length)
and the recorded steps' states, contact caches and island rows bit for bit,
from the initial state on.  The coloured-island mode is checked step by step
against the oracle's replay of its recorded row order (the oracle itself is
pinned to the same fixture in test_oracle_sleep.py)."""
import numpy as np
import pytest

from fixtures import Oracle, golden, world_from_fixture

pytestmark = pytest.mark.gpu


def _rel(a, b):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    return d / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def fx():
    return golden("sleep")


def test_reference_order_sleep_trace_bit_exact(fx):
    print("sleep fixture, reference order, 420 steps", flush=True)
    w = world_from_fixture(fx, pgs_mode=1, sleeping=1)
    w.set_state(fx["initial_state"])
    tr = fx["sleep_trace"]
    rec = set(fx.recorded)
    for k in range(int(fx.meta["steps"])):
        if k == int(fx.meta["wake_step"]):
            w.wake_bodies([int(fx.meta["wake_body"])])
        w.step(1)
        got = w.sleep_state()
        assert np.array_equal(got, tr[k]), f"step {k}: bodies {np.unique(np.argwhere(got != tr[k])[:, 0])}"
        if k not in rec:
            continue
        assert np.array_equal(w.state(), fx[f"s{k}_post"]), f"state at step {k}"
        ci, cm, cf = w.contacts()
        gi, gm, gf = fx.cache(k, "post")
        cols = [0, 1, 2, 3, 5]
        assert np.array_equal(ci[:, cols], gi[:, cols]) and np.array_equal(cm, gm) and np.array_equal(cf, gf), f"cache at step {k}"
        rows, offs = w.last_islands()
        assert np.array_equal(rows, fx[f"s{k}_island_rows"]) and np.array_equal(offs, fx[f"s{k}_island_off"]), f"islands at step {k}"
    w.close()


def test_colored_islands_with_sleeping_replayed_by_oracle(fx):
    print("sleep fixture, coloured islands vs oracle replay, 420 steps", flush=True)
    w = world_from_fixture(fx, pgs_mode=2, sleeping=1)
    w.set_state(fx["initial_state"])
    o = Oracle(fx)
    o.set_sleeping(True)
    o.load(fx["initial_state"])
    worst = 0.0
    for k in range(int(fx.meta["steps"])):
        if k == int(fx.meta["wake_step"]):
            w.wake_bodies([int(fx.meta["wake_body"])])
            o.wake_bodies([int(fx.meta["wake_body"])])
        w.step(1)
        rows, offs = w.last_islands()
        o.step_scheduled(rows, offs)
        worst = max(worst, _rel(w.state(), o.state()))
        assert np.array_equal(w.sleep_state()[:, 0], o.sleep_state()[:, 0]), f"asleep flags at step {k}"
    print(f"  worst state rel {worst:.2e}", flush=True)
    assert worst <= 1e-12
    # the run passes through the same sleep and wake events as the reference's
    asleep = o.sleep_state()[:, 0]
    assert asleep[1:9].all() and not asleep[14]
    w.close()


def test_sleeping_needs_an_island_mode(fx):
    import constraintphysics_amd as cpa
    w = world_from_fixture(fx, pgs_mode=0, sleeping=1)
    w.set_state(fx["initial_state"])
    with pytest.raises(cpa.CpError):
        w.step(1)
    w.close()


def test_sleep_disabled_body_keeps_its_island_awake(fx):
    """RigidBody::setSleepDisabled: a body that is never sleepy keeps its
    island awake (the oracle, with the same flag, agrees bit for bit)."""
    w = world_from_fixture(fx, pgs_mode=1, sleeping=1)
    w.set_state(fx["initial_state"])
    w.set_sleep_disabled([9])  # the lone cube
    o = Oracle(fx)
    o.set_sleeping(True)
    o.load(fx["initial_state"])
    idx = np.array([9], np.int64)
    o.lib.orc_set_sleep_disabled(o.w, 1, idx.ctypes.data, 1)
    for k in range(200):
        w.step(1)
        o.step(fx["s0_node_order"])
    assert np.array_equal(w.sleep_state(), o.sleep_state())
    assert np.array_equal(w.state(), o.state())
    s = w.sleep_state()
    assert s[9, 0] == 0 and s[9, 1] == 0 and s[1:9, 0].all()
    w.close()


def test_snapshot_resumes_sleep_state(fx, tmp_path):
    """A snapshot taken mid-run (bodies asleep, histories part full) resumes
    bit for bit: the loaded world then follows the reference's sleep trace,
    through the column's wake at step 314."""
    w = world_from_fixture(fx, pgs_mode=1, sleeping=1)
    w.set_state(fx["initial_state"])
    for k in range(300):
        if k == int(fx.meta["wake_step"]):
            w.wake_bodies([int(fx.meta["wake_body"])])
        w.step(1)
    path = str(tmp_path / "sleep.cps")
    w.save_snapshot(path)
    w.close()
    v = world_from_fixture(fx, pgs_mode=1, sleeping=1)
    v.set_state(fx["initial_state"])
    v.load_snapshot(path)
    tr = fx["sleep_trace"]
    assert np.array_equal(v.sleep_state(), tr[299])
    for k in range(300, int(fx.meta["steps"])):
        v.step(1)
        assert np.array_equal(v.sleep_state(), tr[k]), f"step {k}"
    assert np.array_equal(v.state(), fx["final_state"])
    v.close()
