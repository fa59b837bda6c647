"""Sleeping in the CPU oracle, pinned to the compiled reference.

The `sleep` fixture runs the reference with sleeping enabled (its default
parameters: sleep_delay 0.5 s, vel_sleep_coeff 0.1, ang_vel_eps 0.1,
accel_sleep_coeff 0.044, non_sleepy_tick_threshold 3; PhysicsEngine.h:218-224)
for 420 steps: a 2x2x2 cube pile, a lone cube and a sliding cube fall asleep
(steps 153, 135 and 220), a column of three cubes falls asleep at step 136
and wakes at step 314 when a sphere dropped from high above lands on it; a
waking action on a pile cube (RigidBody::alertWakingAction, before step 250)
wakes the pile, which falls asleep again at step 339.
Every step's sleep state (asleep, sleep_ready_counter, non_sleepy_tick_count,
history length) and the recorded steps' states, contact caches and island
rows must match bit for bit, from the initial state on."""
import numpy as np
import pytest

from fixtures import Oracle, golden


@pytest.fixture(scope="module")
def run():
    fx = golden("sleep")
    o = Oracle(fx)
    o.set_sleeping(True)
    o.load(fx["initial_state"])
    order = fx["s0_node_order"]
    out = {"trace": [], "post": {}, "cache": {}, "islands": {}}
    rec = set(fx.recorded)
    for k in range(int(fx.meta["steps"])):
        if k == int(fx.meta["wake_step"]):
            o.wake_bodies([int(fx.meta["wake_body"])])
        o.step(order)
        out["trace"].append(o.sleep_state())
        if k in rec:
            out["post"][k] = o.state()
            out["cache"][k] = o.cache()
            out["islands"][k] = o.islands()
    return fx, out


def test_sleep_trace_matches_reference(run):
    fx, out = run
    tr = fx["sleep_trace"]
    for k in range(tr.shape[0]):
        assert np.array_equal(out["trace"][k], tr[k]), f"step {k}: {np.argwhere(out['trace'][k] != tr[k])[:4]}"


def test_scene_sleeps_and_wakes(run):
    """The scenario the fixture is built for actually happens."""
    fx, _ = run
    asleep = fx["sleep_trace"][:, :, 0]
    assert asleep[152, 1:9].sum() == 0 and asleep[153, 1:9].all()  # the pile
    assert asleep[135, 9] and asleep[220, 10] and not asleep[219, 10]  # the lone and the sliding cube
    assert asleep[200, 11:14].all() and not asleep[314, 11:14].any()  # the column, woken by the sphere
    assert not asleep[:, 14].any()  # the sphere never sleeps
    assert asleep[249, 1:9].all() and not asleep[250, 1:9].any() and asleep[339, 1:9].all()  # the pile woken by an action


@pytest.mark.parametrize("k", [100, 135, 136, 153, 154, 200, 220, 221, 250, 251, 300, 313, 314, 315, 330, 339, 419])
def test_recorded_steps_match_reference(run, k):
    fx, out = run
    assert np.array_equal(out["post"][k], fx[f"s{k}_post"])
    ci, cm, cf = out["cache"][k]
    gi, gm, gf = fx.cache(k, "post")
    assert np.array_equal(ci[:, [0, 1, 2, 3, 5]], gi[:, [0, 1, 2, 3, 5]]) and np.array_equal(cm, gm) and np.array_equal(cf, gf)
    rows, offs = out["islands"][k]
    assert np.array_equal(rows, fx[f"s{k}_island_rows"])
    assert np.array_equal(offs, fx[f"s{k}_island_off"])
