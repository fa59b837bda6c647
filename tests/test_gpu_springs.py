"""Springs and per-joint CFM on the GPU, through the C ABI.

Scenes: tests/golden/springs and springs_h0, dumped from the compiled
reference engine: a box hanging from a fixed anchor by a spring whose resting
length is the distance at creation, a second box hanging from it by a spring
of resting length 1.0 that lands on the ground, two boxes on the ground pushed
apart by a compressed spring (a spring pair that also collides), a three-link
ball-socket chain (one holonomic system) whose middle joint has a custom CFM,
and a hinged block with a custom CFM.  The oracle reproduces both runs bit for
bit (test_oracle.py).  Bars: island rows exact; body state within 1e-12
relative and joint state within 1e-10 relative (scale max(1, |x|)) for single
steps; whole runs within 1e-9 / 1e-8.
"""
import numpy as np
import pytest

import constraintphysics_amd as cpa
from fixtures import Oracle, golden, world_from_fixture

pytestmark = pytest.mark.gpu

# springs step 0 founds the holonomic system inside the step; the dump's
# systems are those before it, so that step is covered by the whole run
STEPS = [("springs", k) for k in (1, 10, 50, 99, 100, 150, 199)] + [("springs_h0", k) for k in (0, 1, 50, 100, 199)]


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def _node_order(fx, k):
    no = fx[f"s{k}_node_order"]  # dumped only for steps with a collision
    return no if len(no) else cpa.reference_node_order(int(fx.meta["n_bodies"]))


def _load(w, fx, k):
    w.set_state(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"])
    w.set_contacts(*fx.cache(k, "pre"))
    w.set_holonomic_systems(fx[f"s{k}_hsys_off"], fx[f"s{k}_hsys_og"])
    w.set_joint_state(fx[f"s{k}_joint_pre"])
    w.set_node_order(_node_order(fx, k))


@pytest.mark.parametrize("name,k", STEPS)
def test_step_matches_reference(name, k):
    fx = golden(name)
    print(f"springs step {name} {k}", flush=True)
    w = world_from_fixture(fx, pgs_mode=1)
    _load(w, fx, k)
    w.step(1)
    rows, offs = w.last_islands()
    if len(fx[f"s{k}_island_rows"]):  # the dump holds islands only for steps with a collision
        assert np.array_equal(rows, fx[f"s{k}_island_rows"]) and np.array_equal(offs, fx[f"s{k}_island_off"])
    err = _rel(w.state(), fx[f"s{k}_post"])
    jerr = _rel(w.joint_state(), fx[f"s{k}_joint_post"])
    print(f"  state rel {err:.2e}, joint state rel {jerr:.2e}", flush=True)
    assert err <= 1e-12, err
    assert jerr <= 1e-10, jerr
    w.close()


@pytest.mark.parametrize("name", ["springs", "springs_h0"])
def test_whole_run_from_the_initial_state(name):
    """200 steps from the initial state: the world measures the first
    spring's resting length itself (given as -1: the distance at creation)."""
    fx = golden(name)
    print(f"springs whole run {name}", flush=True)
    w = world_from_fixture(fx, pgs_mode=1)
    w.set_state(fx["initial_state"])
    js0 = w.joint_state()
    assert _rel(js0, fx["s0_joint_pre"]) <= 1e-12
    spring = fx["joint_info"][:, 0] == 13
    assert np.array_equal(js0[spring, 6], fx["s0_joint_pre"][spring, 6])  # resting lengths bit for bit
    rec = set(fx.recorded)
    worst = jworst = 0.0
    for k in range(int(fx.meta["steps"])):
        w.step(1)
        if k in rec:
            worst = max(worst, _rel(w.state(), fx[f"s{k}_post"]))
            jworst = max(jworst, _rel(w.joint_state(), fx[f"s{k}_joint_post"]))
    print(f"  worst state rel {worst:.2e}, joint state rel {jworst:.2e}", flush=True)
    assert worst <= 1e-9 and jworst <= 1e-8
    w.close()


def test_custom_cfm_rows_match_oracle():
    """Rows of every joint after a step (the custom-CFM ball-socket and hinge
    among them) against the oracle's, and against a world whose joints all
    keep the global CFM: the custom value must change exactly those rows'
    inverse blocks."""
    fx = golden("springs_h0")
    k = 50
    w = world_from_fixture(fx, pgs_mode=1)
    _load(w, fx, k)
    w.step(1)
    o = Oracle(fx)
    o.load(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"], fx.cache(k, "pre"))
    o.load_joints(fx[f"s{k}_joint_pre"], fx[f"s{k}_hsys_off"], fx[f"s{k}_hsys_og"])
    o.step(_node_order(fx, k))
    cfm = fx["joint_cfm"]
    for j in range(len(fx["joint_info"])):
        if fx["joint_info"][j, 0] == 13:
            continue
        g, r = w.joint_rows(j), o.joint_rows(j)
        worst = max(_rel(a, b) for a, b in zip(g, r))
        assert worst <= 1e-10, (j, worst)
    # the same step with every joint on the global CFM: only the state of the
    # custom-CFM joints' scene parts changes
    w2 = world_from_fixture(fx, pgs_mode=1)
    for j in np.flatnonzero(~np.isnan(cfm)):
        w2.set_joint_cfm(int(j), float("nan"))
    _load(w2, fx, k)
    w2.step(1)
    assert not np.array_equal(w.joint_state(), w2.joint_state())
    w.close()
    w2.close()


def test_colored_islands_replayed_by_oracle():
    """Coloured islands over 60 steps: each step equals the oracle's
    sequential replay of the recorded row order, springs applied in the
    constraint graph's edge order at every substep on both sides."""
    fx = golden("springs")
    k = 10
    w = world_from_fixture(fx, pgs_mode=2)
    _load(w, fx, k)
    o = Oracle(fx)
    o.load(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"], fx.cache(k, "pre"))
    o.load_joints(fx[f"s{k}_joint_pre"], fx[f"s{k}_hsys_off"], fx[f"s{k}_hsys_og"])
    o.set_node_order(_node_order(fx, k))
    worst = jworst = 0.0
    for s in range(k + 1, k + 61):
        w.step(1)
        rows, offs = w.last_islands()
        o.step_scheduled(rows, offs)
        worst = max(worst, _rel(w.state(), o.state()))
        jworst = max(jworst, _rel(w.joint_state(), o.joint_state()))
    print(f"  60 steps: worst state rel {worst:.2e}, joints {jworst:.2e}", flush=True)
    assert worst <= 1e-12 and jworst <= 1e-10
    w.close()


def test_springs_refuse_sleeping_and_bad_cfm():
    fx = golden("springs")
    w = world_from_fixture(fx, pgs_mode=1, sleeping=1)
    w.set_state(fx["initial_state"])
    with pytest.raises(cpa.CpError):
        w.step(1)
    w.close()
    w = world_from_fixture(fx, pgs_mode=1)
    with pytest.raises(cpa.CpError):
        w.set_joint_cfm(0, -1.0)
    with pytest.raises(cpa.CpError):
        w.set_joint_cfm(99, 1e-3)
    w.close()
