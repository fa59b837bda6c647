"""Cone and twist limits on the GPU, through the C ABI.

Scenes: tests/golden/limits and limits_h0, dumped from the compiled reference
engine: a two-link ball-socket pendulum under cone limits (0.15 and 0.12 rad)
and a hinged block under a twist limit (-0.3, 0.4 rad), kicked so that the
limits engage and release; with and without the holonomic block solver.  The
oracle reproduces both runs bit for bit (test_oracle.py).  Bars: island rows,
the limits' is_inactive flags and the holonomic systems exact; body state
within 1e-12 relative and joint state (impulses, the twist's tracked angle)
within 1e-10 relative (scale max(1, |x|)) for single steps; whole runs from
the initial state within 1e-9 / 1e-8 with every recorded flag equal.
"""
import numpy as np
import pytest

import constraintphysics_amd as cpa
from fixtures import Oracle, golden, world_from_fixture

pytestmark = pytest.mark.gpu

STEPS = [("limits", k) for k in (1, 10, 14, 15, 16, 20, 30, 32, 57, 149)] + [("limits_h0", k) for k in (0, 15, 30, 57, 149)]


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


def _limits(fx):
    return np.nonzero(fx["joint_info"][:, 0] >= 8)[0]


@pytest.mark.parametrize("name,k", STEPS)
def test_step_matches_reference(name, k):
    fx = golden(name)
    print(f"limits step {name} {k}", flush=True)
    w = world_from_fixture(fx, pgs_mode=1)
    _load(w, fx, k)
    w.step(1)
    rows, offs = w.last_islands()
    if len(fx[f"s{k}_island_rows"]):  # the dump holds islands only for steps with a collision
        assert np.array_equal(rows, fx[f"s{k}_island_rows"]) and np.array_equal(offs, fx[f"s{k}_island_off"])
    js, ref = w.joint_state(), fx[f"s{k}_joint_post"]
    lim = _limits(fx)
    assert np.array_equal(js[lim, 6], ref[lim, 6]), (js[lim, 6], ref[lim, 6])
    err = _rel(w.state(), fx[f"s{k}_post"])
    jerr = _rel(js, ref)
    print(f"  state rel {err:.2e}, joint state rel {jerr:.2e}", flush=True)
    assert err <= 1e-12, err
    assert jerr <= 1e-10, jerr
    w.close()


@pytest.mark.parametrize("name", ["limits", "limits_h0"])
def test_whole_run_from_the_initial_state(name):
    """150 steps from the initial state: the limits start inactive, the world
    finds the holonomic system itself (the two ball-sockets; the limits stay
    out of it), and the runs pass through the same engage / release events."""
    fx = golden(name)
    print(f"limits whole run {name}", flush=True)
    w = world_from_fixture(fx, pgs_mode=1)
    w.set_state(fx["initial_state"])
    lim = _limits(fx)
    assert np.array_equal(w.joint_state()[lim, 6], np.ones(len(lim)))
    rec = set(fx.recorded)
    worst = jworst = 0.0
    for k in range(int(fx.meta["steps"])):
        w.step(1)
        if k == 0 and int(fx.meta["hol_iters"]) > 0:
            og = [s[0] for s in w.holonomic_systems()]
            assert len(og) == 1 and np.array_equal(og[0], fx["s1_hsys_og"])
        if k in rec:
            js = w.joint_state()
            assert np.array_equal(js[lim, 6], fx[f"s{k}_joint_post"][lim, 6]), f"flags at step {k}"
            worst = max(worst, _rel(w.state(), fx[f"s{k}_post"]))
            jworst = max(jworst, _rel(js, fx[f"s{k}_joint_post"]))
    print(f"  worst state rel {worst:.2e}, joint state rel {jworst:.2e}", flush=True)
    assert worst <= 1e-9 and jworst <= 1e-8
    w.close()


def test_limit_rows_match_oracle():
    """The rows of every joint after a step with both a cone and the twist
    limit active (step 30)."""
    fx = golden("limits")
    k = 30
    w = world_from_fixture(fx, pgs_mode=1)
    _load(w, fx, k)
    w.step(1)
    o = Oracle(fx)
    o.load(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"], fx.cache(k, "pre"))
    o.load_joints(fx[f"s{k}_joint_pre"], fx[f"s{k}_hsys_off"], fx[f"s{k}_hsys_og"])
    o.step(_node_order(fx, k))
    for j in range(len(fx["joint_info"])):
        g, r = w.joint_rows(j), o.joint_rows(j)
        worst = max(_rel(a, b) for a, b in zip(g, r))
        assert worst <= 1e-10, (j, worst)
    w.close()


def test_colored_islands_replayed_by_oracle():
    """Coloured islands over 40 steps through limit engage and release
    events: each step equals the oracle's sequential replay of the recorded
    row order."""
    fx = golden("limits")
    k = 10
    w = world_from_fixture(fx, pgs_mode=2)
    _load(w, fx, k)
    o = Oracle(fx)
    o.load(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"], fx.cache(k, "pre"))
    o.load_joints(fx[f"s{k}_joint_pre"], fx[f"s{k}_hsys_off"], fx[f"s{k}_hsys_og"])
    worst = jworst = 0.0
    for s in range(40):
        w.step(1)
        rows, offs = w.last_islands()
        o.step_scheduled(rows, offs)
        worst = max(worst, _rel(w.state(), o.state()))
        jworst = max(jworst, _rel(w.joint_state(), o.joint_state()))
    print(f"  40 steps: worst state rel {worst:.2e}, joints {jworst:.2e}", flush=True)
    assert worst <= 1e-12 and jworst <= 1e-10
    w.close()


def test_limits_stay_out_of_holonomic_systems():
    fx = golden("limits")
    w = world_from_fixture(fx, pgs_mode=1)
    lim = _limits(fx)
    with pytest.raises(cpa.CpError):
        w.set_holonomic_systems(np.array([0, 2]), np.array([0, lim[0]]))
    w.close()
