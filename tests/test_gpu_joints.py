"""Joints and the holonomic block solver on the GPU, through the C ABI.

Scenes: tests/golden/chain3 (three free chains of six cylinder links joined by
ball-sockets with every third joint a hinge, holonomic block solve on) and
chain3h0 (the same scene with the block solve off), dumped from the compiled
reference engine.  Bars, as for the contact step in test_gpu_parity.py: island
rows bit-exact, body state within 1e-12 relative, joint impulses and hinge
bases within 1e-10 relative (scale max(1, |x|)); the holonomic systems'
constraint order, elimination order and block table bit-exact against the
reference's dump, their factor within 1e-10 of the oracle's.
"""
import numpy as np
import pytest

import constraintphysics_amd as cpa
from fixtures import Oracle, golden, world_from_fixture

pytestmark = pytest.mark.gpu

CHAINS = [("chain3", k) for k in (1, 2, 45, 70, 99)] + [("chain3h0", k) for k in (1, 45, 70, 99)]


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def _node_order(fx, k):
    no = fx[f"s{k}_node_order"]
    return no if len(no) else cpa.reference_node_order(int(fx.meta["n_bodies"]))


def _systems(fx, k):
    off, og = fx[f"s{k}_hsys_off"], fx[f"s{k}_hsys_og"]
    return [tuple(og[off[q]:off[q + 1]]) for q in range(len(off) - 1)]


def _load(w, fx, k, explicit=True):
    w.set_state(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"])
    w.set_contacts(*fx.cache(k, "pre"))
    if explicit:
        w.set_holonomic_systems(fx[f"s{k}_hsys_off"], fx[f"s{k}_hsys_og"])
    w.set_joint_state(fx[f"s{k}_joint_pre"])
    w.set_node_order(_node_order(fx, k))


@pytest.mark.parametrize("name,k", CHAINS)
def test_chain_step_matches_reference(name, k):
    fx = golden(name)
    print(f"chain step {name} {k}", flush=True)
    w = world_from_fixture(fx, pgs_mode=1)
    _load(w, fx, k)
    if int(fx.meta["hol_iters"]) > 0:
        off, toff = fx[f"s{k}_hsys_off"], fx[f"s{k}_hsys_table_off"]
        for q, (og, order, table, vloc, _) in enumerate(w.holonomic_systems()):
            assert np.array_equal(og, fx[f"s{k}_hsys_og"][off[q]:off[q + 1]])
            assert np.array_equal(order, fx[f"s{k}_hsys_order"][off[q]:off[q + 1]])
            assert np.array_equal(table, fx[f"s{k}_hsys_table"][toff[q]:toff[q + 1]])
    w.step(1)
    rows, offs = w.last_islands()
    if len(fx[f"s{k}_island_rows"]):
        assert np.array_equal(rows, fx[f"s{k}_island_rows"])
    err = _rel(w.state(), fx[f"s{k}_post"])
    jerr = _rel(w.joint_state(), fx[f"s{k}_joint_post"])
    print(f"  state rel {err:.2e}, joint state rel {jerr:.2e}", flush=True)
    assert err <= 1e-12, err
    assert jerr <= 1e-10, jerr


@pytest.mark.parametrize("k", [1, 45, 99])
def test_holonomic_factor_matches_oracle(k):
    """After a step both hold the rows and block LDL factor the next step's
    first substep uses (rebuilt from the post-step state)."""
    fx = golden("chain3")
    w = world_from_fixture(fx, pgs_mode=1)
    _load(w, fx, k)
    w.step(1)
    o = Oracle(fx)
    o.load(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"], fx.cache(k, "pre"))
    o.load_joints(fx[f"s{k}_joint_pre"], fx[f"s{k}_hsys_off"], fx[f"s{k}_hsys_og"])
    o.step(_node_order(fx, k))
    gs, os_ = w.holonomic_systems(), o.hsys()
    assert len(gs) == len(os_) == 3
    for (og, order, table, vloc, buf), (o_order, o_table, o_vloc, o_buf) in zip(gs, os_):
        assert np.array_equal(order, o_order) and np.array_equal(table, o_table) and np.array_equal(vloc, o_vloc)
        assert _rel(buf, o_buf) <= 1e-10, _rel(buf, o_buf)
    for j in range(w.joint_count()):
        for a, b in zip(w.joint_rows(j), o.joint_rows(j)):
            assert _rel(a, b) <= 1e-10


def test_systems_found_like_the_reference():
    """Without an explicit set, the first step finds the holonomic systems
    (maintainHolonomicSystems): the same members in the same constraint
    order as the reference's dump, and the same step result."""
    fx = golden("chain3")
    k = 1
    w = world_from_fixture(fx, pgs_mode=1)
    _load(w, fx, k, explicit=False)
    w.step(1)
    found = sorted(tuple(s[0]) for s in w.holonomic_systems())
    assert found == sorted(_systems(fx, k))
    assert _rel(w.state(), fx[f"s{k}_post"]) <= 1e-12
    assert _rel(w.joint_state(), fx[f"s{k}_joint_post"]) <= 1e-10


def test_chain_run_tracks_oracle():
    """40 steps from the first recorded state, GPU against the oracle."""
    fx = golden("chain3")
    k = 1
    w = world_from_fixture(fx, pgs_mode=1)
    _load(w, fx, k, explicit=False)
    o = Oracle(fx)
    o.load(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"], fx.cache(k, "pre"))
    o.load_joints(fx[f"s{k}_joint_pre"], fx[f"s{k}_hsys_off"], fx[f"s{k}_hsys_og"])
    no = _node_order(fx, k)
    worst = 0.0
    for s in range(40):
        w.step(1)
        o.step(no)
        worst = max(worst, _rel(w.state(), o.state()))
    print(f"40 steps: worst state rel {worst:.2e}", flush=True)
    assert worst <= 1e-9, worst
    assert _rel(w.joint_state(), o.joint_state()) <= 1e-8


def _island_sets(rows, offs):
    return sorted(tuple(sorted(map(tuple, rows[offs[i]:offs[i + 1]]))) for i in range(len(offs) - 1))


@pytest.mark.parametrize("name,k", [("chain3", 1), ("chain3", 45), ("chain3", 99), ("chain3h0", 45)])
def test_island_colored_replayed_by_oracle(name, k):
    """Coloured islands (one wave per island, edges coloured so that no two
    of a colour share a dynamic body) over 10 steps: each step equals the
    oracle's sequential replay of the recorded row order, and the first
    step's islands hold the same rows as the reference-order islands."""
    fx = golden(name)
    print(f"island-coloured {name} {k}", flush=True)
    ref = world_from_fixture(fx, pgs_mode=1)
    _load(ref, fx, k)
    ref.step(1)
    w = world_from_fixture(fx, pgs_mode=2)
    _load(w, fx, k)
    o = Oracle(fx)
    o.load(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"], fx.cache(k, "pre"))
    o.load_joints(fx[f"s{k}_joint_pre"], fx[f"s{k}_hsys_off"], fx[f"s{k}_hsys_og"])
    worst = jworst = 0.0
    for s in range(10):
        w.step(1)
        rows, offs = w.last_islands()
        if s == 0:
            assert _island_sets(rows, offs) == _island_sets(*ref.last_islands())
        o.step_scheduled(rows, offs)
        worst = max(worst, _rel(w.state(), o.state()))
        jworst = max(jworst, _rel(w.joint_state(), o.joint_state()))
    print(f"  10 steps: worst state rel {worst:.2e}, joints {jworst:.2e}", flush=True)
    assert worst <= 1e-12 and jworst <= 1e-10


def _two_boxes(joint):
    from constraintphysics_amd import scenes
    w = cpa.World(dt=1.0 / 120.0, gravity=(0.0, 0.0, 0.0), vel_iters=10, pos_iters=1, substeps=4, pgs_mode=1)
    b = scenes.SceneBuilder(w)
    sid, mp = b.box_shape(0.5, 0.5, 0.5, 1.0)
    b.add(sid, mp, 0, (1.0, 0.2, 0.5, 0.5), 0, np.zeros(3))
    b.add(sid, mp, 0, (1.0, 0.2, 0.5, 0.5), 0, np.array([0.4, 0.0, 0.0]))  # overlapping by 0.1
    b.finish()
    if joint:
        geo = np.zeros(14)
        geo[0], geo[3] = 0.2, -0.2
        geo[12] = 15000.0
        w.add_joint(3, 0, 1, geo)
    return w


def test_jointed_pair_does_not_collide():
    free = _two_boxes(False)
    free.step(1)
    assert len(free.contacts()[0]) > 0
    tied = _two_boxes(True)
    tied.step(1)
    assert len(tied.contacts()[0]) == 0
    assert np.all(np.isfinite(tied.state()))


def test_joints_need_an_island_mode():
    w = _two_boxes(True)
    w.set_config(pgs_mode=0)
    with pytest.raises(cpa.CpError):
        w.step(1)
    w.set_config(pgs_mode=2)
    w.step(1)
    assert np.all(np.isfinite(w.state()))
