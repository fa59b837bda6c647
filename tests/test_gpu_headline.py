"""The headline workload — 4096 batched scenes, bodies log-uniform in
[8, 2048] (seed 20240), the mixed-primitive distribution in box containers,
10 PGS velocity sweeps, graph-coloured PGS — checked against the oracle at
its full size.

The batch is settled in the split precision (the headline's), then one step
is taken from the settled state in FP64 and one in the split precision, each
recording its colour schedule.  Scenes never interact, so each sampled scene
is replayed on its own by the oracle: it loads the scene's state, world
inertia and contact cache, runs its own narrowphase and persistence, and
sweeps the rows in the GPU's recorded colour order (edges of one colour share
no dynamic body, so the coloured sweep is that sequential Gauss-Seidel order;
ConstraintSolver.cpp:80-213).  Bars:

  * FP64 step vs the oracle replay: body state within 1e-12 relative
    (max(1, |x|)); contact cache ids, ages, live flags and MagicIDs exact,
    values within 1e-10;
  * split step vs the same oracle replay: velocities within 1e-4 relative
    (2-norm over the scene), poses within 1e-6 absolute, contact impulses
    within 1e-2 relative (redundant resting contacts trade impulse with no
    effect on velocities);
  * both steps see the same colour schedule (narrowphase and persistence
    are FP64 in both precisions).

The sample holds two scenes above 1,500 bodies and scenes with hull-hull
contacts between irregular hulls.  Size-independent properties of the whole
batch: every state finite, every dynamic body inside its container, the
split solve's FP64 natural-map residual no worse than the FP64 solve's, and a
sampled scene stepped alone from the same snapshot equal to its batched step
bit for bit.
"""
import numpy as np
import pytest

import constraintphysics_amd as cpa
from constraintphysics_amd import scenes
from fixtures import Oracle

pytestmark = pytest.mark.gpu

SEED = 20240
N_SCENES = 4096
SETTLE = 150
CFG = dict(dt=1.0 / 120.0, gravity=(0.0, -6.0, 0.0), vel_iters=10, pos_iters=1, substeps=8, pgs_mode=0)


def _recorded_step(w, snap, precision):
    st, inertia, cache = snap
    w.set_config(precision=precision, diagnostics=3)
    w.set_state(st, inertia)
    w.set_contacts(*cache)
    w.step(1)
    rows, offs = w.last_islands()
    return dict(state=w.state(), contacts=w.contacts(), rows=rows, offs=offs, coff=w.last_scene_contacts(),
                resid=w.last_residual())


@pytest.fixture(scope="module")
def batch():
    sizes = scenes.c3_sizes(N_SCENES, SEED)
    w = cpa.World(precision=1, **CFG)
    info = scenes.build_batched(w, sizes, SEED)
    for i in range(SETTLE):
        w.step(1)
    snap = (w.state(), w.inertia(), w.contacts())
    print(f"settled {N_SCENES} scenes ({info['dynamic_bodies']} dynamic bodies) for {SETTLE} steps: {len(snap[2][0])} contacts",
          flush=True)
    out = dict(sizes=sizes, off=scenes.scene_offsets(sizes), snap=snap)
    out["fp64"] = _recorded_step(w, snap, 0)
    out["split"] = _recorded_step(w, snap, 1)
    w.close()
    return out


def _kinds(sizes, s):
    """Kind per local body index of scene s: -1 wall, 0 hull, 1 sphere, 2 cylinder."""
    lib = scenes.hull_library(SEED)
    sph_r, _ = scenes.sphere_library(SEED)
    _, _, cyl_r = scenes.cylinder_library(SEED)
    kind = scenes.scene_bodies(sizes, SEED, s, lib, sph_r, cyl_r)[0]
    return np.concatenate([np.full(scenes.CONTAINER_WALLS, -1), kind])


def _hull_hull_contacts(sizes, off, contacts, s):
    ii = scenes.scene_contacts(contacts, off, s)[0]
    k = _kinds(sizes, s)
    return int(np.sum((k[ii[:, 0] - 1] == 0) & (k[ii[:, 1] - 1] == 0)))


def _sample(b):
    sizes, off = b["sizes"], b["off"]
    big = [int(s) for s in np.nonzero(sizes > 1500)[0][:2]]
    order = np.argsort(sizes, kind="stable")
    mid = [int(order[int(q * (len(order) - 1))]) for q in (0.1, 0.35, 0.6, 0.85)]
    pick = list(dict.fromkeys(big + mid))
    hh = {s: _hull_hull_contacts(sizes, off, b["fp64"]["contacts"], s) for s in pick}
    for s in order[::-1]:  # add scenes with hull-hull contacts until two sampled scenes have them
        if sum(1 for v in hh.values() if v > 0) >= 2:
            break
        s = int(s)
        if s not in hh:
            c = _hull_hull_contacts(sizes, off, b["fp64"]["contacts"], s)
            if c > 0:
                hh[s] = c
                pick.append(s)
    return pick, hh


def _scene_rows(rec, s):
    rows = rec["rows"][rec["offs"][s]:rec["offs"][s + 1]].copy()
    rows[:, 0] -= rec["coff"][s]
    return rows


def _replay(b, s, rec):
    sizes, off = b["sizes"], b["off"]
    st, cache, inertia = scenes.scene_snapshot(b["snap"][0], b["snap"][2], off, s, b["snap"][1])
    desc = scenes.scene_description(sizes, SEED, s)
    o = Oracle(desc=desc, dt=CFG["dt"], gravity_y=-6.0, vel_iters=10, pos_iters=1, substeps=8, drop_nonfinite=1)
    o.load(st, inertia, cache)
    rows = _scene_rows(rec, s)
    o.step_scheduled(rows, np.array([0, len(rows)], np.int64))
    return o


def _rel_max(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)))) if len(b) else 0.0


def _rel2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_sampled_scenes_match_oracle_replay(batch):
    b = batch
    off = b["off"]
    pick, hh = _sample(b)
    assert sum(1 for s in pick if b["sizes"][s] > 1500) >= 2
    assert sum(1 for v in hh.values() if v > 0) >= 2, hh
    for s in pick:
        o = _replay(b, s, b["fp64"])
        a, z = int(off[s]), int(off[s + 1])
        ref = o.state()
        err64 = _rel_max(b["fp64"]["state"][a:z], ref)
        oi, om, of = o.cache()
        gi, gm, gf = scenes.scene_contacts(b["fp64"]["contacts"], off, s)
        assert np.array_equal(gi[:, [0, 1, 2, 3, 5]], oi[:, [0, 1, 2, 3, 5]]) and np.array_equal(gm, om), s
        cerr = _rel_max(gf, of)
        sv = _rel2(b["split"]["state"][a:z, 7:13], ref[:, 7:13])
        sp = float(np.max(np.abs(b["split"]["state"][a:z, :7] - ref[:, :7])))
        si = _rel2(scenes.scene_contacts(b["split"]["contacts"], off, s)[2][:, 12:15], of[:, 12:15])
        print(f"scene {s}: {b['sizes'][s]} bodies, {len(oi)} contacts ({hh.get(s, 0)} hull-hull); fp64 state {err64:.2e}, "
              f"cache {cerr:.2e}; split velocity {sv:.2e}, pose {sp:.2e}, impulse {si:.2e}", flush=True)
        assert err64 <= 1e-12 and cerr <= 1e-10, s
        assert sv <= 1e-4 and sp <= 1e-6 and si <= 1e-2, s


def test_split_and_fp64_steps_share_the_schedule(batch):
    f, p = batch["fp64"], batch["split"]
    assert np.array_equal(f["offs"], p["offs"]) and np.array_equal(f["coff"], p["coff"])
    assert np.array_equal(f["rows"], p["rows"])


def test_full_batch_properties(batch):
    b = batch
    off, sizes = b["off"], b["sizes"]
    dyn = np.ones(off[-1], bool)
    for a in off[:-1]:
        dyn[a:a + scenes.CONTAINER_WALLS] = False
    scene_of = np.repeat(np.arange(len(sizes)), np.diff(off))
    L = 2.0 * b["snap"][0][off[:-1], 0]  # the floor wall sits at (L/2, -1/2, L/2)
    for name in ("fp64", "split"):
        st = b[name]["state"]
        assert np.all(np.isfinite(st)), name
        p = st[dyn, :3]
        lim = L[scene_of[dyn]][:, None]
        assert np.all(p > -0.25) and np.all(p < lim + 0.25), name
    r64, r32 = b["fp64"]["resid"], b["split"]["resid"]
    print(f"natural-map residual (relative): fp64 {r64[2]:.3e}, split {r32[2]:.3e} over {int(r64[3])} rows", flush=True)
    assert np.isfinite(r64[2]) and r64[2] < 0.1
    assert r32[2] <= 1.5 * r64[2] + 1e-6


@pytest.mark.parametrize("which", ["big", "small"])
def test_sampled_scene_alone_equals_batched(batch, which):
    b = batch
    sizes, off = b["sizes"], b["off"]
    s = int(np.nonzero(sizes > 1500)[0][0]) if which == "big" else int(np.argsort(sizes, kind="stable")[len(sizes) // 4])
    st, cache, inertia = scenes.scene_snapshot(b["snap"][0], b["snap"][2], off, s, b["snap"][1])
    w = cpa.World(precision=1, **CFG)
    scenes.build_batched(w, sizes, SEED, scene_indices=[s])
    w.set_state(st, inertia)
    w.set_contacts(*cache)
    w.step(1)
    alone = w.state()
    w.close()
    assert np.array_equal(alone, b["split"]["state"][off[s]:off[s + 1]]), s


def test_single_large_scene_matches_oracle_replay():
    """The mixed-primitive single scene: 20,000 bodies in a 60 m box, 8 PGS
    velocity sweeps.  After 200 steps (the lowest bodies resting on the floor
    and on each other), one FP64 step and one split step from the same state
    against the oracle's replay of the recorded colour schedule; same bars as
    the batched scenes."""
    cfg = dict(CFG, vel_iters=8)
    sizes = np.array([20000])
    w = cpa.World(precision=1, **cfg)
    scenes.build_batched(w, sizes, 20241)
    for _ in range(200):
        assert w.step(1)
    snap = (w.state(), w.inertia(), w.contacts())
    rec64 = _recorded_step(w, snap, 0)
    rec32 = _recorded_step(w, snap, 1)
    w.close()
    assert np.array_equal(rec64["rows"], rec32["rows"])
    desc = scenes.scene_description(sizes, 20241, 0)
    o = Oracle(desc=desc, dt=cfg["dt"], gravity_y=-6.0, vel_iters=8, pos_iters=1, substeps=8, drop_nonfinite=1)
    o.load(snap[0], snap[1], snap[2])
    o.step_scheduled(rec64["rows"], np.array([0, len(rec64["rows"])], np.int64))
    ref = o.state()
    err64 = _rel_max(rec64["state"], ref)
    oi, om, of = o.cache()
    gi, gm, gf = rec64["contacts"]
    assert np.array_equal(gi[:, [0, 1, 2, 3, 5]], oi[:, [0, 1, 2, 3, 5]]) and np.array_equal(gm, om)
    cerr = _rel_max(gf, of)
    sv = _rel2(rec32["state"][:, 7:13], ref[:, 7:13])
    sp = float(np.max(np.abs(rec32["state"][:, :7] - ref[:, :7])))
    si = _rel2(rec32["contacts"][2][:, 12:15], of[:, 12:15])
    print(f"20000-body scene: {len(oi)} contacts, {len(rec64['rows']) // 2} contact rows; fp64 state {err64:.2e}, cache {cerr:.2e}; "
          f"split velocity {sv:.2e}, pose {sp:.2e}, impulse {si:.2e}", flush=True)
    assert len(oi) > 1000
    assert err64 <= 1e-12 and cerr <= 1e-10
    assert sv <= 1e-4 and sp <= 1e-6 and si <= 1e-2
