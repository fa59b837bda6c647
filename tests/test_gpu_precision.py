"""The FP32/FP64 split solve against the FP64 solve (itself the oracle's
sequential order replayed to 1e-12).

Bar: from the same per-step state, one step of the split path agrees with the
FP64 path within 1e-4 relative on body velocities (2-norm over all bodies),
1e-6 absolute on poses, and 1e-2 relative on the contact impulses.  The
impulse bar is looser on purpose: after 10 sweeps the PGS is not converged
(natural-map residuals of 1e-3 to 1, in FP64 too), and a body resting on
several redundant contact points can shift impulse between them with no
effect on its velocity, so individual impulses amplify FP32 rounding while
the velocities, which the step hands on, do not.  Narrowphase and persistence
stay FP64, so the contact set and MagicIDs are identical (static_ready, the
"friction cone not saturated" flag, may flip on contacts at the cone
boundary).
"""
import numpy as np
import pytest

import constraintphysics_amd as cpa
from constraintphysics_amd import scenes
from fixtures import golden, world_from_fixture

pytestmark = pytest.mark.gpu

TOL = 1e-4
IMP_TOL = 1e-2


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _one_step(make, load, precision, diagnostics=0):
    w = make(precision, diagnostics)
    load(w)
    w.step(1)
    out = (w.state(), w.contacts(), w.last_residual() if diagnostics else None)
    w.close()
    return out


def _compare(make, load):
    s64, (i64, m64, f64), r64 = _one_step(make, load, 0, diagnostics=1)
    s32, (i32, m32, f32), r32 = _one_step(make, load, 1, diagnostics=1)
    # same contacts and MagicIDs; static_ready (column 4, friction cone not
    # saturated) may flip on contacts sitting on the cone boundary
    cols = [0, 1, 2, 3, 5]
    assert np.array_equal(i64[:, cols], i32[:, cols]) and np.array_equal(m64, m32)
    assert np.mean(i64[:, 4] != i32[:, 4]) <= 0.02
    imp = _rel(f32[:, 12:15], f64[:, 12:15])
    vel = _rel(s32[:, 7:13], s64[:, 7:13])
    pos = float(np.max(np.abs(s32[:, :7] - s64[:, :7])))
    return imp, vel, pos, r64, r32


@pytest.mark.parametrize("name,k", [("pile3", 10), ("mix3", 30), ("mix3", 59), ("cyl4", 149)])
def test_split_matches_fp64_on_reference_states(name, k):
    fx = golden(name)

    def make(p, d):
        return world_from_fixture(fx, pgs_mode=0, precision=p, diagnostics=d)

    def load(w):
        w.set_state(fx[f"s{k}_pre"], fx[f"s{k}_pre_inertia"])
        w.set_contacts(*fx.cache(k, "pre"))

    imp, vel, pos, r64, r32 = _compare(make, load)
    print(f"{name} step {k}: impulse rel {imp:.3e}, velocity rel {vel:.3e}, pose abs {pos:.3e}, "
          f"residual fp64 {r64[2]:.3e} split {r32[2]:.3e}", flush=True)
    assert vel <= TOL and pos <= 1e-6 and imp <= IMP_TOL


def test_split_matches_fp64_on_settled_batch():
    sizes = scenes.c3_sizes(48, 5, lo=8, hi=400)
    wb = cpa.World(dt=1.0 / 120.0, vel_iters=10, pos_iters=1, substeps=8, pgs_mode=0)
    scenes.build_batched(wb, sizes, 5, hull_lib_size=32)
    for _ in range(90):
        wb.step(1)
    st, cache = wb.state(), wb.contacts()
    wb.close()

    def make(p, d):
        w = cpa.World(dt=1.0 / 120.0, vel_iters=10, pos_iters=1, substeps=8, pgs_mode=0, precision=p, diagnostics=d)
        scenes.build_batched(w, sizes, 5, hull_lib_size=32)
        return w

    def load(w):
        w.set_state(st)
        w.set_contacts(*cache)

    imp, vel, pos, r64, r32 = _compare(make, load)
    print(f"settled batch: {len(cache[0])} contacts, impulse rel {imp:.3e}, velocity rel {vel:.3e}, pose abs {pos:.3e}, "
          f"residual fp64 {r64[2]:.3e} split {r32[2]:.3e}", flush=True)
    assert vel <= TOL and pos <= 1e-6 and imp <= IMP_TOL
    # the FP32 solve converges as far as the FP64 one (natural-map residual in FP64)
    assert r32[2] <= 1.5 * r64[2] + 1e-6


def test_split_runs_long_and_stays_finite():
    """400 steps of 16 scenes in each precision: the split run stays finite,
    keeps every body inside its container, and settles like the FP64 run
    (trajectories diverge chaotically, so only aggregates are compared)."""
    sizes = scenes.c3_sizes(16, 9, lo=8, hi=300)
    out = {}
    for p in (0, 1):
        w = cpa.World(dt=1.0 / 120.0, vel_iters=10, pos_iters=1, substeps=8, pgs_mode=0, precision=p)
        scenes.build_batched(w, sizes, 9, hull_lib_size=32)
        for _ in range(400):
            w.step(1)
        out[p] = w.state()
        w.close()
    off = np.concatenate([[0], np.cumsum(np.asarray(sizes) + scenes.CONTAINER_WALLS)])
    dyn = np.ones(off[-1], bool)
    for a in off[:-1]:
        dyn[a:a + scenes.CONTAINER_WALLS] = False
    s64, s32 = out[0][dyn], out[1][dyn]
    assert np.all(np.isfinite(s32))
    assert np.min(s32[:, 1]) > 0.0
    v64 = np.linalg.norm(s64[:, 7:10], axis=1)
    v32 = np.linalg.norm(s32[:, 7:10], axis=1)
    print(f"speed median/p90 fp64 {np.median(v64):.3e}/{np.percentile(v64, 90):.3e} split {np.median(v32):.3e}/"
          f"{np.percentile(v32, 90):.3e}; mean height fp64 {s64[:, 1].mean():.4f} split {s32[:, 1].mean():.4f}", flush=True)
    assert np.median(v32) <= 2.0 * np.median(v64) + 0.05
    assert abs(s32[:, 1].mean() - s64[:, 1].mean()) <= 0.05 * s64[:, 1].mean()
