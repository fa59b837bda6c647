"""PGS size classes and workgroup widths on the GPU: they change speed only.

A batch whose scenes span all three LDS classes and the global-memory
fallback is stepped under several class settings (bounds moved so scenes
change class, widths of 64..256, one setting sending the largest scenes to
the global-memory kernel) in both precisions: every run ends in the same
state bit for bit.  The tuner, run on a live batch, returns and applies a
valid setting.
"""
import numpy as np
import pytest

import constraintphysics_amd as cpa
from constraintphysics_amd import autotune, scenes

pytestmark = pytest.mark.gpu

SETTINGS = [
    dict(pgs_class_bodies=(192, 768, 3072), pgs_class_threads=(64, 128, 256)),
    dict(pgs_class_bodies=(40, 90, 3072), pgs_class_threads=(256, 64, 128)),
    dict(pgs_class_bodies=(20, 60, 100), pgs_class_threads=(128, 256, 64)),  # scenes above 100 bodies: global memory
]


def _run(setting, precision, sizes):
    w = cpa.World(dt=1.0 / 120.0, gravity=(0.0, -6.0, 0.0), vel_iters=10, pos_iters=1, substeps=8, pgs_mode=0, precision=precision,
                  **setting)
    scenes.build_batched(w, sizes, 31, hull_lib_size=16)
    for _ in range(50):
        w.step(1)
    return w.state(), w.contacts()[2]


@pytest.mark.parametrize("precision", [0, 1])
def test_class_settings_do_not_change_results(precision):
    sizes = [8, 15, 30, 55, 80, 140, 220]
    print(f"class settings, precision {precision}", flush=True)
    ref_s, ref_c = _run(SETTINGS[0], precision, sizes)
    for st in SETTINGS[1:]:
        s, c = _run(st, precision, sizes)
        assert np.array_equal(s, ref_s), st
        assert np.array_equal(c, ref_c), st


def test_tuner_applies_a_valid_setting():
    sizes = scenes.c3_sizes(48, 5, lo=8, hi=400)
    w = cpa.World(dt=1.0 / 120.0, gravity=(0.0, -6.0, 0.0), vel_iters=10, pos_iters=1, substeps=8, pgs_mode=0, precision=1)
    scenes.build_batched(w, sizes, 5, hull_lib_size=16)
    for _ in range(30):
        w.step(1)
    r = autotune.tune(w, steps_per_trial=1, rounds=1)
    print(f"tuned {r}", flush=True)
    b, t = r["pgs_class_bodies"], r["pgs_class_threads"]
    assert b[0] < b[1] < b[2] <= 3072 and all(x in (64, 128, 256) for x in t)
    assert list(w.cfg.pgs_class_bodies) == b and list(w.cfg.pgs_class_threads) == t
    w.step(1)
    assert np.all(np.isfinite(w.state()))


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replay_matches_eager(precision):
    """The substep loop captured in a HIP graph and replayed step after step
    (contact counts change every step, buffers grow and force recaptures)
    ends every step in the same state as the eager launches, bit for bit;
    switching the graph off and on mid-run changes nothing either."""
    sizes = [8, 15, 30, 55, 80, 140, 220]
    ws = []
    for g in (0, 1):
        w = cpa.World(dt=1.0 / 120.0, gravity=(0.0, -6.0, 0.0), vel_iters=10, pos_iters=1, substeps=8, pgs_mode=0, precision=precision,
                      graph=g)
        scenes.build_batched(w, sizes, 31, hull_lib_size=16)
        ws.append(w)
    for k in range(60):
        if k == 30:
            ws[1].set_config(graph=0)
        if k == 35:
            ws[1].set_config(graph=1)
        for w in ws:
            w.step(1)
        a, b = ws[0].state(), ws[1].state()
        assert np.array_equal(a, b), k
    assert np.array_equal(ws[0].contacts()[2], ws[1].contacts()[2])
    assert len(ws[1].contacts()[0]) > 0
