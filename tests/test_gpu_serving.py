"""Serving with continuous batching (constraintphysics_amd.serving) on the GPU.

Requests (scenes of the batched-scene distribution with their own step counts)
arrive faster than the serving world can hold them, so they queue, are
admitted into freed body ranges while other scenes are mid-flight, and leave
when done.  Each request's final state must equal the same scene stepped alone
in a plain world with the same configuration, bit for bit: scenes occupy
fixed body ranges and every kernel works per scene.
"""
import numpy as np
import pytest

import constraintphysics_amd as cpa
from constraintphysics_amd import scenes, serving

pytestmark = pytest.mark.gpu

CFG = dict(dt=1.0 / 120.0, gravity=(0.0, -6.0, 0.0), vel_iters=10, pos_iters=1, substeps=8, pgs_mode=0)


@pytest.mark.parametrize("precision", [1, 0])
def test_serving_results_equal_solo_runs(precision):
    seed = 99
    reqs, sizes = serving.poisson_requests(24, rate=4000.0, seed=seed, max_steps=30, lo=8, hi=200)
    w = cpa.World(precision=precision, **CFG)
    w.serving_init(900)
    src = serving.SceneSource(w, sizes, seed, hull_lib_size=32)
    out = serving.serve(w, src, reqs, keep_states=True)
    print(f"precision {precision}: {out['requests']} requests, {out['ticks']} ticks, {out['value']:.0f} body-steps/s, "
          f"request p50 {out['p50_request_ms']:.1f} ms p99 {out['p99_request_ms']:.1f} ms", flush=True)
    assert out["ticks"] < sum(r.steps for r in reqs)  # scenes overlapped on the device
    info = w.serving_info()
    assert info[1] == info[0] and info[3] == 0  # every range freed
    w.close()
    for r in reqs:
        a = cpa.World(precision=precision, **CFG)
        scenes.build_batched(a, sizes, seed, scene_indices=[r.scene], hull_lib_size=32)
        a.step(r.steps)
        solo = a.state()
        a.close()
        assert np.array_equal(out["states"][r.rid], solo), r


def test_admission_waits_for_capacity():
    seed = 5
    w = cpa.World(precision=1, **CFG)
    w.serving_init(300)
    reqs, sizes = serving.poisson_requests(6, rate=1e6, seed=seed, max_steps=5, lo=150, hi=180)
    src = serving.SceneSource(w, sizes, seed, hull_lib_size=32)
    sid, mv, mp, mat, st, n = src.payload(0)
    h = w.admit_scene(sid, mv, mp, mat, st)
    assert h is not None
    sid, mv, mp, mat, st, n = src.payload(1)
    assert w.admit_scene(sid, mv, mp, mat, st) is None  # no free range large enough: the request waits
    w.retire_scene(h)
    assert w.admit_scene(sid, mv, mp, mat, st) is not None
    w.step(1)
    w.close()


def test_serving_full_request_distribution_equals_solo_runs():
    """Requests drawn as the serving configuration draws them — scenes of
    8-2048 bodies (log-uniform), step counts log-uniform in [1, 240] —
    arriving faster than a 40,000-slot serving world holds them.  A sample
    spanning the sizes and step counts (the largest scene, the longest run,
    and quantiles of both) equals its scene stepped alone, bit for bit."""
    seed = 2024
    reqs, sizes = serving.poisson_requests(40, rate=2000.0, seed=seed)
    w = cpa.World(precision=1, **CFG)
    w.serving_init(40000)
    src = serving.SceneSource(w, sizes, seed)
    out = serving.serve(w, src, reqs, keep_states=True)
    print(f"{out['requests']} requests ({sum(r.bodies for r in reqs)} bodies, steps {min(r.steps for r in reqs)}-"
          f"{max(r.steps for r in reqs)}), {out['ticks']} ticks, request p50 {out['p50_request_ms']:.1f} ms p99 "
          f"{out['p99_request_ms']:.1f} ms", flush=True)
    assert max(r.bodies for r in reqs) > 1000 and max(r.steps for r in reqs) > 100
    assert out["ticks"] < sum(r.steps for r in reqs)
    w.close()
    by_size = sorted(reqs, key=lambda r: r.bodies)
    by_steps = sorted(reqs, key=lambda r: r.steps)
    sample = {by_size[-1].rid: by_size[-1], by_steps[-1].rid: by_steps[-1]}
    for q in (0.25, 0.5, 0.75):
        for lst in (by_size, by_steps):
            r = lst[int(q * (len(lst) - 1))]
            sample[r.rid] = r
    for r in sample.values():
        a = cpa.World(precision=1, **CFG)
        scenes.build_batched(a, sizes, seed, scene_indices=[r.scene])
        a.step(r.steps)
        solo = a.state()
        a.close()
        assert np.array_equal(out["states"][r.rid], solo), r
