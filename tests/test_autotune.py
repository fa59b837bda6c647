"""The PGS launch-shape autotuner's search (host logic, a synthetic cost).

The tuner only reads World.cfg, calls set_config / step and reads the PGS
time from last_timing(); a stand-in world with a known cost surface checks
that coordinate descent finds its minimum, keeps every trial valid (class
bounds ascending, widths from each class's candidate list, 64..1024) and
leaves the best setting applied.
"""
import types

import numpy as np

from constraintphysics_amd import autotune


class FakeWorld:
    def __init__(self, best):
        self.cfg = types.SimpleNamespace(pgs_class_threads=[64, 128, 256], pgs_class_bodies=[192, 768, 3072])
        self.best = best
        self.seen = []
        self.steps = 0

    def set_config(self, pgs_class_bodies, pgs_class_threads):
        b, t = list(pgs_class_bodies), list(pgs_class_threads)
        assert b[0] < b[1] < b[2] <= 3072
        assert all(x in autotune.THREADS[k] for x, k in zip(t, ("T0", "T1", "T2")))
        self.cfg.pgs_class_bodies, self.cfg.pgs_class_threads = b, t
        self.seen.append((tuple(b), tuple(t)))

    def step(self, n):
        self.steps += n

    def last_timing(self):
        b, t = self.cfg.pgs_class_bodies, self.cfg.pgs_class_threads
        cur = dict(T0=t[0], T1=t[1], T2=t[2], B0=b[0], B1=b[1])
        cost = 10.0 + sum(abs(np.log2(cur[k] / self.best[k])) for k in self.best)
        out = np.zeros(8)
        out[5] = cost
        return out


def test_coordinate_descent_finds_the_minimum():
    best = dict(T0=128, T1=64, T2=1024, B0=256, B1=1024)
    w = FakeWorld(best)
    r = autotune.tune(w, steps_per_trial=3, rounds=3)
    assert r["pgs_class_threads"] == [128, 64, 1024]
    assert r["pgs_class_bodies"] == [256, 1024, 3072]
    assert r["pgs_ms"] == 10.0 and r["default_pgs_ms"] > 10.0
    assert w.cfg.pgs_class_threads == [128, 64, 1024] and w.cfg.pgs_class_bodies == [256, 1024, 3072]
    assert w.steps == r["steps"]


def test_keeps_the_start_when_nothing_is_better():
    best = dict(T0=64, T1=128, T2=256, B0=192, B1=768)
    w = FakeWorld(best)
    r = autotune.tune(w, steps_per_trial=1, rounds=2)
    assert r["pgs_class_threads"] == [64, 128, 256] and r["pgs_class_bodies"] == [192, 768, 3072]
    assert r["pgs_ms"] == r["default_pgs_ms"] == 10.0
