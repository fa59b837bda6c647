"""Autotuner for the launch shapes of the batched coloured PGS solve.

The batched solve (CP_PGS_COLORED) runs one workgroup per scene in three
LDS size classes: scenes of at most B0 < B1 < B2 bodies keep their velocity
deltas in LDS and run with T0, T1, T2 threads per workgroup (larger scenes
use the global-memory kernel).  The three class kernels run concurrently on
separate streams, so how the scenes are split and how wide each class runs
decides how well they share the chip.

Neither setting changes results: edges of one colour share no dynamic body,
so which thread runs which edge does not matter, and every class runs the same
sweep.  The tuner therefore measures candidates on the live world as it
steps: coordinate descent over the candidate values of each knob, each
candidate timed as the median PGS time (HIP events inside the library,
`World.last_timing()[5]`) over a few steps; the best setting is kept and
applied.
"""
import numpy as np

THREADS = {"T0": (64, 128, 256), "T1": (64, 128, 256, 512), "T2": (128, 256, 512, 1024)}
BOUND0 = (96, 128, 192, 256, 384)
BOUND1 = (512, 640, 768, 1024, 1536)
BOUND2 = 3072
KNOBS = ("T0", "T1", "T2", "B0", "B1")


def _config(p):
    return dict(pgs_class_bodies=(p["B0"], p["B1"], BOUND2), pgs_class_threads=(p["T0"], p["T1"], p["T2"]))


def _candidates(knob, p):
    if knob.startswith("T"):
        return THREADS[knob]
    if knob == "B0":
        return tuple(v for v in BOUND0 if v < p["B1"])
    return tuple(v for v in BOUND1 if v > p["B0"])


def measure(world, p, steps):
    """Median PGS milliseconds per step of `steps` steps under setting p
    (the world advances `steps` steps)."""
    world.set_config(**_config(p))
    ms = []
    for _ in range(steps):
        world.step(1)
        ms.append(float(world.last_timing()[5]))
    return float(np.median(ms))


def tune(world, steps_per_trial=3, rounds=2, log=None):
    """Coordinate descent over KNOBS on `world` (stepping it); applies and
    returns the best setting with the trial record."""
    c = world.cfg
    p = dict(T0=c.pgs_class_threads[0], T1=c.pgs_class_threads[1], T2=c.pgs_class_threads[2], B0=c.pgs_class_bodies[0],
             B1=c.pgs_class_bodies[1])
    start = dict(p)
    trials = []
    n_meas = 1
    best = measure(world, p, steps_per_trial)
    base = best
    trials.append((dict(p), best))
    for _ in range(rounds):
        changed = False
        for knob in KNOBS:
            for v in _candidates(knob, p):
                if v == p[knob]:
                    continue
                q = dict(p)
                q[knob] = v
                t = measure(world, q, steps_per_trial)
                n_meas += 1
                trials.append((dict(q), t))
                if log:
                    log(f"[autotune] {knob}={v}: {t:.3f} ms (best {best:.3f})")
                if t < best:
                    best, p, changed = t, q, True
        if not changed:
            break
    # the incumbent is re-measured once: a setting kept on one noisy trial
    # must still beat the starting one
    if p != start:
        again = measure(world, p, steps_per_trial)
        start_again = measure(world, start, steps_per_trial)
        n_meas += 2
        if start_again <= again:
            p, best = start, start_again
        else:
            best = again
    world.set_config(**_config(p))
    return dict(pgs_class_bodies=[p["B0"], p["B1"], BOUND2], pgs_class_threads=[p["T0"], p["T1"], p["T2"]], pgs_ms=round(best, 4),
                default_pgs_ms=round(base, 4), trials=len(trials), steps=n_meas * steps_per_trial)
