"""Reference defects the product handles, pinned on the CPU oracle.

Clip NaN: getEdgeIntersectionWithClippingEdge (CollisionDetect.cpp:444-467)
divides by clip_edge_norm.dot(d); when the clipped edge is parallel to the
clipping edge within rounding that is 0 and the contact point is NaN, after
which the body's state becomes NaN (the reference propagates it silently).
The oracle (the reference restated bit for bit) reproduces the NaN; with the
product's defect fix (drop non-finite points) the manifold keeps its finite
points.  The GPU side is checked in test_gpu_defects.py.
"""
import os

import numpy as np
import pytest

from fixtures import GOLDEN, Oracle, clip_nan_description

HERE = os.path.join(GOLDEN, "clip_nan")


def _oracle(drop):
    desc = clip_nan_description()
    o = Oracle(desc=desc, dt=1 / 120, gravity_y=-6.0, vel_iters=10, pos_iters=1, substeps=8, drop_nonfinite=drop)
    ld = lambda k: np.load(os.path.join(HERE, k + ".npy"))
    o.load(ld("st"), None, (ld("ii"), ld("mg"), ld("ff")))
    return o


def test_reference_clip_emits_nan_point():
    head, pts, mg = _oracle(0).narrowphase(1, 495)
    assert len(pts) == 4
    assert np.isnan(pts).any(axis=1).sum() == 1
    assert np.all(np.isfinite(head))


def test_defect_fix_drops_only_the_nan_point():
    h0, p0, m0 = _oracle(0).narrowphase(1, 495)
    h1, p1, m1 = _oracle(1).narrowphase(1, 495)
    fin = np.isfinite(p0).all(axis=1)
    assert np.array_equal(h1, h0)
    assert np.array_equal(p1, p0[fin]) and np.array_equal(m1, m0[fin])


def test_defect_fix_keeps_scene_finite():
    o = _oracle(1)
    o.step()
    assert np.all(np.isfinite(o.state()))
