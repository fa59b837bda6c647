"""GPU side of tests/test_defects.py: the narrowphase is bug-compatible with
the reference when defect fixes are off (same NaN point, bit for bit) and
drops exactly that point when they are on."""
import os

import numpy as np
import pytest

from fixtures import GOLDEN, Oracle, clip_nan_description

pytestmark = pytest.mark.gpu
HERE = os.path.join(GOLDEN, "clip_nan")


@pytest.mark.parametrize("fix", [0, 1])
def test_clip_nan_case(fix):
    from fixtures import world_from_description
    desc = clip_nan_description()
    ld = lambda k: np.load(os.path.join(HERE, k + ".npy"))
    desc["state"] = ld("st")
    w = world_from_description(desc, dt=1 / 120, vel_iters=10, pos_iters=1, substeps=8, defect_fixes=fix)
    counts, heads, pts, mg = w.narrowphase(np.array([[1, 495]]))
    o = Oracle(desc=desc, dt=1 / 120, gravity_y=-6.0, vel_iters=10, pos_iters=1, substeps=8, drop_nonfinite=fix)
    o.load(ld("st"), None, (ld("ii"), ld("mg"), ld("ff")))
    h, p, m = o.narrowphase(1, 495)
    n = counts[0]
    assert n == len(p)
    np.testing.assert_array_equal(heads[0], h)
    np.testing.assert_array_equal(pts[0, :n], p)
    np.testing.assert_array_equal(mg[0, :n], m)
    assert (np.isnan(pts[0, :n]).any(axis=1).sum() == 1) if fix == 0 else np.all(np.isfinite(pts[0, :n]))
