"""The fused substep kernels (cp_config.fused: each scene's workgroup builds
its rows, solves, writes its impulses back and integrates its bodies,
substep after substep) against one launch per stage and substep, bit for bit:
same arithmetic, same order, per scene."""
import numpy as np
import pytest

import constraintphysics_amd as cpa
from constraintphysics_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [1, 0])
def test_fused_substeps_equal_per_stage_launches(precision):
    sizes = np.concatenate([scenes.c3_sizes(30, 77, lo=8, hi=900), [3500]])  # the last scene uses the global-memory kernel
    out = {}
    for fused in (1, 0):
        w = cpa.World(dt=1.0 / 120.0, vel_iters=10, pos_iters=1, substeps=8, pgs_mode=0, precision=precision, fused=fused)
        scenes.build_batched(w, sizes, 77, hull_lib_size=32)
        for _ in range(70):
            w.step(1)
        out[fused] = (w.state(), w.contacts())
        w.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert all(np.array_equal(a, b) for a, b in zip(out[0][1], out[1][1]))
    assert len(out[0][1][0]) > 1000
