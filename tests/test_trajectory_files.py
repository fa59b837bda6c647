"""Trajectory files (host-only readers of the C ABI): frames of 7 doubles per
body, the layout of the reference's recorded runs (ImageDemo.h:98-133)."""
import numpy as np
import pytest

import constraintphysics_amd as cpa


def test_reads_frames_written_in_the_reference_layout(tmp_path):
    rng = np.random.default_rng(3)
    frames = rng.normal(size=(5, 11, 7))
    path = tmp_path / "t.traj"
    frames.astype("<f8").tofile(path)
    assert cpa.trajectory_frames(str(path), 11) == 5
    for f in range(5):
        assert np.array_equal(cpa.trajectory_read(str(path), 11, f), frames[f])


def test_rejects_truncated_files_and_missing_frames(tmp_path):
    path = tmp_path / "t.traj"
    np.zeros(7 * 4 * 3 + 2).tofile(path)  # not a whole number of frames
    with pytest.raises(cpa.CpError):
        cpa.trajectory_frames(str(path), 4)
    np.zeros(7 * 4 * 3).tofile(path)
    assert cpa.trajectory_frames(str(path), 4) == 3
    with pytest.raises(cpa.CpError):
        cpa.trajectory_read(str(path), 4, 3)
    with pytest.raises(cpa.CpError):
        cpa.trajectory_frames(str(tmp_path / "missing.traj"), 4)
