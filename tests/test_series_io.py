"""The time-series module's file helpers (importParticles.h:8-32 format) on
the CPU: round trip of steps of different sizes (an empty one included) and
the importer's error on a truncated step file."""
import os

import numpy as np
import pytest

import aarbf
from aarbf import series


def test_series_file_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    steps = [("0", rng.normal(size=(64, 4)).astype(np.float32)), ("5", rng.normal(size=(3, 4)).astype(np.float32)),
             ("9", np.zeros((0, 4), np.float32))]
    path = str(tmp_path / "s.txt")
    series.export_particles(path, steps)
    back = series.import_particles(path)
    assert [n for n, _ in back] == ["0", "5", "9"]
    for (_, a), (_, b) in zip(back, steps):
        assert a.dtype == np.float32 and np.array_equal(a, b)


def test_series_truncated_step_raises(tmp_path):
    steps = [("0", np.ones((10, 4), np.float32))]
    path = str(tmp_path / "s.txt")
    series.export_particles(path, steps)
    step_file = path + "_0"
    assert os.path.exists(step_file)
    with open(step_file, "r+b") as f:
        f.truncate(8 + 16 * 5)
    with pytest.raises(aarbf.AarbfError):
        series.import_particles(path)
