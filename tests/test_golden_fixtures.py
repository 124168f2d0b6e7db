"""Regression fixtures: the oracle still reproduces the committed cells."""
import glob
import os

import numpy as np
import pytest

import tess
from oracle import oracle as orc

FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "uniform_*.npz")))


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p) for p in FIX])
def test_oracle_matches_fixture(path):
    f = np.load(path, allow_pickle=False)
    n = len(f["points"])
    seed = int(os.path.basename(path).split("seed")[1].split(".")[0])
    assert np.array_equal(tess.uniform(n, seed), f["points"])  # generator unchanged
    r = orc.compute(f["points"], np.zeros(3), np.ones(3), density=float(f["density"]))
    for k in ("volume", "face_count", "face_offset", "neighbor", "area"):
        assert np.array_equal(r[k], f[k]), k
