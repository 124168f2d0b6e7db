"""Writes the regression fixtures of small seeded uniform sets: points and
the oracle's cells (volume, face counts, CSR neighbours and areas).

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "the-tessellator_amd")]

import tess  # noqa: E402
from oracle import oracle as orc  # noqa: E402

for n, seed, density in [(100, 100, 1.25), (1000, 1000, 1.25), (1000, 1001, 5.0)]:
    pts = tess.uniform(n, seed)
    r = orc.compute(pts, np.zeros(3), np.ones(3), density=density)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"uniform_{n}_seed{seed}.npz")
    np.savez(path, points=pts, density=np.float64(density), volume=r["volume"], face_count=r["face_count"],
             face_offset=r["face_offset"], neighbor=r["neighbor"], area=r["area"])
    print(path)
