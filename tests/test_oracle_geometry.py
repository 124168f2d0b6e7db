"""Face rings of the oracle (orc_compute_geometry: compute_face_vertices,
polyhedron.rs:895-919), pinned by what the geometry fixes: each ring's fan
area is the face area, its normal points away from the site, it starts at
its lexicographically smallest vertex, every cell is a closed polyhedron
(V - E + F = 2 with shared edges) and, for a simple cubic lattice, every
interior face is a unit square of the known corners."""
import numpy as np

import tess
from oracle import oracle as orc

LO, HI = np.zeros(3), np.ones(3)


def _faces(g):
    ro = g["ring_offset"].astype(np.int64)
    return [g["ring_xyz"][ro[f]:ro[f + 1]] for f in range(len(ro) - 1)]


def test_rings_area_orientation_start():
    pts = tess.uniform(800, 21)
    g = orc.compute_geometry(pts, LO, HI)
    for f, ring in enumerate(_faces(g)):
        w = np.zeros(3)
        for k in range(1, len(ring) - 1):
            w += np.cross(ring[k] - ring[0], ring[k + 1] - ring[0])
        assert abs(0.5 * np.linalg.norm(w) - g["area"][f]) <= 1e-12 * max(g["area"][f], 1e-12)
        assert np.dot(w, ring[0]) > 0
        assert tuple(ring[0]) == min(map(tuple, ring))


def test_closed_polyhedra():
    pts = tess.clustered(1500, 22)
    g = orc.compute_geometry(pts, LO, HI)
    faces = _faces(g)
    fo = g["face_offset"].astype(np.int64)
    for c in range(len(pts)):
        edges = {}
        verts = set()
        for ring in faces[fo[c]:fo[c + 1]]:
            t = [tuple(v) for v in ring]
            verts.update(t)
            for a, b in zip(t, t[1:] + t[:1]):
                edges[(a, b)] = edges.get((a, b), 0) + 1
        # every directed edge once, and its reverse once: a closed, consistently oriented surface
        assert all(v == 1 for v in edges.values())
        assert all((b, a) in edges for a, b in edges)
        assert len(verts) - len(edges) // 2 + (fo[c + 1] - fo[c]) == 2


def test_simple_cubic_faces_are_unit_squares():
    m = 4
    ax = (np.arange(m) + 0.5) / m
    pts = np.array(np.meshgrid(ax, ax, ax, indexing="ij")).reshape(3, -1).T.copy()
    g = orc.compute_geometry(pts, LO, HI)
    assert np.allclose(g["volume"], 1.0 / m ** 3, rtol=1e-12)
    h = 0.5 / m
    for ring in _faces(g):
        assert len(ring) == 4
        span = ring.max(0) - ring.min(0)
        assert np.sum(span < 1e-6) == 1 and np.allclose(np.sort(span)[1:], 2 * h, atol=1e-6)
