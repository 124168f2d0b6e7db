"""Whole cells from the oracle: closed forms, invariants, the identity of the
security-radius stop with streaming every point, the identity of the two
clip restatements, and an independent check against Qhull."""
import numpy as np
import pytest
from scipy.spatial import ConvexHull, Voronoi

import tess
from oracle import oracle as orc

LO, HI = np.zeros(3), np.ones(3)


def csr(res, i):
    a, b = int(res["face_offset"][i]), int(res["face_offset"][i + 1])
    return res["neighbor"][a:b], res["area"][a:b]


def test_one_site():
    r = orc.compute(np.array([[0.3, 0.6, 0.2]]), LO, HI, clip=orc.CLIP_FACES | orc.CHECK)
    assert r["volume"][0] == 1.0
    assert list(r["neighbor"]) == [-6, -5, -4, -3, -2, -1] and np.all(r["area"] == 1.0)


def test_two_sites():
    r = orc.compute(np.array([[0.25, 0.5, 0.5], [0.75, 0.5, 0.5]]), LO, HI, clip=orc.CLIP_FACES | orc.CHECK)
    assert np.array_equal(r["volume"], [0.5, 0.5])
    nb0, ar0 = csr(r, 0)
    assert list(nb0) == [-6, -5, -4, -3, -1, 1]
    assert ar0[list(nb0).index(1)] == 1.0


@pytest.mark.parametrize("m", [3, 4, 5])
def test_simple_cubic_lattice(m):
    ax = (np.arange(m) + 0.5) / m
    pts = np.array([[a, b, c] for a in ax for b in ax for c in ax])
    r = orc.compute(pts, LO, HI, clip=orc.CLIP_FACES | orc.CHECK)
    assert np.all(r["stats"]["status"] == 0)
    assert np.allclose(r["volume"], 1 / m ** 3, rtol=1e-13, atol=0)
    assert np.all(r["face_count"] == 6)
    for i, p in enumerate(pts):  # face-adjacent neighbours only
        nb, ar = csr(r, i)
        for j in nb[nb >= 0]:
            assert np.isclose(np.abs(pts[j] - p).sum(), 1 / m)
        assert np.allclose(ar, 1 / m ** 2, rtol=1e-13)


def _interior(pts, spacing, margin):
    return np.all((pts > margin * spacing) & (pts < 1 - margin * spacing), axis=1)


def test_bcc_truncated_octahedra():
    m = 4
    a = 1.0 / m
    base = np.array([[i, j, k] for i in range(m) for j in range(m) for k in range(m)], float) * a
    pts = np.vstack([base + 0.25 * a, base + 0.75 * a])
    r = orc.compute(pts, LO, HI, clip=orc.CLIP_FACES | orc.CHECK)
    inner = _interior(pts, a, 1.0)
    assert inner.sum() > 0
    assert np.allclose(r["volume"][inner], a ** 3 / 2, rtol=1e-12)
    assert np.all(r["face_count"][inner] == 14)
    assert abs(r["volume"].sum() - 1) < 1e-12


def test_fcc_rhombic_dodecahedra():
    m = 3
    a = 1.0 / m
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pts = np.array([(np.array([i, j, k]) + b + 0.25) * a for i in range(m) for j in range(m) for k in range(m)
                    for b in basis])
    r = orc.compute(pts, LO, HI, clip=orc.CLIP_FACES | orc.CHECK)
    inner = _interior(pts, a, 0.8)
    assert inner.sum() > 0
    assert np.allclose(r["volume"][inner], a ** 3 / 4, rtol=1e-12)
    assert np.all(r["face_count"][inner] == 12)
    assert abs(r["volume"].sum() - 1) < 1e-12


@pytest.mark.parametrize("n,seed", [(60, 1), (300, 2), (1500, 3)])
def test_security_stop_equals_all_points(n, seed):
    pts = tess.uniform(n, seed)
    a = orc.compute(pts, LO, HI, mode=orc.MODE_ALL_POINTS)
    b = orc.compute(pts, LO, HI, mode=orc.MODE_SECURITY)
    for k in ("volume", "face_count", "face_offset", "neighbor", "area"):
        assert np.array_equal(a[k], b[k]), k
    assert np.all(b["stats"]["streamed"] <= a["stats"]["streamed"])


@pytest.mark.parametrize("n,seed,density", [(500, 4, 1.25), (3000, 5, 5.0), (3000, 6, 0.5)])
def test_walk_equals_per_face_clip(n, seed, density):
    pts = tess.uniform(n, seed)
    a = orc.compute(pts, LO, HI, density=density, clip=orc.CLIP_WALK | orc.CHECK)
    b = orc.compute(pts, LO, HI, density=density, clip=orc.CLIP_FACES | orc.CHECK)
    assert np.all(a["stats"]["status"] == 0) and np.all(b["stats"]["status"] == 0)
    for k in ("volume", "face_count", "face_offset", "neighbor", "area"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("streamed", "classified", "cuts", "flops"):
        assert np.array_equal(a["stats"][k], b["stats"][k]), k


@pytest.mark.parametrize("n,seed", [(2000, 7), (20000, 8)])
def test_uniform_invariants(n, seed):
    pts = tess.uniform(n, seed)
    r = orc.compute(pts, LO, HI, density=5.0)
    assert abs(r["volume"].sum() - 1.0) < 1e-10
    nbrs = {}
    for i in range(n):
        nb, ar = csr(r, i)
        assert np.all(np.diff(nb) > 0)  # sorted, unique
        nbrs[i] = dict(zip(nb.tolist(), ar.tolist()))
    for i in range(n):
        for j, a in nbrs[i].items():
            if j >= 0:
                assert i in nbrs[j]  # symmetry
                assert abs(nbrs[j][i] - a) <= 1e-9 * max(a, 1.0 / n)  # shared face (absolute near 0)


def test_against_qhull_interior_cells():
    """Interior cells (whole Voronoi cell inside the box) against Qhull's
    Voronoi diagram: same neighbour sets, same volumes."""
    n = 3000
    pts = tess.uniform(n, 21)
    r = orc.compute(pts, LO, HI)
    vor = Voronoi(pts)
    adj = {i: set() for i in range(n)}
    for a, b in vor.ridge_points:
        adj[a].add(int(b))
        adj[b].add(int(a))
    checked = 0
    for i in range(n):
        reg = vor.regions[vor.point_region[i]]
        if -1 in reg or not reg:
            continue
        verts = vor.vertices[reg]
        if np.any(verts <= 0) or np.any(verts >= 1):
            continue
        nb, _ = csr(r, i)
        assert set(nb.tolist()) == adj[i], i
        assert abs(ConvexHull(verts).volume - r["volume"][i]) <= 1e-9 * r["volume"][i]
        checked += 1
    assert checked > n // 2


def test_poisson_mean_face_count():
    """Interior cells of a Poisson-Voronoi tessellation have 2 + 48 pi^2 / 35
    = 15.535 faces on average (Meijering 1953)."""
    n = 40000
    pts = tess.uniform(n, 31)
    r = orc.compute(pts, LO, HI, density=5.0)
    inner = np.all((pts > 0.15) & (pts < 0.85), axis=1)
    fc = []
    for i in np.nonzero(inner)[0]:
        nb, _ = csr(r, i)
        fc.append(int((nb >= 0).sum()))
    assert abs(np.mean(fc) - (2 + 48 * np.pi ** 2 / 35)) < 0.08


def test_incident_only_cuts_counted():
    """A simple cubic lattice: the bisectors pass through vertices, so some
    cuts cut off a region bordered by Incident vertices only.  The reference's
    find_outgoing_edge (polyhedron.rs:418-432) needs an Inside -> Outside edge
    and skips them -- the walk clip, which keeps that rule, counts none and
    gets the wrong cells -- while the per-face clip cuts them (exact 1/64
    volumes) and counts each one."""
    g = (np.arange(4) + 0.5) / 4
    pts = np.array([(x, y, z) for x in g for y in g for z in g])
    faces = orc.compute(pts, LO, HI, clip=orc.CLIP_FACES)
    walk = orc.compute(pts, LO, HI, clip=orc.CLIP_WALK)
    assert faces["stats"]["incident_cuts"].sum() > 0
    assert np.abs(faces["volume"] - 1 / 64).max() < 1e-15
    assert walk["stats"]["incident_cuts"].sum() == 0
    assert np.abs(walk["volume"] - 1 / 64).max() > 1e-3
    # uniform points in general position never cut that way
    u = orc.compute(tess.uniform(2000, 5), LO, HI)
    assert u["stats"]["incident_cuts"].sum() == 0
