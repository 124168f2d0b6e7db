"""Radius queries of the oracle (celery.rs:700-855, interface.rs:348-365)
against the reference's tests (celery.rs:1459-1901) and against brute force.

The reference's four query tests use 79 points in [-2, 2]^3 with anchors at
(2, 2, 2) and (-2, -2, -2), so that the grid is 4x4x4 with whole-number cell
bounds; the random points are seeded here.
"""
import itertools

import numpy as np
import pytest

import tess
from oracle import oracle as orc


def _celery_79(seed=9, extra=()):
    """generate_random_points(-2, 2, ...) + big_pt + little_pt (+ extra)."""
    n_random = 79 - 2 - len(extra)
    pts = tess.uniform(n_random, seed, [-2.0] * 3, [2.0] * 3)
    pts = np.vstack([pts, [[2.0, 2.0, 2.0], [-2.0, -2.0, -2.0]]] + ([np.array(extra, float)] if extra else []))
    g, bins, srt, dl = orc.grid(pts)
    assert g.cpd == 4 and (g.xmin, g.xmax, g.sx) == (-2.0, 2.0, 1.0)
    return pts, g, bins, srt, dl


def _touching_ge1(i, j, k):
    return (i == 0 and j >= 1 and k >= 1) or (i >= 1 and j == 0 and k >= 1) or \
        (i >= 1 and j >= 1 and k == 0) or (i >= 1 and j >= 1 and k >= 1)


def test_check_cell_in_range():
    """celery.rs:1489-1665."""
    _, g, _, _, _ = _celery_79()
    for i, j, k in itertools.product(range(4), repeat=3):
        for p in (0.1, 0.9):
            inr = lambda r: orc.check_cell_in_range(g, p, p, p, r, i, j, k)  # noqa: E731
            assert inr(0.5) == (i != 0 and j != 0 and k != 0)
            assert inr(1.1) == _touching_ge1(i, j, k)
            assert inr(1.7) == (i + j + k > 0)
            assert inr(1.8)


def test_find_cells_in_radius():
    """celery.rs:1667-1764."""
    _, g, _, _, _ = _celery_79()
    cell = lambda i, j, k: 16 * i + 4 * j + k  # noqa: E731
    q = lambda p, r: set(orc.find_cells_in_radius(g, p, p, p, r).tolist())  # noqa: E731
    close1, close2, med1, med2 = q(0.1, 0.5), q(0.9, 0.5), q(0.1, 1.2), q(0.9, 1.2)
    far1, far2, all1, all2 = q(0.1, 1.7), q(0.9, 1.7), q(0.1, 1.8), q(0.9, 1.8)
    for i, j, k in itertools.product(range(4), repeat=3):
        c = cell(i, j, k)
        assert (c in close1) == (i in (1, 2) and j in (1, 2) and k in (1, 2))
        assert (c in close2) == (i in (2, 3) and j in (2, 3) and k in (2, 3))
        assert (c in med1) == _touching_ge1(i, j, k)
        assert (c in med2) == (i >= 1 and j >= 1 and k >= 1)
        assert (c in far1) == (i + j + k > 0)
        assert (c in far2) == (i != 0 and j != 0 and k != 0)
        assert (c in all2) == (i != 0 and j != 0 and k != 0)
        assert c in all1


_SPECIAL = {  # index: (cell-radius test point, real-radius test point)
    73: (0.1, 0.1), 74: (-0.3, -0.1), 75: (-0.7, -0.7), 76: (-1.1, -1.1), 77: (1.3, 1.3), 78: (1.7, 1.7)}


@pytest.mark.parametrize("real", [False, True])
def test_find_neighbors_in_radius(real):
    """celery.rs:1766-1833 (cell radius) and :1835-1901 (real radius)."""
    extra = [[v[int(real)]] * 3 for _, v in sorted(_SPECIAL.items())]
    pts, g, _, srt, dl = _celery_79(seed=11, extra=extra)
    nb = orc.find_neighbors(pts, g, srt, dl, 0.5, 0.5, 0.5, 1.73, real=real).tolist()
    want = {73: 1, 74: 1, 75: 0 if real else 1, 76: 0, 77: 1, 78: 0 if real else 1}
    for i, c in want.items():
        assert nb.count(i) == c, i


def test_neighbor_cloud_known_answers():
    """Cell::compute_neighbor_cloud is expand_all_in_radius from the cell's
    position (celery.rs:1459-1487's answers, through the batched query)."""
    pts, _, _, _, _ = _celery_79()
    off, ids = orc.queries(orc.QUERY_CLOUD, pts, [[0.0, 0.0, 0.0]] * 2, [0.5, 10.0])
    small, large = ids[off[0]:off[1]].tolist(), ids[off[1]:off[2]].tolist()
    assert 77 in small and 78 not in small and len(small) < 79
    assert sorted(large) == list(range(79))
    assert small == orc.expand_all(pts, (0, 0, 0), max_radius=0.5).tolist()


# ---- brute force: what the geometry fixes, independent of the grid code ----
def _brute_setup(n=3000, seed=5, density=1.25):
    pts = tess.clustered(n, seed)
    g, bins, srt, dl = orc.grid(pts, density)
    return pts, g, bins


def _queries(pts, seed, m=60):
    rng = np.random.default_rng(seed)
    inside = pts[rng.integers(0, len(pts), m // 2)] + rng.normal(0, 0.01, (m // 2, 3))
    outside = rng.uniform(-0.3, 1.3, (m - m // 2, 3))  # clamped bins outside the box
    q = np.vstack([inside, outside])
    r = rng.uniform(0.0, 0.25, m)
    return q, r


def _cell_ijk(g, p):
    return tuple(orc.axis_index(v, lo, hi, inv, g.cpd) for v, lo, hi, inv in
                 zip(p, (g.xmin, g.ymin, g.zmin), (g.xmax, g.ymax, g.zmax), (g.ix, g.iy, g.iz)))


def test_real_radius_is_the_ball_in_grid_order():
    """For a query inside the points' box the box of cells covers the ball,
    so the real-radius result is exactly the brute-force ball, listed by
    (cell id, position in the cell); the cell-radius result contains it."""
    pts, g, bins = _brute_setup()
    q, r = _queries(pts, 1)
    off_r, ids_r = orc.queries(orc.QUERY_REAL_RADIUS, pts, q, r)
    off_c, ids_c = orc.queries(orc.QUERY_CELL_RADIUS, pts, q, r)
    lo = np.array([g.xmin, g.ymin, g.zmin])
    hi = np.array([g.xmax, g.ymax, g.zmax])
    for a in range(len(q)):
        got = ids_r[off_r[a]:off_r[a + 1]]
        cell_pts = set(ids_c[off_c[a]:off_c[a + 1]].tolist())
        assert set(got.tolist()) <= cell_pts
        if np.all(q[a] >= lo) and np.all(q[a] <= hi):
            d2 = np.sum((pts - q[a]) ** 2, axis=1)
            ball = np.nonzero(d2 <= r[a] ** 2)[0]
            want = ball[np.lexsort((ball, bins[ball]))]
            assert np.array_equal(got, want)


def test_cells_in_radius_brute_force():
    """Cell ids ascend (i, j, k loop order); a cell is listed iff it lies in
    the clamped index box and its gap distance is within the radius."""
    pts, g, _ = _brute_setup()
    q, r = _queries(pts, 2)
    off, ids = orc.queries(orc.QUERY_CELLS, pts, q, r)
    cpd = g.cpd
    s = np.array([g.sx, g.sy, g.sz])
    for a in range(len(q)):
        got = ids[off[a]:off[a + 1]]
        assert np.all(np.diff(got.astype(np.int64)) > 0)
        lo_ix = _cell_ijk(g, np.maximum(q[a] - r[a], [g.xmin, g.ymin, g.zmin]))
        hi_ix = _cell_ijk(g, np.minimum(q[a] + r[a], [g.xmax, g.ymax, g.zmax]))
        home = np.array(_cell_ijk(g, q[a]))
        want = []
        for c in range(cpd ** 3):
            ijk = np.array([c // (cpd * cpd), (c // cpd) % cpd, c % cpd])
            if np.any(ijk < lo_ix) or np.any(ijk > hi_ix):
                continue
            gx, gy, gz = np.maximum(np.abs(ijk - home) - 1, 0) * s
            if (gx * gx + gy * gy) + gz * gz <= r[a] * r[a]:
                want.append(c)
        assert got.tolist() == want


def test_cloud_is_every_bin_within_the_key():
    """expand_all_in_radius lists the points of the bins whose squared gap
    distance (the search-order key) is <= r (compared unsquared, the quirk of
    celery.rs:1036), bins in (key, generation) order; a group filter keeps
    the subsequence of that group."""
    pts, g, bins = _brute_setup(2000, 7)
    rng = np.random.default_rng(3)
    groups = rng.integers(0, 3, len(pts)).astype(np.int64)
    q, r = _queries(pts, 4, 20)
    r = r ** 2 * 40  # keys are squared lengths
    off, ids = orc.queries(orc.QUERY_CLOUD, pts, q, r)
    off_g, ids_g = orc.queries(orc.QUERY_CLOUD, pts, q, r, groups=groups, target_group=1)
    cpd = g.cpd
    s = np.array([g.sx, g.sy, g.sz])
    for a in range(len(q)):
        home = np.array(_cell_ijk(g, q[a]))
        got = ids[off[a]:off[a + 1]]
        ijk = np.stack([bins // (cpd * cpd), (bins // cpd) % cpd, bins % cpd], 1).astype(np.int64)
        o = ijk - home
        key = np.sum((np.maximum(np.abs(o) - 1, 0) * s) ** 2, axis=1)
        key[np.all(o == 0, axis=1)] = -1.0
        assert set(got.tolist()) == set(np.nonzero(key <= r[a])[0].tolist())
        k = key[got]
        assert np.all(np.diff(k) >= 0)  # bins in key order
        same_bin = bins[got][1:] == bins[got][:-1]
        assert np.all(np.diff(got.astype(np.int64))[same_bin] > 0)  # a bin's points in sorted order
        gg = ids_g[off_g[a]:off_g[a + 1]]
        assert np.array_equal(gg, got[groups[got] == 1])


def test_query_edge_cases():
    """NaN radius: no cells for the box queries (ds <= NaN is false), every
    point for the cloud (key > NaN is false); negative radius keeps the home
    cell of a box query only when the flipped box still holds it."""
    pts, g, bins = _brute_setup(500, 8)
    q = [[0.5, 0.5, 0.5]]
    for kind in (orc.QUERY_CELLS, orc.QUERY_CELL_RADIUS, orc.QUERY_REAL_RADIUS):
        off, _ = orc.queries(kind, pts, q, [np.nan])
        assert off[-1] == 0
    off, ids = orc.queries(orc.QUERY_CLOUD, pts, q, [np.nan])
    assert sorted(ids.tolist()) == list(range(len(pts)))
    off, ids = orc.queries(orc.QUERY_CLOUD, pts, q, [-1.0])  # the centre bin only (key -1)
    i, j, k = _cell_ijk(g, q[0])
    assert ids.tolist() == np.nonzero(bins == (i * g.cpd + j) * g.cpd + k)[0].tolist()
    off, ids = orc.queries(orc.QUERY_CLOUD, pts, q, [-1.5])
    assert off[-1] == 0
    off, ids = orc.queries(orc.QUERY_CELLS, pts, q, [-1e-9])
    assert len(ids) <= 1
    off, ids = orc.queries(orc.QUERY_CELLS, pts, np.zeros((0, 3)), [])
    assert len(off) == 1 and off[0] == 0
