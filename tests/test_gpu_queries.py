"""Radius queries on the GPU (tess_diagram_query): find_cells_in_radius,
find_neighbors_in_{cell,real}_radius (celery.rs:700-855) and
Cell::compute_neighbor_cloud (interface.rs:348-365).

Ids and their order must equal the oracle's exactly: the reference's known
answers (celery.rs:1459-1901) through the device path, seeded random queries
against the oracle at sizes it finishes in seconds, and at a million points
the properties the geometry fixes (the real-radius result is the
brute-force ball, listed by grid cell).
"""
import numpy as np
import pytest

import tess
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KINDS = (tess.QUERY_CELLS, tess.QUERY_CELL_RADIUS, tess.QUERY_REAL_RADIUS, tess.QUERY_CLOUD)


def _celery_79(seed=9, extra=()):
    n_random = 79 - 2 - len(extra)
    pts = tess.uniform(n_random, seed, [-2.0] * 3, [2.0] * 3)
    return np.vstack([pts, [[2.0, 2.0, 2.0], [-2.0, -2.0, -2.0]]] + ([np.array(extra, float)] if extra else []))


def _diagram(ctx, pts, density, groups=None):
    lo, hi = pts.min(0) - 0.1, pts.max(0) + 0.1
    return tess.Diagram(ctx, pts, lo, hi, groups=groups, density=density)


def test_known_answers(ctx):
    """celery.rs:1459-1487 (cloud), :1667-1764 (cells), :1766-1901 (neighbours)."""
    pts = _celery_79()
    d = _diagram(ctx, pts, 1.25)
    cloud = d.query(tess.QUERY_CLOUD, [[0, 0, 0]] * 2, [0.5, 10.0])
    small = cloud[1][cloud[0][0]:cloud[0][1]].tolist()
    large = cloud[1][cloud[0][1]:cloud[0][2]].tolist()
    assert 77 in small and 78 not in small and len(small) < 79 and sorted(large) == list(range(79))
    cell = lambda i, j, k: 16 * i + 4 * j + k  # noqa: E731
    close1 = set(d.find_cells_in_radius(0.1, 0.1, 0.1, 0.5).tolist())
    far2 = set(d.find_cells_in_radius(0.9, 0.9, 0.9, 1.7).tolist())
    all1 = set(d.find_cells_in_radius(0.1, 0.1, 0.1, 1.8).tolist())
    for i in range(4):
        for j in range(4):
            for k in range(4):
                c = cell(i, j, k)
                assert (c in close1) == (i in (1, 2) and j in (1, 2) and k in (1, 2))
                assert (c in far2) == (i != 0 and j != 0 and k != 0)
                assert c in all1
    d.close()
    for real, extra, want in [
            (False, [-0.3, -0.7, -1.1, 1.3, 1.7], {73: 1, 74: 1, 75: 1, 76: 0, 77: 1, 78: 1}),
            (True, [-0.1, -0.7, -1.1, 1.3, 1.7], {73: 1, 74: 1, 75: 0, 76: 0, 77: 1, 78: 0})]:
        pts = _celery_79(11, [[0.1] * 3] + [[v] * 3 for v in extra])
        d = _diagram(ctx, pts, 1.25)
        f = d.find_neighbors_in_real_radius if real else d.find_neighbors_in_cell_radius
        nb = f(0.5, 0.5, 0.5, 1.73).tolist()
        for i, c in want.items():
            assert nb.count(i) == c, (real, i)
        d.close()


def _random_queries(pts, seed, m):
    rng = np.random.default_rng(seed)
    lo, hi = pts.min(0), pts.max(0)
    near = pts[rng.integers(0, len(pts), m // 2)] + rng.normal(0, 0.01, (m // 2, 3))
    anywhere = lo + (hi - lo) * rng.uniform(-0.2, 1.2, (m - m // 2, 3))  # some outside the grid
    q = np.vstack([near, anywhere])
    r = rng.uniform(0, 0.15, m) * (hi - lo).max()
    r[:4] = [np.nan, -0.5, 0.0, np.inf]
    return q, r


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("workload", ["uniform", "clustered"])
@pytest.mark.parametrize("density", [1.25, 5.0])
def test_parity_with_oracle(ctx, kind, workload, density):
    n = 20000
    pts = tess.uniform(n, 41) if workload == "uniform" else tess.clustered(n, 42)
    d = _diagram(ctx, pts, density)
    q, r = _random_queries(pts, 43, 300)
    if kind == tess.QUERY_CLOUD:
        r = np.where(np.isfinite(r), np.sign(r) * r * r * 2.0, r)  # keys are squared lengths
        r[3] = 1e300  # a finite radius beyond every key: the full order
    off, ids = d.query(kind, q, r)
    ro, rid = orc.queries(kind, pts, q, r, density=density)
    assert np.array_equal(off, ro)
    assert np.array_equal(ids, rid)
    # one radius for all queries
    off1, ids1 = d.query(kind, q, 0.05)
    ro1, rid1 = orc.queries(kind, pts, q, 0.05, density=density)
    assert np.array_equal(off1, ro1) and np.array_equal(ids1, rid1)
    d.close()


def test_cloud_groups(ctx):
    n = 15000
    pts = tess.clustered(n, 7)
    groups = np.random.default_rng(8).integers(0, 4, n).astype(np.int64)
    d = _diagram(ctx, pts, 5.0, groups=groups)
    q, r = _random_queries(pts, 9, 200)
    r = np.where(np.isfinite(r), np.sign(r) * r * r * 2.0, r)
    for g in (0, 3, 7):
        off, ids = d.query(tess.QUERY_CLOUD, q, r, target_group=g)
        ro, rid = orc.queries(tess.QUERY_CLOUD, pts, q, r, density=5.0, groups=groups, target_group=g)
        assert np.array_equal(off, ro) and np.array_equal(ids, rid)
        assert np.all(groups[ids] == g)
    # the neighbour cloud of a cell is the query at its point's position
    nb = d.neighbor_cloud(123, 0.002, target_group=1)
    _, rid = orc.queries(tess.QUERY_CLOUD, pts, pts[123:124], 0.002, density=5.0, groups=groups, target_group=1)
    assert np.array_equal(nb, rid)
    d.close()


def test_query_after_compute_and_empty(ctx):
    """Queries share the grid of compute; no queries is an empty CSR."""
    pts = tess.uniform(5000, 3)
    d = _diagram(ctx, pts, 5.0)
    d.compute()
    before = d.cells()
    off, ids = d.query(tess.QUERY_REAL_RADIUS, pts[:50], 0.05)
    ro, rid = orc.queries(tess.QUERY_REAL_RADIUS, pts, pts[:50], 0.05, density=5.0)
    assert np.array_equal(off, ro) and np.array_equal(ids, rid)
    off, ids = d.query(tess.QUERY_CELLS, np.zeros((0, 3)), np.zeros(0))
    assert off.tolist() == [0] and len(ids) == 0
    after = d.cells()
    for k in before:
        assert np.array_equal(before[k], after[k])
    d.close()


def test_query_errors(ctx):
    pts = tess.uniform(1000, 4)
    d = _diagram(ctx, pts, 5.0)
    with pytest.raises(tess.TessError):
        d.query(7, pts[:3], 0.1)
    with pytest.raises(tess.TessError):
        d.query(tess.QUERY_CLOUD, pts[:3], 0.1, target_group=1)  # no groups
    with pytest.raises(tess.TessError):
        d.query(tess.QUERY_CELLS, pts[:3], [0.1, 0.2])  # 2 radii for 3 queries
    g = _diagram(ctx, pts, 5.0, groups=np.zeros(1000, np.int64))
    with pytest.raises(tess.TessError):
        g.query(tess.QUERY_REAL_RADIUS, pts[:3], 0.1, target_group=0)  # groups filter clouds only
    d.close()
    g.close()


def test_million_points_properties(ctx):
    """At 1M clustered points: CSR well formed, the real-radius result is the
    brute-force ball in (grid cell, position) order, the cell-radius result
    contains it, and a cloud over every key lists every point once."""
    n = 1_000_000
    pts = tess.clustered(n, 77)
    d = _diagram(ctx, pts, 5.0)
    rng = np.random.default_rng(78)
    q = pts[rng.integers(0, n, 2000)]
    off_r, ids_r = d.query(tess.QUERY_REAL_RADIUS, q, 0.01)
    off_c, ids_c = d.query(tess.QUERY_CELL_RADIUS, q, 0.01)
    assert np.all(np.diff(off_r.astype(np.int64)) >= 0) and off_r[-1] == len(ids_r)
    bins, _, _ = d.copy_grid()
    for a in range(0, 2000, 97):
        got = ids_r[off_r[a]:off_r[a + 1]]
        d2 = np.sum((pts - q[a]) ** 2, axis=1)
        ball = np.nonzero(d2 <= 0.01 * 0.01)[0]
        assert np.array_equal(got, ball[np.lexsort((ball, bins[ball]))])
        assert set(got.tolist()) <= set(ids_c[off_c[a]:off_c[a + 1]].tolist())
    off, ids = d.query(tess.QUERY_CLOUD, q[:2], np.nan)
    for a in range(2):
        seg = ids[off[a]:off[a + 1]]
        assert len(seg) == n and np.array_equal(np.sort(seg), np.arange(n, dtype=np.uint32))
    d.close()


def test_device_inputs(ctx):
    """tess_diagram_query_device: positions and radii already in HBM give the
    host-input results, and the results stay on the device."""
    pts = tess.clustered(20000, 5)
    d = _diagram(ctx, pts, 5.0)
    q, r = _random_queries(pts, 6, 256)
    qt = tess.DeviceArray(ctx, q)
    for kind in KINDS:
        rr = np.where(np.isfinite(r), np.sign(r) * r * r * 2.0, r) if kind == tess.QUERY_CLOUD else r
        rt = tess.DeviceArray(ctx, rr)
        res = d.query_device(kind, qt.ptr, len(q), rt.ptr, len(rr))
        off, ids = res.copy()
        t = res.timings()
        assert t["count_ms"] > 0 and t["fill_ms"] > 0 and t["total_ms"] >= t["fill_ms"]
        res.close()
        ro, rid = orc.queries(kind, pts, q, rr, density=5.0)
        assert np.array_equal(off, ro) and np.array_equal(ids, rid)
        rt.close()
    qt.close()
    d.close()


def test_device_resident_diagram(ctx):
    """tess_diagram_create_device (points and groups already in HBM) gives
    the cells and query results of the host-input diagram."""
    pts = tess.clustered(30000, 12)
    groups = np.random.default_rng(13).integers(0, 2, len(pts)).astype(np.int64)
    lo, hi = np.zeros(3), np.ones(3)
    ref = tess.Diagram(ctx, pts, lo, hi, groups=groups, density=5.0)
    ref.compute()
    want = ref.cells()
    pd, gd = tess.DeviceArray(ctx, pts), tess.DeviceArray(ctx, groups)
    d = tess.Diagram(ctx, pd, lo, hi, groups=gd, density=5.0)
    d.compute()
    got = d.cells()
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    q = pts[:100]
    a = d.query(tess.QUERY_CLOUD, q, 1e-4, target_group=1)
    b = ref.query(tess.QUERY_CLOUD, q, 1e-4, target_group=1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    d.close()
    ref.close()
    pd.close()
    gd.close()


@pytest.mark.parametrize("kind", KINDS)
def test_large_query_sets_in_grid_order(ctx, kind):
    """Sets of 4096 queries and more run in the grid order of their positions;
    the CSR is still in input order and bitwise the oracle's."""
    pts = tess.clustered(20000, 31)
    d = _diagram(ctx, pts, 5.0)
    q, r = _random_queries(pts, 32, 6000)
    if kind == tess.QUERY_CLOUD:
        r = np.where(np.isfinite(r), np.sign(r) * r * r * 2.0, r)
    off, ids = d.query(kind, q, r)
    ro, rid = orc.queries(kind, pts, q, r, density=5.0)
    assert np.array_equal(off, ro) and np.array_equal(ids, rid)
    d.close()
