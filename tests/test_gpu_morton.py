"""The reference's internal (Morton) point order on the GPU: index maps
(Diagram::sort_particles / morton_index, interface.rs:88-183) and diagrams
whose grid cells list their points in that order (TESS_ORDER_MORTON),
against the oracle's restatement, bitwise."""
import numpy as np
import pytest

import tess
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

LO, HI = np.zeros(3), np.ones(3)


@pytest.mark.parametrize("workload", ["uniform", "clustered", "shifted"])
def test_index_maps_match_oracle(ctx, workload):
    if workload == "uniform":
        pts = tess.uniform(50000, 91)
    elif workload == "clustered":
        pts = tess.clustered(50000, 92)
    else:  # negative coordinates: the box grows to contain the origin
        pts = tess.uniform(50000, 93, [-3.0, -1.0, 0.5], [-0.5, 2.0, 4.0])
    pts[10] = pts[20]  # a tie
    d = tess.Diagram(ctx, pts, pts.min(0) - 0.1, pts.max(0) + 0.1, density=5.0)
    ooi, ioo = d.index_maps()
    r_ooi, r_ioo = orc.index_maps(pts)
    assert np.array_equal(ooi, r_ooi) and np.array_equal(ioo, r_ioo)


@pytest.mark.parametrize("workload", ["uniform", "clustered"])
@pytest.mark.parametrize("density", [1.25, 5.0])
def test_morton_order_cells_match_oracle(ctx, workload, density):
    pts = tess.uniform(20000, 94) if workload == "uniform" else tess.clustered(20000, 95)
    d = tess.Diagram(ctx, pts, LO, HI, density=density, point_order=tess.ORDER_MORTON, collect_stats=True)
    d.compute()
    got = d.cells()
    ref = orc.compute(pts, LO, HI, density=density, order=orc.ORDER_MORTON)
    for k in ("volume", "face_count", "face_offset", "neighbor", "area"):
        assert np.array_equal(got[k], ref[k]), k
    st = d.stats()
    for k in ("streamed", "classified", "cuts", "entries", "flops", "incident_cuts"):
        assert np.array_equal(st[k], ref["stats"][k]), k
    _, srt, dl = d.copy_grid()
    _, _, r_srt, r_dl = orc.grid(pts, density, order=orc.ORDER_MORTON)
    assert np.array_equal(srt, r_srt) and np.array_equal(dl, r_dl)
    # the same cells as in input order, to rounding
    d0 = tess.Diagram(ctx, pts, LO, HI, density=density)
    d0.compute()
    c0 = d0.cells()
    assert np.array_equal(c0["neighbor"], got["neighbor"])
    assert np.allclose(c0["volume"], got["volume"], rtol=1e-12, atol=0)


def test_morton_order_options(ctx):
    pts = tess.uniform(2000, 96)
    with pytest.raises(tess.TessError):
        tess.Diagram(ctx, pts, LO, HI, point_order=7)
    d = tess.Diagram(ctx, pts, LO, HI, point_order=tess.ORDER_MORTON)
    off, ids = d.query(tess.QUERY_CELL_RADIUS, [[0.5, 0.5, 0.5]], 0.1)
    bins, srt, dl = d.copy_grid()
    cells = d.find_cells_in_radius(0.5, 0.5, 0.5, 0.1)
    want = np.concatenate([srt[dl[c]:dl[c + 1]] for c in cells])
    assert np.array_equal(ids, want)  # a cell's points in the internal order


@pytest.mark.parametrize("kind", [tess.QUERY_CELL_RADIUS, tess.QUERY_REAL_RADIUS, tess.QUERY_CLOUD])
def test_morton_order_queries_match_oracle(ctx, kind):
    """In the internal order a cell's points come by ascending internal index
    in every query kind that lists points."""
    pts = tess.clustered(20000, 97)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0, point_order=tess.ORDER_MORTON)
    rng = np.random.default_rng(98)
    q = pts[rng.integers(0, len(pts), 5000)]
    r = 0.02 if kind != tess.QUERY_CLOUD else 4e-4
    off, ids = d.query(kind, q, r)
    ro, rid = orc.queries(kind, pts, q, r, density=5.0, order=orc.ORDER_MORTON)
    assert np.array_equal(off, ro) and np.array_equal(ids, rid)


@pytest.mark.parametrize("computed", [False, True])
def test_morton_order_geometry_matches_oracle(ctx, computed):
    """Vertices and face rings of a diagram in the internal order, staged by
    its compute pass or by the separate staging pass, bitwise against the
    oracle's rings in that order."""
    pts = tess.clustered(20000, 97)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0, point_order=tess.ORDER_MORTON)
    if computed:
        d.compute()
    geo = d.geometry()
    cells = d.cells()
    ref = orc.compute_geometry(pts, LO, HI, density=5.0, order=orc.ORDER_MORTON)
    for k in ("volume", "neighbor", "area"):
        assert np.array_equal(cells[k], ref[k]), k
    assert np.array_equal(geo["ring_offset"], ref["ring_offset"])
    fc = cells["face_count"].astype(np.int64)
    entry_cell = np.repeat(np.repeat(np.arange(len(fc)), fc), np.diff(geo["ring_offset"].astype(np.int64)))
    xyz = geo["vertices"][geo["vertex_offset"][entry_cell].astype(np.int64) + geo["ring"]]
    assert np.array_equal(xyz, ref["ring_xyz"])


@pytest.mark.parametrize("m", [3000, 20000])  # sites in input order; in grid order
def test_morton_order_sites_match_oracle(ctx, m):
    """Cells of outside sites over a diagram in the internal order, bitwise
    against the oracle in that order, counters included."""
    pts = tess.clustered(20000, 98)
    sites = np.random.default_rng(99).uniform(0, 1, (m, 3))
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0, point_order=tess.ORDER_MORTON, collect_stats=True)
    got = d.cells_at_particles(sites, stats=True)
    ref = orc.compute_at_points(pts, sites, LO, HI, density=5.0, order=orc.ORDER_MORTON)
    for k in ("volume", "face_count", "face_offset", "neighbor", "area"):
        assert np.array_equal(got[k], ref[k]), k
    for k in ("streamed", "classified", "cuts", "entries", "flops", "incident_cuts"):
        assert np.array_equal(got["stats"][k], ref["stats"][k]), k
