"""Vertices and face rings of every cell on the GPU (Cell::compute_vertices,
VoronoiFace::compute_vertices; interface.rs:367-370, :403-406,
polyhedron.rs:885-919) against the oracle's rings (orc_compute_geometry):
every ring's coordinates bitwise, faces in neighbour-list order, rings from
the lexicographically smallest vertex; plus what the topology fixes."""
import numpy as np
import pytest

import tess
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

LO, HI = np.zeros(3), np.ones(3)


def _ring_coords(d, cells, geo):
    fc = cells["face_count"].astype(np.int64)
    ro = geo["ring_offset"].astype(np.int64)
    face_cell = np.repeat(np.arange(len(fc)), fc)
    entry_cell = np.repeat(face_cell, np.diff(ro))
    return geo["vertices"][geo["vertex_offset"][entry_cell].astype(np.int64) + geo["ring"]], entry_cell


@pytest.mark.parametrize("workload", ["uniform", "clustered"])
def test_rings_match_oracle(ctx, workload):
    pts = tess.uniform(20000, 81) if workload == "uniform" else tess.clustered(20000, 82)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0, collect_stats=True)
    d.compute()
    before = (d.cells(), d.stats())
    geo = d.geometry()
    cells = d.cells()
    for k in before[0]:  # the geometry passes leave the cell outputs and counters alone
        assert np.array_equal(before[0][k], cells[k]), k
    assert np.array_equal(before[1], d.stats())
    if workload == "clustered":
        assert d.timings()["capacity_reruns"] > 0  # the capacity-256 geometry rerun is exercised
    ref = orc.compute_geometry(pts, LO, HI, density=5.0)
    assert np.array_equal(cells["neighbor"], ref["neighbor"])
    assert np.array_equal(geo["ring_offset"], ref["ring_offset"])
    xyz, entry_cell = _ring_coords(d, cells, geo)
    assert np.array_equal(xyz, ref["ring_xyz"])
    # every vertex of a cell is on one of its rings, and each ring entry is a vertex of the cell
    nv = np.diff(geo["vertex_offset"].astype(np.int64))
    assert np.all(geo["ring"] < nv[entry_cell])
    used = np.zeros(len(geo["vertices"]), bool)
    used[geo["vertex_offset"][entry_cell].astype(np.int64) + geo["ring"]] = True
    assert used.all()
    # Euler: V - E + F = 2 with E = ring entries / 2
    ne = np.bincount(entry_cell, minlength=len(nv)) // 2
    assert np.all(nv - ne + cells["face_count"] == 2)


def test_geometry_staged_by_compute(ctx):
    """Geometry of a diagram not yet computed: the compute's own cell passes
    (first pass and capacity rerun) stage it. Cell outputs and counters equal
    a plain compute's; rings equal the oracle's."""
    pts = tess.clustered(20000, 86)
    plain = tess.Diagram(ctx, pts, LO, HI, density=5.0, collect_stats=True)
    plain.compute()
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0, collect_stats=True)
    geo = d.geometry()
    cells = d.cells()
    ref_cells = plain.cells()
    for k in ref_cells:
        assert np.array_equal(ref_cells[k], cells[k]), k
    assert np.array_equal(plain.stats(), d.stats())
    assert d.timings()["capacity_reruns"] > 0
    assert geo["timings"]["stage_ms"] == d.timings()["cells_ms"]
    ref = orc.compute_geometry(pts, LO, HI, density=5.0)
    assert np.array_equal(geo["ring_offset"], ref["ring_offset"])
    xyz, _ = _ring_coords(d, cells, geo)
    assert np.array_equal(xyz, ref["ring_xyz"])
    again = d.geometry()  # computed now: the separate staging pass
    for k in ("vertex_offset", "vertices", "ring_offset", "ring"):
        assert np.array_equal(geo[k], again[k]), k


def test_geometry_at_capacity_256(ctx):
    """A diagram whose compute runs at vertex capacity 256 takes the separate
    staging pass; its rings equal those staged by a default compute."""
    pts = tess.clustered(8000, 87)
    ref = tess.Diagram(ctx, pts, LO, HI, density=5.0)
    g_ref = ref.geometry()
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0, vertex_capacity=256)
    geo = d.geometry()
    cells = d.cells()
    assert np.array_equal(cells["neighbor"], ref.cells()["neighbor"])
    assert np.array_equal(geo["ring_offset"], g_ref["ring_offset"])
    assert np.array_equal(_ring_coords(d, cells, geo)[0], _ring_coords(ref, ref.cells(), g_ref)[0])


def test_ring_areas_and_orientation(ctx):
    """The fan area of every ring equals the face area; rings are
    counter-clockwise seen from outside (the site is behind every face)."""
    pts = tess.uniform(5000, 83)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0)
    geo = d.geometry()
    cells = d.cells()
    xyz, _ = _ring_coords(d, cells, geo)
    ro = geo["ring_offset"].astype(np.int64)
    for f in range(0, len(ro) - 1, 37):
        ring = xyz[ro[f]:ro[f + 1]]
        w = np.zeros(3)
        for k in range(1, len(ring) - 1):
            w += np.cross(ring[k] - ring[0], ring[k + 1] - ring[0])
        area = cells["area"][f]
        assert abs(0.5 * np.linalg.norm(w) - area) <= 1e-12 * max(area, 1e-12)
        if area > 1e-9:
            assert np.dot(w, ring[0]) > 0  # outward normal, site at the origin


def _geometry_sizes(d):
    import ctypes as C
    nv, nr = C.c_uint64(), C.c_uint64()
    d.ctx.check(d.lib.tess_diagram_geometry_sizes(d.h, C.byref(nv), C.byref(nr)))
    return nv.value, nr.value


def test_geometry_follows_recompute(ctx):
    pts = tess.uniform(3000, 84)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0)
    with pytest.raises(tess.TessError):
        _geometry_sizes(d)      # nothing computed yet
    g1 = d.geometry()           # runs compute first
    d.compute()                 # invalidates the geometry
    with pytest.raises(tess.TessError):
        _geometry_sizes(d)
    g2 = d.geometry()
    assert g2["timings"]["stage_ms"] > 0 and g2["timings"]["compact_ms"] > 0
    for k in ("vertex_offset", "vertices", "ring_offset", "ring"):
        assert np.array_equal(g1[k], g2[k]), k


def test_geometry_million_points(ctx):
    """1M clustered points: the geometry is consistent at full size."""
    pts = tess.clustered(1_000_000, 85)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0)
    geo = d.geometry()
    cells = d.cells()
    nv = np.diff(geo["vertex_offset"].astype(np.int64))
    ro = geo["ring_offset"].astype(np.int64)
    fc = cells["face_count"].astype(np.int64)
    entry_cell = np.repeat(np.repeat(np.arange(len(fc)), fc), np.diff(ro))
    ne = np.bincount(entry_cell, minlength=len(nv)) // 2
    assert np.all(nv - ne + fc == 2)
    assert np.all(geo["ring"] < nv[entry_cell])
