"""Cells past the LDS kernels' limits, against the oracle (which streams the
whole shell order, celery.rs:418-679, :971-1018, and keeps every face ring
with no size limit, pool.rs:85-110):

* a security ball that reaches past the default shell table (|o|_inf <= 12):
  a container much larger than the points' bounding box, a void wider than
  12 grid cells -- recomputed with longer tables;
* more than 256 vertices / 128 faces: a site ringed by hundreds of points on
  a sphere -- the global-memory spill kernel.

Outputs bitwise (volumes and areas included) and the search counters equal;
the radius limit of expand_all_in_radius ends a search without a longer
table."""
import numpy as np
import pytest

import tess
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
LO, HI = np.zeros(3), np.ones(3)
COUNTERS = ("streamed", "classified", "cuts", "entries", "flops", "incident_cuts")


def same_cells(got, ref):
    assert np.array_equal(got["face_count"], ref["face_count"])
    assert np.array_equal(got["face_offset"], ref["face_offset"])
    assert np.array_equal(got["neighbor"], ref["neighbor"])
    assert np.array_equal(got["volume"], ref["volume"])
    assert np.array_equal(got["area"], ref["area"])


def same_counters(st, ref_st, idx=None):
    idx = np.arange(len(st)) if idx is None else idx
    for k in COUNTERS:
        assert np.array_equal(st[k][idx].astype(np.uint64), ref_st[k][idx].astype(np.uint64)), k


def sphere_ring(m=200, seed=7, r=0.3, bg=3000):
    """A centre point, m points on a sphere of radius r around it, and
    uniform background points farther than 1.5 r: the centre's cell is the
    m-faced polytope around the inscribed sphere of radius r / 2."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(m, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    c = np.full(3, 0.5)
    b = tess.uniform(bg, seed + 1)
    b = b[np.linalg.norm(b - c, axis=1) > 1.5 * r]
    return np.vstack([c[None, :], c + r * v, b])


def test_container_larger_than_points(ctx):
    pts = tess.uniform(20000, 2020)
    lo, hi = np.full(3, -1.0), np.full(3, 2.0)
    d = tess.Diagram(ctx, pts, lo, hi, collect_stats=True)
    d.compute()
    got, st = d.cells(), d.stats()
    ref = orc.compute(pts, lo, hi)
    same_cells(got, ref)
    same_counters(st, ref["stats"])
    assert abs(got["volume"].sum() - 27.0) <= 1e-10 * 27.0
    t = d.timings()
    long = (st["status"] & tess.CELL_LONG_TABLE) != 0
    print(f"table reruns {t['table_reruns']}, max entries {st['entries'].max()}")
    assert t["table_reruns"] > 0 and long.sum() > 0
    assert np.all(st["status"] & ~np.uint32(tess.CELL_LONG_TABLE | tess.CELL_RERUN) == 0)


def test_void_wider_than_table(ctx):
    """Two dense blobs in opposite corners: the cells facing the void reach
    across it."""
    a = tess.uniform(6000, 31, lo=(0, 0, 0), hi=(0.1, 0.1, 0.1))
    b = tess.uniform(6000, 32, lo=(0.9, 0.9, 0.9), hi=(1, 1, 1))
    pts = np.vstack([a, b])
    d = tess.Diagram(ctx, pts, LO, HI, collect_stats=True)
    d.compute()
    ref = orc.compute(pts, LO, HI)
    same_cells(d.cells(), ref)
    same_counters(d.stats(), ref["stats"])
    assert d.timings()["table_reruns"] > 0


def test_radius_limit_is_a_valid_stop(ctx):
    """expand_all_in_radius stops at the radius (a squared key against r,
    celery.rs:1036): a limit below the table's end is no failure."""
    pts = tess.uniform(20000, 2020)
    lo, hi = np.full(3, -1.0), np.full(3, 2.0)
    d = tess.Diagram(ctx, pts, lo, hi)
    corner = int(np.argmin(pts.sum(axis=1)))
    for r in (0.002, 0.5, 3.0):
        c = d.get_cell_at_index(corner, search_radius=r)
        ref = orc.compute(pts, lo, hi, cells=[corner], max_radius=r)
        assert np.array_equal(c.compute_neighbors(), ref["neighbor"])
        assert c.compute_volume() == ref["volume"][0]


def test_single_cell_long_table(ctx):
    pts = tess.uniform(20000, 2020)
    lo, hi = np.full(3, -1.0), np.full(3, 2.0)
    d = tess.Diagram(ctx, pts, lo, hi)
    corner = int(np.argmin(pts.sum(axis=1)))
    c = d.get_cell_at_index(corner)
    ref = orc.compute(pts, lo, hi, cells=[corner])
    assert np.array_equal(c.compute_neighbors(), ref["neighbor"])
    assert c.compute_volume() == ref["volume"][0]


@pytest.mark.parametrize("capacity", [0, 256])
def test_sphere_ring_spills(ctx, capacity):
    pts = sphere_ring()
    d = tess.Diagram(ctx, pts, LO, HI, collect_stats=True, vertex_capacity=capacity)
    d.compute()
    got, st = d.cells(), d.stats()
    ref = orc.compute(pts, LO, HI)
    same_cells(got, ref)
    same_counters(st, ref["stats"])
    assert got["face_count"][0] == 200
    assert np.array_equal(got["neighbor"][:200], np.arange(1, 201))
    assert st["status"][0] & tess.CELL_SPILL
    assert st["max_vertices"][0] > 256
    assert d.timings()["spill_cells"] >= 1
    assert abs(got["volume"].sum() - 1.0) <= 1e-10


@pytest.mark.parametrize("batched", [False, True])
def test_sphere_ring_of_100_reruns_twice(ctx, batched):
    """100 faces, about 196 vertices: past the first pass (64 vertices) and
    the first rerun (128), finished by the second rerun at 256 vertices."""
    pts = sphere_ring(m=100, seed=13, r=0.3, bg=3000)
    ref = orc.compute(pts, LO, HI)
    if batched:
        for eager in (0, 1):
            b = tess.Batch(ctx, [(pts, LO, HI), (tess.uniform(2000, 14), LO, HI)])
            b.set_config(eager=eager)
            b.compute()
            same_cells(b.cells(0), ref)
            assert b.timings()["capacity_reruns"] >= 1 and b.timings()["spill_cells"] == 0
            b.close()
        return
    d = tess.Diagram(ctx, pts, LO, HI, collect_stats=True)
    d.compute()
    got, st = d.cells(), d.stats()
    same_cells(got, ref)
    same_counters(st, ref["stats"])
    assert got["face_count"][0] == 100
    assert 128 < st["max_vertices"][0] <= 256
    assert st["status"][0] & tess.CELL_RERUN
    assert not st["status"][0] & tess.CELL_SPILL
    assert d.timings()["spill_cells"] == 0


def test_sphere_ring_of_600_grows_the_workspace(ctx):
    """600 faces, about 1200 vertices: past the first spill workspace (512
    faces), so the kernel reruns the cell with a larger one."""
    pts = sphere_ring(m=600, seed=9, r=0.25, bg=2000)
    d = tess.Diagram(ctx, pts, LO, HI, collect_stats=True)
    d.compute()
    got = d.cells()
    ref = orc.compute(pts, LO, HI)
    same_cells(got, ref)
    same_counters(d.stats(), ref["stats"])
    assert got["face_count"][0] == 600


def test_sites_past_the_table(ctx):
    pts = tess.uniform(20000, 2020)
    lo, hi = np.full(3, -1.0), np.full(3, 2.0)
    d = tess.Diagram(ctx, pts, lo, hi, collect_stats=True)
    sites = np.array([[-0.9, -0.9, -0.9], [1.9, 1.9, 1.9], [0.5, 0.5, 0.5], [-0.5, 1.5, 0.2]])
    got = d.cells_at_particles(sites, stats=True)
    ref = orc.compute_at_points(pts, sites, lo, hi)
    for k in ("face_count", "neighbor", "volume", "area"):
        assert np.array_equal(got[k], ref[k]), k
    same_counters(got["stats"], ref["stats"])


def test_geometry_with_long_tables(ctx):
    a = tess.uniform(3000, 41, lo=(0, 0, 0), hi=(0.1, 0.1, 0.1))
    b = tess.uniform(3000, 42, lo=(0.9, 0.9, 0.9), hi=(1, 1, 1))
    pts = np.vstack([a, b])
    for staged_by_compute in (True, False):
        d = tess.Diagram(ctx, pts, LO, HI)
        if not staged_by_compute:
            d.compute()
        geo = d.geometry()
        cells = d.cells()
        ref = orc.compute_geometry(pts, LO, HI)
        assert np.array_equal(cells["neighbor"], ref["neighbor"])
        assert np.array_equal(geo["ring_offset"], ref["ring_offset"])
        fc = cells["face_count"].astype(np.int64)
        ro = geo["ring_offset"].astype(np.int64)
        entry_cell = np.repeat(np.repeat(np.arange(len(fc)), fc), np.diff(ro))
        xyz = geo["vertices"][geo["vertex_offset"][entry_cell].astype(np.int64) + geo["ring"]]
        assert np.array_equal(xyz, ref["ring_xyz"])


@pytest.mark.parametrize("eager", [0, 1])
def test_batch_with_long_tables_and_spills(ctx, eager):
    sets = [(tess.uniform(4000, 51), LO, HI),
            (sphere_ring(seed=11), LO, HI),
            (tess.uniform(5000, 52), np.full(3, -1.0), np.full(3, 2.0)),
            (tess.uniform(3000, 53, lo=(0.2, 0.3, 0.1), hi=(1.2, 0.8, 1.9)), np.array([0.2, 0.3, 0.1]),
             np.array([1.2, 0.8, 1.9]))]
    b = tess.Batch(ctx, sets)
    b.set_config(eager=eager)
    b.compute()
    t = b.timings()
    assert t["table_reruns"] > 0 and t["spill_cells"] >= 1
    for s, (p, lo, hi) in enumerate(sets):
        same_cells(b.cells(s), orc.compute(p, lo, hi))


def test_learned_radius_routes_long_searches(ctx):
    """The model's predicted final radius R sends the cells whose security
    ball will reach past the default table (4 R^2 >= kcut) straight to a pass
    with a table long enough for the largest such prediction; the others run
    as planned, and a cell the prediction under-sized still gets its exact
    longer-table rerun (counted as a fallback).  Trained on three sets of the
    same kind (20,000 uniform points in a [-1, 2]^3 container) by the GPU
    kernels; outputs and counters bitwise equal to the plain path."""
    lo, hi = np.full(3, -1.0), np.full(3, 2.0)
    m = tess.Model(ctx, seed=5)
    for s in (2101, 2102, 2103):
        t = tess.Diagram(ctx, tess.uniform(20000, s), lo, hi, collect_stats=True)
        t.compute()
        m.add_samples(t)
        t.close()
    m.train(epochs=150, batch=2048, lr=2e-3, seed=1)
    pts = tess.uniform(20000, 2104)
    plain = tess.Diagram(ctx, pts, lo, hi, collect_stats=True)
    plain.compute()
    d = tess.Diagram(ctx, pts, lo, hi, collect_stats=True)
    d.compute_learned(m)
    same_cells(d.cells(), plain.cells())
    same_counters(d.stats(), plain.stats())
    tp, tl = plain.timings(), d.timings()
    print(f"plain: {tp['table_reruns']} long-table reruns, {tp['total_ms']:.2f} ms; learned: "
          f"{tl['planned_long']} cells routed long, {tl['table_reruns']} reruns, {tl['fallback_cells']} fallbacks, "
          f"{tl['total_ms']:.2f} ms")
    assert tl["planned_long"] > 0
    assert tl["table_reruns"] < tp["table_reruns"]
    st = d.stats()["status"]
    flagged = int(((st & tess.CELL_FALLBACK) != 0).sum())
    assert flagged <= tl["fallback_cells"] and (flagged > 0) == (tl["fallback_cells"] > 0)
