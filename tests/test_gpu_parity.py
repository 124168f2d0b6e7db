"""GPU parity against the CPU oracle, through the C ABI (libtess.so)."""
import numpy as np
import pytest

import tess
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

UNIT_LO, UNIT_HI = np.zeros(3), np.ones(3)


def _compare(got, ref, vol_tol=1e-12):
    assert np.array_equal(got["face_count"], ref["face_count"])
    assert np.array_equal(got["face_offset"], ref["face_offset"])
    assert np.array_equal(got["neighbor"], ref["neighbor"])
    rv = np.abs(got["volume"] - ref["volume"]) / np.abs(ref["volume"])
    assert rv.max() <= vol_tol, rv.max()
    ra = np.abs(got["area"] - ref["area"]) / np.maximum(np.abs(ref["area"]), 1e-300)
    assert ra.max() <= vol_tol, ra.max()


def test_arith_correctly_rounded(ctx):
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.random(20000), rng.random(20000) * 1e-8, np.exp(rng.normal(0, 30, 20000))])
    b = np.concatenate([rng.random(20000) + 0.5, np.exp(rng.normal(0, 10, 40000))])
    s, r, q = ctx.diag_arith(a, b)
    assert np.array_equal(s, np.sqrt(a))
    assert np.array_equal(r, 1.0 / a)
    assert np.array_equal(q, a / b)


@pytest.mark.parametrize("n,seed,density", [(10_000, 0, 1.25), (200_000, 1, 5.0)])
def test_binning_matches_oracle(ctx, n, seed, density):
    pts = tess.uniform(n, seed)
    d = tess.Diagram(ctx, pts, UNIT_LO, UNIT_HI, density=density)
    d.compute()
    b, s, cpd = d.grid()
    g, bins, srt, dl = orc.grid(pts, density)
    assert cpd == g.cpd
    assert np.array_equal(b, [g.xmin, g.xmax, g.ymin, g.ymax, g.zmin, g.zmax])
    assert np.array_equal(s, [g.sx, g.sy, g.sz])
    cell, sorted_, delim = d.copy_grid()
    assert np.array_equal(cell, bins)
    assert np.array_equal(sorted_, srt)
    assert np.array_equal(delim, dl)


@pytest.mark.parametrize("order", [tess.ORDER_INPUT, tess.ORDER_MORTON])
def test_binning_dense_cells_matches_oracle(ctx, order):
    """Grid cells far past the small-cell insertion sort: a ball of 6,000
    points inside one grid cell (larger than the LDS-staged ranking holds, so
    the global ranking runs; in Morton order all of them share one code and
    tie by index) and a Gaussian cluster whose cells hold tens to hundreds of
    points (the LDS-staged ranking), in a uniform background, shuffled."""
    rng = np.random.default_rng(11)
    ball = np.array([0.31, 0.52, 0.77]) + rng.uniform(-1e-4, 1e-4, (6000, 3))
    clus = np.clip(np.array([0.6, 0.3, 0.4]) + rng.normal(0.0, 0.03, (4000, 3)), 0.0, 1.0)
    pts = rng.permutation(np.concatenate([rng.uniform(0.0, 1.0, (20000, 3)), ball, clus]))
    d = tess.Diagram(ctx, pts, UNIT_LO, UNIT_HI, density=5.0, point_order=order)
    d.compute()
    _, bins, srt, dl = orc.grid(pts, 5.0, order=order)
    counts = np.diff(dl)
    assert counts.max() > 4096 and np.count_nonzero((counts > 32) & (counts <= 4096)) >= 10
    cell, sorted_, delim = d.copy_grid()
    assert np.array_equal(cell, bins)
    assert np.array_equal(delim, dl)
    assert np.array_equal(sorted_, srt)
    d.close()


def test_single_point_closed_form(ctx):
    d = tess.Diagram(ctx, np.array([[0.3, 0.6, 0.2]]), UNIT_LO, UNIT_HI)
    d.compute()
    c = d.cells()
    assert c["volume"][0] == 1.0
    assert list(c["neighbor"]) == [-6, -5, -4, -3, -2, -1]
    assert np.all(c["area"] == 1.0)


def test_two_points_closed_form(ctx):
    d = tess.Diagram(ctx, np.array([[0.25, 0.5, 0.5], [0.75, 0.5, 0.5]]), UNIT_LO, UNIT_HI)
    d.compute()
    c = d.cells()
    assert np.array_equal(c["volume"], [0.5, 0.5])
    assert list(c["face_count"]) == [6, 6]
    assert list(c["neighbor"][:6]) == [-6, -5, -4, -3, -1, 1]
    assert list(c["neighbor"][6:]) == [-6, -5, -4, -3, -2, 0]


@pytest.mark.parametrize("m", [4, 6])
def test_cubic_lattice(ctx, m):
    ax = (np.arange(m) + 0.5) / m
    pts = np.array([[a, b, c] for a in ax for b in ax for c in ax])
    d = tess.Diagram(ctx, pts, UNIT_LO, UNIT_HI)
    d.compute()
    got = d.cells()
    ref = orc.compute(pts, UNIT_LO, UNIT_HI)
    _compare(got, ref)
    assert np.allclose(got["volume"], 1.0 / m**3, rtol=1e-12, atol=0)
    assert np.all(got["face_count"] == 6)


@pytest.mark.parametrize("n,seed,density", [(100, 3, 1.25), (1000, 4, 1.25), (10_000, 0, 1.25), (20_000, 7, 5.0)])
def test_uniform_cells_match_oracle(ctx, n, seed, density):
    pts = tess.uniform(n, seed)
    d = tess.Diagram(ctx, pts, UNIT_LO, UNIT_HI, density=density, collect_stats=True)
    d.compute()
    got = d.cells()
    ref = orc.compute(pts, UNIT_LO, UNIT_HI, density=density)
    _compare(got, ref)
    assert abs(got["volume"].sum() - 1.0) <= 1e-10
    st, rs = d.stats(), ref["stats"]
    for k in ("streamed", "classified", "cuts", "entries", "flops", "incident_cuts"):
        assert np.array_equal(st[k].astype(np.uint64), rs[k].astype(np.uint64)), k


def test_thin_slab_box_matches_oracle(ctx):
    """A box 250 times longer than it is thin ([0, 50] x [0, 1] x [0, 0.2]):
    grid cells of 3.3 x 0.067 x 0.013, so the shell order's keys weigh the
    three axes very differently and the security balls span many grid cells
    along the thin axis.  Bitwise against the oracle, counters included, on
    the single-set path and in a batch beside a unit-cube set."""
    lo, hi = np.zeros(3), np.array([50.0, 1.0, 0.2])
    pts = tess.uniform(20_000, 41, lo, hi)
    d = tess.Diagram(ctx, pts, lo, hi, density=5.0, collect_stats=True)
    d.compute()
    got = d.cells()
    ref = orc.compute(pts, lo, hi, density=5.0)
    for k in ("face_count", "face_offset", "neighbor", "volume", "area"):
        assert np.array_equal(got[k], ref[k]), k
    assert abs(got["volume"].sum() - 10.0) <= 1e-10 * 10.0
    st, rs = d.stats(), ref["stats"]
    for k in ("streamed", "classified", "cuts", "entries", "flops", "incident_cuts"):
        assert np.array_equal(st[k].astype(np.uint64), rs[k].astype(np.uint64)), k
    b = tess.Batch(ctx, [(tess.uniform(5_000, 42), UNIT_LO, UNIT_HI), (pts, lo, hi)], density=5.0)
    b.compute()
    got_b = b.cells(1)
    for k in got:
        assert np.array_equal(got_b[k], got[k]), k
    b.close()
    d.close()


def test_incident_only_cuts_counted(ctx):
    """Cuts bordered by Incident vertices only (the reference's
    find_outgoing_edge skips them, polyhedron.rs:418-432): counted per cell
    exactly as the oracle counts them, and the cells exact."""
    g = (np.arange(4) + 0.5) / 4
    pts = np.array([(x, y, z) for x in g for y in g for z in g])
    d = tess.Diagram(ctx, pts, UNIT_LO, UNIT_HI, collect_stats=True)
    d.compute()
    ref = orc.compute(pts, UNIT_LO, UNIT_HI)
    st = d.stats()
    assert np.array_equal(st["incident_cuts"].astype(np.int64), ref["stats"]["incident_cuts"].astype(np.int64))
    assert st["incident_cuts"].sum() > 0
    assert d.timings()["incident_cuts"] == st["incident_cuts"].sum()
    assert np.abs(d.cells()["volume"] - 1 / 64).max() < 1e-15
