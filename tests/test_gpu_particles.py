"""Cells of sites that are not points of the diagram on the GPU
(Diagram::get_cell_at_particle + Cell::compute_voronoi_cell,
interface.rs:211-313), against the oracle's restatement
(orc_compute_at_points): neighbour lists, face counts and volumes bitwise,
counters exact; sites outside the points' bounding box, on a point of the
set, with a search radius and with a target group."""
import numpy as np
import pytest

import tess
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

LO, HI = np.zeros(3), np.ones(3)


def _same(got, want, stats=True):
    for k in ("volume", "face_count", "face_offset", "neighbor", "area"):
        assert np.array_equal(got[k], want[k]), k
    if stats:
        for k in ("streamed", "classified", "cuts", "entries", "flops", "incident_cuts"):
            assert np.array_equal(got["stats"][k], want["stats"][k]), k


def _sites(pts, seed, m):
    rng = np.random.default_rng(seed)
    s = rng.uniform(0, 1, (m, 3))
    s[0], s[1] = 0.0, 1.0                  # box corners: outside the points' box, clamped grid cells
    s[2] = pts[7]                          # on a point of the set (its bisector is a no-op)
    s[3:13] = pts[rng.integers(0, len(pts), 10)] + rng.normal(0, 1e-3, (10, 3)).clip(-0.5, 0.5)
    return s.clip(0, 1)


@pytest.mark.parametrize("m", [3000, 20000])  # input order; grid order (4096 sites and more)
@pytest.mark.parametrize("workload", ["uniform", "clustered"])
def test_sites_match_oracle(ctx, workload, m):
    pts = tess.uniform(30000, 61) if workload == "uniform" else tess.clustered(30000, 62)
    sites = _sites(pts, 63, m)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0, collect_stats=True)
    got = d.cells_at_particles(sites, stats=True)
    want = orc.compute_at_points(pts, sites, LO, HI, density=5.0)
    _same(got, want)
    assert abs(got["volume"].sum() - want["volume"].sum()) < 1e-9


def test_site_on_removed_point_is_that_points_cell(ctx):
    """The cell of a site at point i among the other points is point i's cell."""
    pts = tess.uniform(5000, 64)
    idx = np.arange(0, 5000, 250)
    keep = np.setdiff1d(np.arange(5000), idx)
    full = tess.Diagram(ctx, pts, LO, HI, density=5.0)
    full.compute()
    ref = full.cells()
    d = tess.Diagram(ctx, pts[keep], LO, HI, density=5.0)
    got = d.cells_at_particles(pts[idx])
    checked = 0
    for a, i in enumerate(idx):
        nb_ref = ref["neighbor"][ref["face_offset"][i]:ref["face_offset"][i + 1]]
        if np.isin(nb_ref, idx).any():
            continue  # another removed point was a neighbour: a different cell
        nb = got["neighbor"][got["face_offset"][a]:got["face_offset"][a + 1]]
        mapped = np.where(nb >= 0, keep[np.maximum(nb, 0)], nb)
        assert sorted(mapped.tolist()) == sorted(nb_ref.tolist())
        assert abs(got["volume"][a] - ref["volume"][i]) <= 1e-12 * ref["volume"][i]
        checked += 1
    assert checked >= len(idx) // 2


@pytest.mark.parametrize("m", [500, 5000])  # input order; grid order
def test_sites_with_radius_and_group(ctx, m):
    pts = tess.clustered(20000, 65)
    groups = np.random.default_rng(66).integers(0, 3, len(pts)).astype(np.int64)
    sites = _sites(pts, 67, m)
    d = tess.Diagram(ctx, pts, LO, HI, groups=groups, density=5.0, collect_stats=True)
    for r, g in ((0.002, None), (None, 2), (0.01, 0)):
        got = d.cells_at_particles(sites, search_radius=r, target_group=g, stats=True)
        want = orc.compute_at_points(pts, sites, LO, HI, density=5.0, max_radius=r, groups=groups, target_group=g)
        _same(got, want)


def test_site_errors(ctx):
    pts = tess.uniform(1000, 68)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0)
    with pytest.raises(tess.TessError):
        d.cells_at_particles([[0.5, 0.5, 1.5]])      # outside the container
    with pytest.raises(tess.TessError):
        d.cells_at_particles([[0.5, np.nan, 0.5]])
    with pytest.raises(tess.TessError):
        d.cells_at_particles([[0.5, 0.5, 0.5]], target_group=1)  # no groups
    with pytest.raises(tess.TessError):
        d.cells_at_particles(np.zeros((0, 3)))
    d2 = tess.Diagram(ctx, pts, LO, HI, density=5.0)
    with pytest.raises(tess.TessError):
        d2.cells_at_particles([[0.5, 0.5, 0.5]], stats=True)  # created without collect_stats


def test_cell_at_particle_single(ctx):
    """tess_cell_at_particle: the Cell API for one outside site equals the
    batched result (and the oracle's), and leaves the diagram's cells alone."""
    pts = tess.clustered(20000, 69)
    sites = _sites(pts, 70, 40)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0)
    d.compute()
    before = d.cells()
    batch = d.cells_at_particles(sites)
    want = orc.compute_at_points(pts, sites, LO, HI, density=5.0)
    fo = batch["face_offset"]
    for i in range(len(sites)):
        c = d.get_cell_at_particle(sites[i])
        assert c.original_index() == -1
        assert c.compute_volume() == batch["volume"][i] == want["volume"][i]
        nb = c.compute_neighbors()
        assert np.array_equal(nb, batch["neighbor"][fo[i]:fo[i + 1]])
        for f in range(c.face_count()):
            area, _ = c.face(f)
            ring = c.face_vertices(f)
            w = np.zeros(3)
            for k in range(1, len(ring) - 1):
                w += np.cross(ring[k] - ring[0], ring[k + 1] - ring[0])
            assert abs(0.5 * np.linalg.norm(w) - area) <= 1e-12 * max(area, 1e-12)
    after = d.cells()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    with pytest.raises(tess.TessError):
        d.get_cell_at_particle([2.0, 0.5, 0.5])
