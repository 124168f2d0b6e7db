"""Cells of sites that are not points of the set in the oracle
(orc_compute_at_points: Diagram::get_cell_at_particle, interface.rs:211-232).

Pinned by: the cell of a site at point i among the other points is point i's
own cell (same neighbours; volume to rounding) whenever no other removed
point neighboured it; the security-radius stop is bit-identical to streaming
every point (expand_all_no_radius) also for sites outside the points'
bounding box, whose grid cell is clamped; a site on a point of the set
leaves that point's bisector a no-op.
"""
import numpy as np

import tess
from oracle import oracle as orc

LO, HI = np.zeros(3), np.ones(3)


def test_site_at_removed_point_is_its_cell():
    pts = tess.uniform(3000, 11)
    idx = np.arange(0, 3000, 150)
    keep = np.setdiff1d(np.arange(3000), idx)
    full = orc.compute(pts, LO, HI, cells=idx)
    got = orc.compute_at_points(pts[keep], pts[idx], LO, HI)
    checked = 0
    for a in range(len(idx)):
        nb_ref = full["neighbor"][full["face_offset"][a]:full["face_offset"][a + 1]]
        if np.isin(nb_ref, idx).any():
            continue
        nb = got["neighbor"][got["face_offset"][a]:got["face_offset"][a + 1]]
        mapped = np.where(nb >= 0, keep[np.maximum(nb, 0)], nb)
        assert sorted(mapped.tolist()) == sorted(nb_ref.tolist())
        assert abs(got["volume"][a] - full["volume"][a]) <= 1e-12 * full["volume"][a]
        checked += 1
    assert checked >= len(idx) // 2


def test_security_stop_equals_all_points_for_outside_sites():
    pts = tess.clustered(4000, 12) * 0.8 + 0.1  # points' box well inside the container
    rng = np.random.default_rng(13)
    sites = np.vstack([rng.uniform(0, 1, (200, 3)), [[0, 0, 0], [1, 1, 1], [0.05, 0.5, 0.95]], pts[:3]])
    a = orc.compute_at_points(pts, sites, LO, HI, mode=orc.MODE_SECURITY)
    b = orc.compute_at_points(pts, sites, LO, HI, mode=orc.MODE_ALL_POINTS)
    for k in ("volume", "face_count", "face_offset", "neighbor", "area"):
        assert np.array_equal(a[k], b[k]), k
    assert np.all(a["stats"]["status"] == 0)


def test_site_on_a_point_and_duplicate_free_cell():
    """A site on point j: the bisector with j is degenerate (a no-op), so the
    cell is j's own cell computed against the other points."""
    pts = tess.uniform(2000, 14)
    j = 321
    on = orc.compute_at_points(pts, pts[j:j + 1], LO, HI)
    own = orc.compute(pts, LO, HI, cells=[j])
    for k in ("volume", "face_count", "neighbor", "area"):
        assert np.array_equal(on[k], own[k]), k
