"""Morton order and index maps of the oracle (Diagram::morton_index and
sort_particles, interface.rs:52-183), pinned independently of the C
restatement: a bit-interleaving written here with numpy, the reference's
bounding box (grown from the origin), the quirks of its `as usize` cast, and
the invariants of the internal order."""
import numpy as np
import pytest

import tess
from oracle import oracle as orc


def _interleave(ix, iy, iz):
    m = np.zeros(len(ix), np.uint64)
    for level in range(10):
        for a, v in enumerate((ix, iy, iz)):
            m |= ((v >> level) & 1).astype(np.uint64) << np.uint64(3 * level + a)
    return m.astype(np.uint32)


def test_codes_match_an_independent_interleave():
    pts = tess.uniform(5000, 31, [-1.0, 0.5, 2.0], [3.0, 1.5, 7.0])
    lo = np.minimum(pts.min(0), 0.0)  # the default box sits at the origin
    hi = np.maximum(pts.max(0), 0.0)
    side = (hi - lo) * (1.0 / 1024.0)
    idx = np.floor((pts - lo) / side).astype(np.uint64)
    want = _interleave(idx[:, 0], idx[:, 1], idx[:, 2])
    assert np.array_equal(orc.morton_codes(pts), want)


def test_code_quirks():
    """The largest coordinate maps to 1024, whose low 10 bits are zero; the
    box always contains the origin."""
    pts = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.5, 0.25, 0.75], [1.0, 0.0, 0.5]])
    c = orc.morton_codes(pts)
    assert c[0] == 0 and c[1] == 0
    assert c[2] == _interleave(np.array([512]), np.array([256]), np.array([768]))[0]
    assert c[3] == _interleave(np.array([0]), np.array([0]), np.array([512]))[0]
    # shifting the set away from the origin changes the box and the codes
    assert not np.array_equal(orc.morton_codes(pts + 5.0), c)


def test_index_maps_are_the_stable_code_order():
    pts = tess.clustered(4000, 32)
    pts[100] = pts[200]  # a tie: equal codes keep input order
    ooi, ioo = orc.index_maps(pts)
    codes = orc.morton_codes(pts)
    assert np.array_equal(ooi, np.argsort(codes, kind="stable"))
    assert np.array_equal(ioo[ooi], np.arange(len(pts)))


@pytest.mark.parametrize("density", [1.25, 5.0])
def test_grid_in_internal_order(density):
    pts = tess.uniform(3000, 33)
    g0, b0, s0, d0 = orc.grid(pts, density)
    g1, b1, s1, d1 = orc.grid(pts, density, order=orc.ORDER_MORTON)
    assert np.array_equal(b0, b1) and np.array_equal(d0, d1)
    _, ioo = orc.index_maps(pts)
    for c in range(len(d1) - 1):
        seg = s1[d1[c]:d1[c + 1]]
        assert sorted(seg.tolist()) == s0[d0[c]:d0[c + 1]].tolist()
        assert np.all(np.diff(ioo[seg].astype(np.int64)) > 0)


def test_cells_in_internal_order_are_the_same_cells():
    """The candidate order changes the rounding only: the same neighbours,
    volumes to 1e-12, and the all-points check still bit-identical."""
    pts = tess.uniform(2000, 34)
    lo, hi = np.zeros(3), np.ones(3)
    a = orc.compute(pts, lo, hi)
    b = orc.compute(pts, lo, hi, order=orc.ORDER_MORTON)
    c = orc.compute(pts, lo, hi, order=orc.ORDER_MORTON, mode=orc.MODE_ALL_POINTS)
    assert np.array_equal(a["neighbor"], b["neighbor"])
    assert np.allclose(a["volume"], b["volume"], rtol=1e-12, atol=0)
    for k in ("volume", "neighbor", "area"):
        assert np.array_equal(b[k], c[k]), k


def test_geometry_in_internal_order():
    """Rings in the internal order: the cells of compute() in that order
    (bitwise), the same rings as in input order up to rounding, and fan areas
    equal to the face areas."""
    pts = tess.uniform(2000, 35)
    lo, hi = np.zeros(3), np.ones(3)
    g = orc.compute_geometry(pts, lo, hi, density=5.0, order=orc.ORDER_MORTON)
    c = orc.compute(pts, lo, hi, density=5.0, order=orc.ORDER_MORTON)
    for k in ("volume", "neighbor", "area"):
        assert np.array_equal(g[k], c[k]), k
    g0 = orc.compute_geometry(pts, lo, hi, density=5.0)
    assert np.array_equal(g["ring_offset"], g0["ring_offset"])
    assert np.allclose(g["ring_xyz"], g0["ring_xyz"], rtol=0, atol=1e-12)
    assert not np.array_equal(g["ring_xyz"], g0["ring_xyz"])  # the order did reach the clip
    ro = g["ring_offset"].astype(np.int64)
    for f in range(0, len(ro) - 1, 53):
        ring = g["ring_xyz"][ro[f]:ro[f + 1]]
        w = sum(np.cross(ring[k] - ring[0], ring[k + 1] - ring[0]) for k in range(1, len(ring) - 1))
        assert abs(0.5 * np.linalg.norm(w) - g["area"][f]) <= 1e-12 * max(g["area"][f], 1e-12)


def test_sites_in_internal_order():
    """Cells of outside sites over a grid in the internal order: the same
    neighbours as in input order, volumes to 1e-12, and the order reaches the
    clip (some volume differs in its last bits)."""
    pts = tess.uniform(3000, 36)
    sites = np.random.default_rng(37).uniform(0, 1, (500, 3))
    lo, hi = np.zeros(3), np.ones(3)
    a = orc.compute_at_points(pts, sites, lo, hi, density=5.0)
    b = orc.compute_at_points(pts, sites, lo, hi, density=5.0, order=orc.ORDER_MORTON)
    assert np.array_equal(a["neighbor"], b["neighbor"])
    assert np.allclose(a["volume"], b["volume"], rtol=1e-12, atol=0)
    assert not np.array_equal(a["volume"], b["volume"])
