"""GPU path through the C ABI: fixtures, the single-cell query API, the
reference's query options, clustered / jittered-lattice parity, capacity
reruns, degenerate inputs and full-size invariants."""
import glob
import os

import numpy as np
import pytest

import tess
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
LO, HI = np.zeros(3), np.ones(3)
FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "uniform_*.npz")))


def same_cells(got, ref, rel=1e-12):
    assert np.array_equal(got["face_count"], ref["face_count"])
    assert np.array_equal(got["neighbor"], ref["neighbor"])
    for k in ("volume", "area"):
        r = np.abs(got[k] - ref[k]) / np.maximum(np.abs(ref[k]), 1e-300)
        assert r.max(initial=0) <= rel, (k, r.max())


def gpu_cells(ctx, pts, lo=LO, hi=HI, **kw):
    d = tess.Diagram(ctx, pts, lo, hi, **kw)
    d.compute()
    return d, d.cells()


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p) for p in FIX])
def test_golden_fixtures(ctx, path):
    f = np.load(path, allow_pickle=False)
    _, got = gpu_cells(ctx, f["points"], density=float(f["density"]))
    same_cells(got, {k: f[k] for k in ("face_count", "neighbor", "volume", "area")}, rel=0.0)


def test_matches_streaming_every_point(ctx):
    """The security-radius stop against the reference's expand_all_no_radius."""
    pts = tess.uniform(400, 77)
    _, got = gpu_cells(ctx, pts)
    ref = orc.compute(pts, LO, HI, mode=orc.MODE_ALL_POINTS)
    same_cells(got, ref, rel=0.0)


def test_cell_at_index(ctx):
    pts = tess.uniform(3000, 5)
    d, allc = gpu_cells(ctx, pts)
    fo = allc["face_offset"]
    for i in [0, 1, 17, 999, 2999]:
        c = d.get_cell_at_index(i)
        assert c.original_index() == i
        assert c.compute_volume() == allc["volume"][i]
        nb = c.compute_neighbors()
        assert np.array_equal(nb, allc["neighbor"][fo[i]:fo[i + 1]])
        assert c.face_count() == len(nb)
        v = c.compute_vertices()
        assert len(v) >= 4 and np.all(np.abs(v) < 1)  # site-relative
        for f in range(c.face_count()):
            area, nbr = c.face(f)
            assert nbr == nb[f] and area == allc["area"][fo[i] + f]
            ring = c.face_vertices(f)
            w = np.zeros(3)
            for k in range(1, len(ring) - 1):
                w += np.cross(ring[k] - ring[0], ring[k + 1] - ring[0])
            assert abs(0.5 * np.linalg.norm(w) - area) <= 1e-12 * max(area, 1e-9)


def test_cell_query_options(ctx):
    pts = tess.uniform(2000, 6)
    groups = np.arange(2000) % 3
    d = tess.Diagram(ctx, pts, LO, HI, groups=groups)
    for i in [3, 100, 1500]:
        c = d.get_cell_at_index(i, target_group=1)
        ref = orc.compute(pts, LO, HI, cells=[i], groups=groups, target_group=1)
        assert np.array_equal(c.compute_neighbors(), ref["neighbor"])
        assert c.compute_volume() == ref["volume"][0]
        c = d.get_cell_at_index(i, search_radius=0.004)
        ref = orc.compute(pts, LO, HI, cells=[i], max_radius=0.004)
        assert np.array_equal(c.compute_neighbors(), ref["neighbor"])
        assert c.compute_volume() == ref["volume"][0]


def test_clustered_parity(ctx):
    pts = tess.clustered(100_000, 25)
    _, got = gpu_cells(ctx, pts, density=5.0)
    ref = orc.compute(pts, LO, HI, density=5.0)
    same_cells(got, ref)
    assert abs(got["volume"].sum() - 1) < 1e-10


def test_capacity_reruns_change_nothing(ctx):
    pts = tess.clustered(50_000, 26)
    d64, a = gpu_cells(ctx, pts, density=5.0, vertex_capacity=64)
    d256, b = gpu_cells(ctx, pts, density=5.0, vertex_capacity=256)
    same_cells(a, b, rel=0.0)


def test_fcc_jitter_parity(ctx):
    pts = tess.fcc(4000, ncell=12, a=1.0, sigma=1e-3)
    lo, hi = np.zeros(3), np.full(3, 12.0)
    _, got = gpu_cells(ctx, pts, lo, hi, density=5.0)
    ref = orc.compute(pts, lo, hi, density=5.0)
    same_cells(got, ref)
    assert abs(got["volume"].sum() - 12.0 ** 3) < 1e-10 * 12.0 ** 3


def test_duplicate_points(ctx):
    pts = tess.uniform(200, 8)
    pts[10] = pts[11]  # the reference's plane of a zero vector is NaN: no cut
    _, got = gpu_cells(ctx, pts)
    ref = orc.compute(pts, LO, HI)
    same_cells(got, ref, rel=0.0)


def test_points_on_box_walls(ctx):
    pts = tess.uniform(500, 9)
    pts[:20, 0] = 0.0
    pts[20:40, 1] = 1.0
    _, got = gpu_cells(ctx, pts)
    ref = orc.compute(pts, LO, HI)
    same_cells(got, ref, rel=0.0)


def test_invalid_inputs(ctx):
    with pytest.raises(tess.TessError):
        tess.Diagram(ctx, np.zeros((0, 3)), LO, HI)
    with pytest.raises(tess.TessError):
        tess.Diagram(ctx, np.zeros((3, 3)), HI, LO)


def test_full_size_uniform_1m(ctx):
    """Config of the benchmark (1M uniform, about 5 points per grid cell):
    size-independent properties on every cell, oracle parity on a sample."""
    n = 1_000_000
    pts = tess.uniform(n, 1)
    _, got = gpu_cells(ctx, pts, density=5.0)
    assert abs(got["volume"].sum() - 1) < 1e-10
    fo = got["face_offset"].astype(np.int64)
    owner = np.repeat(np.arange(n), np.diff(fo))
    nb, ar = got["neighbor"], got["area"]
    inner = nb >= 0
    a = np.stack([owner[inner], nb[inner]], 1)
    key_ab = a[:, 0].astype(np.int64) * n + a[:, 1]
    key_ba = a[:, 1].astype(np.int64) * n + a[:, 0]
    o1, o2 = np.argsort(key_ab), np.argsort(key_ba)
    assert np.array_equal(key_ab[o1], key_ba[o2])  # neighbour symmetry
    # shared faces: both cells' areas agree to rounding, measured against the
    # typical face area (tiny faces carry absolute, not relative, rounding)
    scale = np.maximum(ar[inner][o1], np.median(ar[inner]))
    assert (np.abs(ar[inner][o1] - ar[inner][o2]) / scale).max() < 1e-9
    sample = np.sort(np.random.default_rng(0).choice(n, 3000, replace=False))
    ref = orc.compute(pts, LO, HI, density=5.0, cells=sample)
    sub = {"face_count": got["face_count"][sample], "volume": got["volume"][sample]}
    sub["neighbor"] = np.concatenate([nb[fo[i]:fo[i + 1]] for i in sample])
    sub["area"] = np.concatenate([ar[fo[i]:fo[i + 1]] for i in sample])
    same_cells(sub, ref, rel=0.0)



def test_large_downloads_staged_and_concurrent(ctx):
    """Results above the staging threshold (16 MiB) come down through the
    context's pinned double buffer, chunk by chunk: every shared face has the
    same area and neighbour pair from both sides and the volumes sum to the
    box, and two host threads downloading from one context at the same time
    (the staging is taken in turn) get the same arrays."""
    import threading
    n = 2_000_000
    pts = tess.uniform(n, 9)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0)
    d.compute()
    first = d.cells()
    assert first["area"].nbytes > (128 << 20) and first["neighbor"].nbytes > (16 << 20)
    assert abs(first["volume"].sum() - 1) < 1e-10
    fo = first["face_offset"].astype(np.int64)
    owner = np.repeat(np.arange(n, dtype=np.int64), np.diff(fo))
    nb, ar = first["neighbor"].astype(np.int64), first["area"]
    inner = nb >= 0
    k1, k2 = owner[inner] * n + nb[inner], nb[inner] * n + owner[inner]
    o1, o2 = np.argsort(k1), np.argsort(k2)
    assert np.array_equal(k1[o1], k2[o2])
    scale = np.maximum(ar[inner][o1], np.median(ar[inner]))
    assert (np.abs(ar[inner][o1] - ar[inner][o2]) / scale).max() < 1e-9
    out = [None, None]

    def work(k):
        out[k] = d.cells()

    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for k in range(2):
        for key in first:
            assert np.array_equal(out[k][key], first[key]), (k, key)
    d.close()
