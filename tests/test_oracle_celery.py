"""Grid (Celery) of the oracle against the reference's tests
(celery.rs:1174-1487) and closed forms of the search order."""
import json
import os

import numpy as np
import pytest

import tess
from oracle import oracle as orc

KA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))


def _points(n, box, seed):
    lo = [b[0] for b in box]
    hi = [b[1] for b in box]
    return tess.uniform(n, seed, lo, hi)


def check_delimiters(bins, srt, dl):
    """celery.rs:1131-1172 plus the CSR identity delimiters == scan(histogram)."""
    n = len(srt)
    assert dl[0] == 0 and dl[-1] == n
    assert len(set(srt.tolist())) == n
    hist = np.bincount(bins, minlength=len(dl) - 1)
    assert np.array_equal(dl[1:], np.cumsum(hist))
    for c in range(len(dl) - 1):
        seg = srt[dl[c]:dl[c + 1]]
        assert np.all(bins[seg] == c)
        assert np.all(np.diff(seg.astype(np.int64)) > 0)  # canonical order inside a cell


@pytest.mark.parametrize("case", KA["celery"]["grid_sizes"]["cases"][:7])
def test_grid_sizes(case):
    pts = _points(case["n"], case["box"], 11)
    g, bins, srt, dl = orc.grid(pts)
    assert g.cpd == case["cpd"]
    assert len(dl) == case["delimiters"]
    assert (2 * g.cpd - 1) ** 3 == case["search_order"]
    key, off, gen, kc = orc.search_order(g.cpd, g.sx, g.sy, g.sz)
    assert len(key) == case["search_order"]
    check_delimiters(bins, srt, dl)


def test_grid_size_one_million_and_config0():
    c = KA["celery"]["grid_sizes"]["cases"][7]
    assert orc.cells_per_dimension(c["n"]) == c["cpd"]
    assert c["cpd"] ** 3 + 1 == c["delimiters"] and (2 * c["cpd"] - 1) ** 3 == c["search_order"]
    c0 = KA["config0_grid"]
    assert orc.cells_per_dimension(c0["n"], c0["density"]) == c0["cpd"]  # exact cube root edge


def test_axis_index_rules():
    # x >= max -> cpd-1 (celery.rs:276-278); saturating cast below min -> 0
    assert orc.axis_index(1.0, 0.0, 1.0, 4.0, 4) == 3
    assert orc.axis_index(2.0, 0.0, 1.0, 4.0, 4) == 3
    assert orc.axis_index(-0.3, 0.0, 1.0, 4.0, 4) == 0
    assert orc.axis_index(0.2499999, 0.0, 1.0, 4.0, 4) == 0
    assert orc.axis_index(0.25, 0.0, 1.0, 4.0, 4) == 1


def _reference_generation_order(M):
    """Offsets in the order the loops of celery.rs:437-673 push them."""
    out = [(0, 0, 0)]
    for i in range(M):
        for j in range(M):
            for k in range(M):
                a, b, c = i + 1, j + 1, k + 1
                out += [(a, b, c), (a, b, -c), (a, -b, c), (-a, b, c), (a, -b, -c), (-a, b, -c), (-a, -b, c),
                        (-a, -b, -c)]
    for i in range(M):
        for j in range(M):
            out += [(i + 1, j + 1, 0), (i + 1, -j - 1, 0), (-i - 1, j + 1, 0), (-i - 1, -j - 1, 0)]
    for i in range(M):
        for k in range(M):
            out += [(i + 1, 0, k + 1), (i + 1, 0, -k - 1), (-i - 1, 0, k + 1), (-i - 1, 0, -k - 1)]
    for j in range(M):
        for k in range(M):
            out += [(0, j + 1, k + 1), (0, j + 1, -k - 1), (0, -j - 1, k + 1), (0, -j - 1, -k - 1)]
    for i in range(M):
        out += [(i + 1, 0, 0), (-i - 1, 0, 0)]
    for j in range(M):
        out += [(0, j + 1, 0), (0, -j - 1, 0)]
    for k in range(M):
        out += [(0, 0, k + 1), (0, 0, -k - 1)]
    return out


@pytest.mark.parametrize("cpd", [1, 2, 5, 7])
def test_generation_index_matches_reference_loops(cpd):
    order = _reference_generation_order(cpd - 1)
    assert len(order) == (2 * cpd - 1) ** 3
    for idx, o in enumerate(order):
        assert orc.generation_index(cpd, *o) == idx


@pytest.mark.parametrize("sizes", [(0.25, 0.25, 0.25), (0.1, 0.37, 0.05), (1.0 / 3, 0.2, 0.7)])
def test_search_order_is_sorted_stable_and_complete(sizes):
    cpd = 6
    key, off, gen, kc = orc.search_order(cpd, *sizes)
    assert kc == float("inf")
    assert key[0] == -1 and tuple(off[0]) == (0, 0, 0)
    # a permutation of all offsets, sorted by key, ties by generation order
    assert len({tuple(o) for o in off}) == (2 * cpd - 1) ** 3
    assert np.all(np.diff(key) >= 0)
    tie = np.diff(key) == 0
    assert np.all(np.diff(gen.astype(np.int64))[tie] > 0)
    # key = squared gap distance between cells (celery.rs:419-427)
    gap = np.maximum(np.abs(off) - 1, 0) * np.array(sizes)
    assert np.allclose(key[1:], (gap[1:] ** 2).sum(1), rtol=1e-15, atol=0)
    # the 26 neighbours share key 0 and come right after the centre
    assert np.all(key[1:27] == 0) and key[27] > 0


def test_truncated_search_order_is_a_prefix():
    cpd, s = 20, (0.05, 0.07, 0.04)
    full = orc.search_order(cpd, *s)
    for L in (3, 6, 9):
        key, off, gen, kc = orc.search_order(cpd, *s, L=L)
        m = len(key)
        assert np.array_equal(key, full[0][:m]) and np.array_equal(off, full[1][:m])
        assert full[0][m] >= kc > key[-1]


def test_compare_distance_index_ties():
    # equal keys order by generation (the canonical choice for sort_unstable's ties)
    key, off, gen, _ = orc.search_order(4, 0.5, 0.5, 0.5)
    groups = {}
    for k, g in zip(key, gen):
        groups.setdefault(k, []).append(int(g))
    for g in groups.values():
        assert g == sorted(g)


@pytest.mark.parametrize("site,cell", [((0.5, 0.5, 0.5), (2, 2, 2)), ((0.0, 0.0, 0.0), (0, 0, 0))])
def test_expanding_search(site, cell):
    k = KA["celery"]["expanding_search"]
    pts = _points(k["n"], k["box"], 5)
    g, bins, srt, dl = orc.grid(pts)
    assert g.cpd == 5 and len(dl) == 126
    got = tuple(orc.axis_index(v, lo, hi, inv, g.cpd) for v, lo, hi, inv in
                zip(site, (g.xmin, g.ymin, g.zmin), (g.xmax, g.ymax, g.zmax), (g.ix, g.iy, g.iz)))
    if site == (0.5, 0.5, 0.5):
        assert got == cell
    out = orc.expand_all(pts, site)
    assert len(out) == k["found"] and sorted(out.tolist()) == list(range(k["n"]))


def test_expand_all_in_radius():
    k = KA["celery"]["expand_all_in_radius"]
    pts = np.vstack([_points(k["n_random"], k["box"], 9), k["anchors"]])
    g, _, _, _ = orc.grid(pts)
    assert g.cpd == 4 and g.xmin == -2 and g.xmax == 2
    small = orc.expand_all(pts, k["site"], max_radius=k["r_small"])
    for i in k["small_contains"]:
        assert i in small
    for i in k["small_excludes"]:
        assert i not in small
    assert len(small) < 79
    large = orc.expand_all(pts, k["site"], max_radius=k["r_large"])
    assert len(large) == k["large_count"]
