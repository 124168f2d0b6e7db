"""Polyhedron (container cube + clipping) of the oracle against the
reference's tests (polyhedron.rs:951-1118) and closed forms."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc

KA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))["polyhedron"]
CLIPS = [orc.CLIP_WALK, orc.CLIP_FACES]


@pytest.mark.parametrize("name", ["new", "reset"])
def test_new_and_reset(name):
    k = KA[name]
    p = orc.Poly(k["lo"], k["hi"], orc.CLIP_WALK)
    nv, ne, nf = p.counts()
    assert (nv, ne, nf) == (k["vertices"], k["half_edges"], k["faces"])
    for v in p.vertices():
        for a in range(3):
            assert v[a] in (k["lo"][a], k["hi"][a])
    assert p.check()  # flip(flip(e)) == e, closed faces, Euler 2 (build_cube_test)


@pytest.mark.parametrize("clip", CLIPS)
def test_cube_faces_outward(clip):
    p = orc.Poly([-1, -2, -3], [1, 2, 3], clip)
    tag, area, wn, deg = p.faces()
    assert list(tag) == [-3, -2, -4, -1, -6, -5]  # creation order F R B L U D (polyhedron.rs:73-80)
    assert list(deg) == [4] * 6
    # outward weighted normals, 2 * area each
    want = {-1: [-1, 0, 0], -2: [1, 0, 0], -3: [0, -1, 0], -4: [0, 1, 0], -5: [0, 0, -1], -6: [0, 0, 1]}
    for t, a, w in zip(tag, area, wn):
        assert np.allclose(w / np.linalg.norm(w), want[t])
        assert a == np.linalg.norm(w) / 2
    assert p.volume() == 2 * 4 * 6


def test_find_outgoing_edge_no_cut():
    k = KA["find_outgoing_edge_no_cut"]
    p = orc.Poly(k["lo"], k["hi"], orc.CLIP_WALK)
    plane = orc.build_from_non_unit_normal_and_point(k["normal"], k["point"])
    assert p.outgoing(plane) is None


def test_find_outgoing_edge_with_cut():
    k = KA["find_outgoing_edge_with_cut"]
    p = orc.Poly(k["lo"], k["hi"], orc.CLIP_WALK)
    plane = orc.build_from_non_unit_normal_and_point(k["normal"], k["point"])
    src, dst = p.outgoing(plane)
    assert list(dst) == k["target"]
    assert list(src) in k["source_one_of"]


@pytest.mark.parametrize("clip", CLIPS)
def test_cut_with_plane_corner_closed_form(clip):
    """The reference's cut_with_plane scenario with its closed form: the
    bisector of (1,1,1) cuts the corner tetrahedron with legs 1.5."""
    k = KA["cut_with_plane"]
    p = orc.Poly(k["lo"], k["hi"], clip)
    assert p.cut(k["tag"], orc.halfway_from_origin_to(k["point"])) == 1
    assert p.check()
    nv, ne, nf = p.counts()
    assert (nv, ne // 2, nf) == (10, 15, 7)
    assert p.volume() == 8 - 1.5 ** 3 / 6  # 7.4375, exact in binary
    tag, area, wn, deg = p.faces()
    i = list(tag).index(k["tag"])
    assert deg[i] == 3
    assert math.isclose(area[i], math.sqrt(3) / 4 * 4.5, rel_tol=1e-14)


@pytest.mark.parametrize("clip", CLIPS)
def test_plane_through_edge_keeps_incident_vertices(clip):
    """Plane x + z/2 = 1 through the cube edge x = 1, z = 0 of [0,1]^3:
    the two edge vertices are Incident and stay, face x = 1 disappears."""
    p = orc.Poly([0, 0, 0], [1, 1, 1], clip)
    plane = orc.build_from_non_unit_normal_and_point([1.0, 0.0, 0.5], [1.0, 0.0, 0.0])
    assert p.cut(7, plane) == 1
    assert p.check()
    tag, area, wn, deg = p.faces()
    assert -2 not in tag  # +x wall fully cut away
    assert 7 in tag and deg[list(tag).index(7)] == 4
    assert math.isclose(p.volume(), 0.75, rel_tol=1e-14)


@pytest.mark.parametrize("clip", CLIPS)
def test_outside_region_bounded_by_incident_vertices(clip):
    """Outside vertices whose neighbours are all Incident.  The reference's
    outgoing-edge rule (polyhedron.rs:418-432) finds no Inside->Outside edge
    here and skips the cut; the per-face clip cuts it."""
    p = orc.Poly([0, 0, 0], [1, 1, 1], clip)
    # corner x=y=z=1 cut by the plane through (1,1,0),(1,0,1),(0,1,1)
    plane = orc.build_from_non_unit_normal_and_point([1.0, 1.0, 1.0], [1.0, 1.0, 0.0])
    r = p.cut(3, plane)
    if clip == orc.CLIP_WALK:
        assert r == 0
        assert p.volume() == 1.0
    else:
        assert r == 1 and p.check()
        assert math.isclose(p.volume(), 1 - 1 / 6, rel_tol=1e-14)


@pytest.mark.parametrize("clip", CLIPS)
def test_random_cuts_stay_consistent(clip):
    rng = np.random.default_rng(3)
    p = orc.Poly([-1, -1, -1], [1, 1, 1], clip)
    for t in range(60):
        q = rng.normal(size=3) * 1.2
        p.cut(t, orc.halfway_from_origin_to(q))
        assert p.check()
    # every vertex is inside every face plane it is not on
    assert 0 < p.volume() < 8
