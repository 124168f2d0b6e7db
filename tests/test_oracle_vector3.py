"""Vector and plane arithmetic of the oracle against the reference's
known-answer tests (vector3.rs:315-476) and closed forms."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc

KA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))["vector3"]
LOC = {"Outside": orc.OUTSIDE, "Incident": orc.INCIDENT, "Inside": orc.INSIDE}


def test_dot():
    k = KA["dot"]
    assert orc.dot(k["a"], k["b"]) == k["expect"]
    assert orc.dot(k["b"], k["a"]) == k["expect"]


def test_cross_closed_form():
    assert np.array_equal(orc.cross([1, 0, 0], [0, 1, 0]), [0, 0, 1])
    assert np.array_equal(orc.cross([0, 1, 0], [1, 0, 0]), [0, 0, -1])
    a, b = np.array([1.5, -2.0, 0.25]), np.array([3.0, 0.5, -4.0])
    assert np.array_equal(orc.cross(a, b), np.cross(a, b))


def test_location():
    k = KA["location"]
    for sd, want in k["cases"]:
        assert orc.location(sd, k["tol"]) == LOC[want]


def test_vector_location():
    k = KA["vector_location"]
    for v, want in k["cases"]:
        assert orc.vector_location(k["plane"], v, k["tol"]) == LOC[want]


def test_intersection_exact():
    k = KA["intersection"]
    assert np.array_equal(orc.intersection(k["plane"], k["a"], k["b"]), k["expect"])


@pytest.mark.parametrize("i", range(3))
def test_halfway_from_origin_to(i):
    p, half = KA["halfway_from_origin_to"]["cases"][i]
    got = orc.halfway_from_origin_to(p)
    want = orc.build_from_non_unit_normal_and_point(half, half)
    assert np.array_equal(got, want)  # bitwise, as assert_eq! on Float64


def test_halfway_plane_is_the_bisector():
    rng = np.random.default_rng(1)
    for p in rng.normal(size=(200, 3)):
        pl = orc.halfway_from_origin_to(p)
        assert abs(np.linalg.norm(pl[:3]) - 1) < 1e-15
        # the midpoint is on the plane, the origin inside, p outside
        assert orc.vector_location(pl, p / 2, 1e-12) == orc.INCIDENT
        assert orc.vector_location(pl, [0, 0, 0], 1e-12) == orc.INSIDE
        assert orc.vector_location(pl, p, 1e-12) == orc.OUTSIDE
