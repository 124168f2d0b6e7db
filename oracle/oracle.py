"""ctypes wrapper of liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU restatement of the-tessellator's per-cell path (see oracle.cpp for the
file:line map).  Only tests, the smoke check and the benchmark's CPU-baseline
leg use it, as the checker; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

OUTSIDE, INCIDENT, INSIDE = 0, 1, 2
MODE_ALL_POINTS, MODE_SECURITY = 0, 1
CLIP_FACES, CLIP_WALK, CHECK = 0, 1, 2


class Grid(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax",
                                          "sx", "sy", "sz", "ix", "iy", "iz")] + [("cpd", C.c_int64)]


class CellStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("streamed", "classified", "cuts", "new_vertices", "entries", "flops")] + \
               [("r2", C.c_double), ("stop_key", C.c_double)] + \
               [(k, C.c_int32) for k in ("max_vertices", "max_faces", "max_edges", "status", "incident_cuts",
                                         "reserved")]


STATS_DTYPE = np.dtype([(k, "<u8") for k in ("streamed", "classified", "cuts", "new_vertices", "entries", "flops")] +
                       [("r2", "<f8"), ("stop_key", "<f8")] +
                       [(k, "<i4") for k in ("max_vertices", "max_faces", "max_edges", "status", "incident_cuts",
                                             "reserved")])
assert STATS_DTYPE.itemsize == C.sizeof(CellStats)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, dp, sz = C.c_void_p, C.c_void_p, C.c_size_t
        sig = {
            "orc_dot": (C.c_double, [dp, dp]),
            "orc_cross": (None, [dp, dp, dp]),
            "orc_halfway_from_origin_to": (None, [dp, dp]),
            "orc_build_from_non_unit_normal_and_point": (None, [dp, dp, dp]),
            "orc_location": (C.c_int, [C.c_double, C.c_double]),
            "orc_vector_location": (C.c_int, [dp, dp, C.c_double]),
            "orc_intersection": (None, [dp, dp, dp, dp]),
            "orc_cells_per_dimension": (C.c_int64, [sz, C.c_double]),
            "orc_grid_info": (None, [dp, dp, dp, sz, C.c_double, C.POINTER(Grid)]),
            "orc_axis_index": (C.c_int64, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64]),
            "orc_grid_build": (None, [dp, dp, dp, sz, C.POINTER(Grid), dp, dp, dp]),
            "orc_search_order": (sz, [C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int64, dp, dp, dp, dp]),
            "orc_generation_index": (C.c_uint64, [C.c_int64, C.c_int32, C.c_int32, C.c_int32]),
            "orc_expand_all": (sz, [dp, dp, dp, sz, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                    dp]),
            "orc_check_cell_in_range": (C.c_int, [C.POINTER(Grid)] + [C.c_double] * 4 + [C.c_int64] * 3),
            "orc_find_cells_in_radius": (sz, [C.POINTER(Grid)] + [C.c_double] * 4 + [dp]),
            "orc_find_neighbors": (sz, [dp, dp, dp, C.POINTER(Grid), dp, dp] + [C.c_double] * 4 + [C.c_int, dp]),
            "orc_queries": (None, [C.c_int, dp, dp, dp, sz, C.c_double, dp, C.c_int64, dp, sz, dp, dp, dp]),
            "orc_queries_ordered": (None, [C.c_int, dp, dp, dp, sz, C.c_double, dp, C.c_int64, dp, sz, dp, dp, dp,
                                           C.c_int]),
            "orc_morton_codes": (None, [dp, dp, dp, sz, dp]),
            "orc_index_maps": (None, [dp, dp, dp, sz, dp, dp]),
            "orc_grid_build_ordered": (None, [dp, dp, dp, sz, C.POINTER(Grid), C.c_int, dp, dp, dp]),
            "orc_compute_ordered": (vp, [dp, dp, dp, sz, dp, dp, C.c_double, C.c_int, C.c_int, dp, sz, C.c_int,
                                         C.c_double, dp, C.c_int64, C.c_int]),
            "orc_poly_new": (vp, [C.c_double] * 6),
            "orc_poly_new_faces": (vp, [C.c_double] * 6),
            "orc_poly_free": (None, [vp]),
            "orc_poly_cut": (C.c_int, [vp, C.c_int64, dp]),
            "orc_poly_counts": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "orc_poly_volume": (C.c_double, [vp]),
            "orc_poly_faces": (C.c_int, [vp, dp, dp, dp, dp, C.c_int]),
            "orc_poly_vertices": (C.c_int, [vp, dp, C.c_int]),
            "orc_poly_check": (C.c_int, [vp]),
            "orc_poly_outgoing": (C.c_int, [vp, dp, dp, dp]),
            "orc_compute": (vp, [dp, dp, dp, sz, dp, dp, C.c_double, C.c_int, C.c_int, dp, sz, C.c_int]),
            "orc_compute_query": (vp, [dp, dp, dp, sz, dp, dp, C.c_double, C.c_int, C.c_int, dp, sz, C.c_int,
                                       C.c_double, dp, C.c_int64]),
            "orc_compute_at_points": (vp, [dp, dp, dp, sz, dp, dp, C.c_double, C.c_int, C.c_int, dp, sz, C.c_int,
                                           C.c_double, dp, C.c_int64]),
            "orc_compute_at_points_ordered": (vp, [dp, dp, dp, sz, dp, dp, C.c_double, C.c_int, C.c_int, dp, sz,
                                                   C.c_int, C.c_double, dp, C.c_int64, C.c_int]),
            "orc_compute_geometry": (vp, [dp, dp, dp, sz, dp, dp, C.c_double, dp, sz, C.c_int]),
            "orc_compute_geometry_ordered": (vp, [dp, dp, dp, sz, dp, dp, C.c_double, dp, sz, C.c_int, C.c_int]),
            "orc_result_ring_entries": (C.c_uint64, [vp]),
            "orc_result_copy_rings": (None, [vp, dp, dp]),
            "orc_result_cells": (sz, [vp]),
            "orc_result_total_faces": (C.c_uint64, [vp]),
            "orc_result_copy": (None, [vp, dp, dp, dp, dp, dp]),
            "orc_result_stats": (None, [vp, dp]),
            "orc_result_seconds": (C.c_double, [vp]),
            "orc_result_free": (None, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _v(x):
    return np.ascontiguousarray(x, np.float64)


# ---- vector / plane (vector3.rs) -----------------------------------------
def dot(a, b):
    return lib().orc_dot(_p(_v(a)), _p(_v(b)))


def cross(a, b):
    out = np.empty(3)
    lib().orc_cross(_p(_v(a)), _p(_v(b)), _p(out))
    return out


def halfway_from_origin_to(p):
    out = np.empty(4)
    lib().orc_halfway_from_origin_to(_p(_v(p)), _p(out))
    return out


def build_from_non_unit_normal_and_point(n, pt):
    out = np.empty(4)
    lib().orc_build_from_non_unit_normal_and_point(_p(_v(n)), _p(_v(pt)), _p(out))
    return out


def location(sd, tol):
    return lib().orc_location(sd, tol)


def vector_location(plane, v, tol):
    return lib().orc_vector_location(_p(_v(plane)), _p(_v(v)), tol)


def intersection(plane, a, b):
    out = np.empty(3)
    lib().orc_intersection(_p(_v(plane)), _p(_v(a)), _p(_v(b)), _p(out))
    return out


# ---- grid (celery.rs) -------------------------------------------------------
def cells_per_dimension(n, density=1.25):
    return lib().orc_cells_per_dimension(n, density)


def soa(xyz):
    xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
    return [np.ascontiguousarray(xyz[:, k]) for k in range(3)]


ORDER_INPUT, ORDER_MORTON = 0, 1


def grid(xyz, density=1.25, order=ORDER_INPUT):
    """(grid info, bins, sorted_indices, delimiters)."""
    x, y, z = soa(xyz)
    n = len(x)
    g = Grid()
    lib().orc_grid_info(_p(x), _p(y), _p(z), n, density, C.byref(g))
    nb = g.cpd ** 3
    bins, srt, dl = np.empty(n, np.uint32), np.empty(n, np.uint32), np.empty(nb + 1, np.uint64)
    lib().orc_grid_build_ordered(_p(x), _p(y), _p(z), n, C.byref(g), order, _p(bins), _p(srt), _p(dl))
    return g, bins, srt, dl


def morton_codes(xyz):
    """Diagram::morton_index of every point (interface.rs:151-183)."""
    x, y, z = soa(xyz)
    out = np.empty(len(x), np.uint32)
    lib().orc_morton_codes(_p(x), _p(y), _p(z), len(x), _p(out))
    return out


def index_maps(xyz):
    """(original_of_internal, internal_of_original) of sort_particles."""
    x, y, z = soa(xyz)
    a, b = np.empty(len(x), np.uint32), np.empty(len(x), np.uint32)
    lib().orc_index_maps(_p(x), _p(y), _p(z), len(x), _p(a), _p(b))
    return a, b


def axis_index(v, vmin, vmax, inv, cpd):
    return lib().orc_axis_index(v, vmin, vmax, inv, cpd)


def search_order(cpd, sx, sy, sz, L=None):
    """Canonical shell table: (keys, offsets (m,3), generation indices, kcut)."""
    if L is None:
        L = cpd - 1
    m = lib().orc_search_order(cpd, sx, sy, sz, L, None, None, None, None)
    key, off, gen, kc = np.empty(m), np.empty((m, 3), np.int32), np.empty(m, np.uint64), C.c_double()
    lib().orc_search_order(cpd, sx, sy, sz, L, _p(key), _p(off), _p(gen), C.byref(kc))
    return key, off, gen, kc.value


def generation_index(cpd, ox, oy, oz):
    return lib().orc_generation_index(cpd, ox, oy, oz)


def expand_all(xyz, site, density=1.25, max_radius=None):
    """expand_all_no_radius (max_radius None) or expand_all_in_radius(max_radius)."""
    x, y, z = soa(xyz)
    out = np.empty(len(x), np.uint32)
    r = float("nan") if max_radius is None else float(max_radius)
    m = lib().orc_expand_all(_p(x), _p(y), _p(z), len(x), density, float(site[0]), float(site[1]),
                             float(site[2]), r, _p(out))
    return out[:m]


# ---- radius queries (celery.rs:700-855, interface.rs:348-365) ------------------
QUERY_CELLS, QUERY_CELL_RADIUS, QUERY_REAL_RADIUS, QUERY_CLOUD = 0, 1, 2, 3


def check_cell_in_range(g, x, y, z, radius, i, j, k):
    return bool(lib().orc_check_cell_in_range(C.byref(g), x, y, z, radius, i, j, k))


def find_cells_in_radius(g, x, y, z, radius):
    m = lib().orc_find_cells_in_radius(C.byref(g), x, y, z, radius, None)
    out = np.empty(m, np.uint32)
    lib().orc_find_cells_in_radius(C.byref(g), x, y, z, radius, _p(out))
    return out


def find_neighbors(xyz, g, srt, dl, x, y, z, radius, real=False):
    """find_neighbors_in_cell_radius (real False) / _in_real_radius (True)."""
    px, py, pz = soa(xyz)
    srt = np.ascontiguousarray(srt, np.uint32)
    dl = np.ascontiguousarray(dl, np.uint64)
    args = (_p(px), _p(py), _p(pz), C.byref(g), _p(srt), _p(dl), x, y, z, radius, int(real))
    m = lib().orc_find_neighbors(*args, None)
    out = np.empty(m, np.uint32)
    lib().orc_find_neighbors(*args, _p(out))
    return out


def queries(kind, xyz, q, radius, density=1.25, groups=None, target_group=-1, order=0):
    """Many queries of one kind over the grid of xyz: (offsets (nq+1), ids)."""
    px, py, pz = soa(xyz)
    q = np.ascontiguousarray(np.asarray(q, np.float64).reshape(-1, 3))
    nq = len(q)
    r = np.ascontiguousarray(np.broadcast_to(np.asarray(radius, np.float64), (nq,)))
    g = None if groups is None else np.ascontiguousarray(groups, np.int64)
    off = np.empty(nq + 1, np.uint64)
    args = [kind, _p(px), _p(py), _p(pz), len(px), density, _p(g), int(target_group), _p(q), nq, _p(r), _p(off)]
    lib().orc_queries_ordered(*args, None, order)
    ids = np.empty(int(off[-1]), np.uint32)
    lib().orc_queries_ordered(*args, _p(ids), order)
    return off, ids


# ---- polyhedron (polyhedron.rs) ----------------------------------------------
class Poly:
    def __init__(self, lo, hi, clip=CLIP_WALK):
        f = lib().orc_poly_new if clip == CLIP_WALK else lib().orc_poly_new_faces
        self.h = f(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_poly_free(self.h)
            self.h = None

    def cut(self, tag, plane):
        return lib().orc_poly_cut(self.h, tag, _p(_v(plane)))

    def counts(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        lib().orc_poly_counts(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def volume(self):
        return lib().orc_poly_volume(self.h)

    def faces(self):
        k = lib().orc_poly_faces(self.h, None, None, None, None, 0)
        tag, area, wn, deg = np.empty(k, np.int64), np.empty(k), np.empty((k, 3)), np.empty(k, np.int32)
        lib().orc_poly_faces(self.h, _p(tag), _p(area), _p(wn), _p(deg), k)
        return tag, area, wn, deg

    def vertices(self):
        k = lib().orc_poly_vertices(self.h, None, 0)
        out = np.empty((k, 3))
        lib().orc_poly_vertices(self.h, _p(out), k)
        return out

    def check(self):
        return bool(lib().orc_poly_check(self.h))

    def outgoing(self, plane):
        s, d = np.empty(3), np.empty(3)
        ok = lib().orc_poly_outgoing(self.h, _p(_v(plane)), _p(s), _p(d))
        return (s, d) if ok else None


# ---- cells (interface.rs) ------------------------------------------------
def compute(xyz, lo, hi, density=1.25, mode=MODE_SECURITY, clip=CLIP_FACES, cells=None, threads=0,
            max_radius=None, groups=None, target_group=None, order=0):
    """All (or the listed) cells: dict(volume, face_count, face_offset, neighbor, area, stats, seconds)."""
    x, y, z = soa(xyz)
    n = len(x)
    lo, hi = _v(lo), _v(hi)
    cl = None if cells is None else np.ascontiguousarray(cells, np.uint32)
    nc = n if cl is None else len(cl)
    g = None if groups is None else np.ascontiguousarray(groups, np.int64)
    r = lib().orc_compute_ordered(_p(x), _p(y), _p(z), n, _p(lo), _p(hi), density, mode, clip, _p(cl), nc, threads,
                                  float("nan") if max_radius is None else float(max_radius), _p(g),
                                  -1 if target_group is None else int(target_group), order)
    return _result(r, nc)


def compute_at_points(xyz, sites, lo, hi, density=1.25, mode=MODE_SECURITY, clip=CLIP_FACES, threads=0,
                      max_radius=None, groups=None, target_group=None, order=0):
    """Cells of sites that are not points of the set (get_cell_at_particle), same dict as compute;
    order as compute()."""
    x, y, z = soa(xyz)
    n = len(x)
    lo, hi = _v(lo), _v(hi)
    st = np.ascontiguousarray(np.asarray(sites, np.float64).reshape(-1, 3))
    g = None if groups is None else np.ascontiguousarray(groups, np.int64)
    r = lib().orc_compute_at_points_ordered(_p(x), _p(y), _p(z), n, _p(lo), _p(hi), density, mode, clip, _p(st),
                                            len(st), threads,
                                            float("nan") if max_radius is None else float(max_radius), _p(g),
                                            -1 if target_group is None else int(target_group), order)
    return _result(r, len(st))


def compute_geometry(xyz, lo, hi, density=1.25, cells=None, threads=0, order=0):
    """compute() plus every face's vertex ring: ring_offset (faces + 1) into
    ring_xyz (entries x 3), faces in output order, rings from the
    lexicographically smallest vertex, site-relative. order as compute()."""
    x, y, z = soa(xyz)
    n = len(x)
    lo, hi = _v(lo), _v(hi)
    cl = None if cells is None else np.ascontiguousarray(cells, np.uint32)
    nc = n if cl is None else len(cl)
    r = lib().orc_compute_geometry_ordered(_p(x), _p(y), _p(z), n, _p(lo), _p(hi), density, _p(cl), nc, threads,
                                           order)
    m = lib().orc_result_ring_entries(r)
    f = lib().orc_result_total_faces(r)
    ro, rx = np.empty(f + 1, np.uint64), np.empty((m, 3))
    lib().orc_result_copy_rings(r, _p(ro), _p(rx))
    out = _result(r, nc)
    out["ring_offset"], out["ring_xyz"] = ro, rx
    return out


def _result(r, nc):
    try:
        f = lib().orc_result_total_faces(r)
        vol, fc = np.empty(nc), np.empty(nc, np.uint32)
        fo, nb, ar = np.empty(nc + 1, np.uint64), np.empty(f, np.int32), np.empty(f)
        lib().orc_result_copy(r, _p(vol), _p(fc), _p(fo), _p(nb), _p(ar))
        st = np.empty(nc, STATS_DTYPE)
        lib().orc_result_stats(r, _p(st))
        sec = lib().orc_result_seconds(r)
    finally:
        lib().orc_result_free(r)
    return dict(volume=vol, face_count=fc, face_offset=fo, neighbor=nb, area=ar, stats=st, seconds=sec)
