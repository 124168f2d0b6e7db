"""ctypes binding of libtess (include/tess.h) and of the synthetic workloads.

The classes mirror the reference's interface.rs surface:
``Diagram`` (interface.rs:24-232), ``Cell`` (interface.rs:236-390) and the
per-face accessors of ``VoronoiFace`` (interface.rs:393-417).  Every call goes
through the C ABI of ``libtess.so``; a missing library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TESS_LIBRARY") or os.path.join(HERE, "libtess.so")
DATA_PATH = os.path.join(HERE, "libtessdata.so")

TESS_OK = 0
STATUS = {0: "TESS_OK", 1: "TESS_E_INVALID", 2: "TESS_E_NOMEM", 3: "TESS_E_DEVICE",
          4: "TESS_E_CAPACITY", 5: "TESS_E_DEGENERATE", 6: "TESS_E_STATE", 7: "TESS_E_NOTFOUND"}
MODE_PLAIN, MODE_LEARNED = 0, 1
CELL_CAPACITY, CELL_TABLE, CELL_DEGENERATE, CELL_FALLBACK, CELL_RERUN = 1, 2, 4, 8, 16
CELL_LONG_TABLE, CELL_SPILL, CELL_ARENA = 32, 64, 128
NOALLOC_MAX_POINTS = 16384

EXPORTS = [
    "tess_opts_default", "tess_ctx_create", "tess_ctx_destroy", "tess_last_error", "tess_ctx_stream",
    "tess_ctx_synchronize", "tess_diagram_create", "tess_diagram_create_device", "tess_diagram_destroy",
    "tess_diagram_size", "tess_diagram_compute", "tess_diagram_total_faces", "tess_diagram_copy_cells",
    "tess_diagram_device_cells", "tess_diagram_copy_stats", "tess_diagram_timings", "tess_diagram_grid",
    "tess_diagram_copy_grid", "tess_cell_at_index", "tess_cell_destroy", "tess_cell_original_index",
    "tess_cell_volume", "tess_cell_neighbors", "tess_cell_face_count", "tess_cell_face",
    "tess_cell_face_vertices", "tess_cell_vertices", "tess_model_create", "tess_model_destroy",
    "tess_model_get", "tess_model_set", "tess_model_save", "tess_model_load", "tess_model_add_samples",
    "tess_model_sample_count", "tess_model_train", "tess_diagram_predict", "tess_diagram_compute_learned",
    "tess_diagram_compute_scheduled", "tess_diagram_features", "tess_mlp_forward", "tess_mlp_gradient",
    "tess_mfma_probe", "tess_batch_create", "tess_batch_destroy", "tess_batch_sets", "tess_batch_set_config",
    "tess_batch_get_config", "tess_batch_compute", "tess_batch_tune", "tess_batch_total_faces",
    "tess_batch_copy_cells", "tess_batch_timings", "tess_diag_arith", "tess_diagram_query",
    "tess_diagram_query_device", "tess_query_count", "tess_query_total", "tess_query_copy", "tess_query_device",
    "tess_query_timings", "tess_query_destroy", "tess_diagram_cells_at_particles", "tess_site_cells_count",
    "tess_site_cells_total_faces", "tess_site_cells_copy", "tess_site_cells_copy_stats", "tess_site_cells_timings",
    "tess_site_cells_destroy", "tess_diagram_compute_geometry", "tess_diagram_geometry_sizes",
    "tess_diagram_copy_geometry", "tess_diagram_geometry_timings", "tess_diagram_copy_index_maps",
    "tess_cell_at_particle", "tess_compute_noalloc",
]
QUERY_CELLS, QUERY_CELL_RADIUS, QUERY_REAL_RADIUS, QUERY_CLOUD = 0, 1, 2, 3
ORDER_INPUT, ORDER_MORTON = 0, 1
MLP_IN, MLP_PARAMS = 32, 7312


class TessError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Opts(C.Structure):
    _fields_ = [("cell_density", C.c_double), ("mode", C.c_int32), ("vertex_capacity", C.c_int32),
                ("collect_stats", C.c_int32), ("xcd_chunk", C.c_int32), ("point_order", C.c_int32),
                ("reserved", C.c_int32)]


class CellStats(C.Structure):
    _fields_ = [("streamed", C.c_uint32), ("classified", C.c_uint32), ("cuts", C.c_uint32),
                ("entries", C.c_uint32), ("flops", C.c_uint64), ("r2", C.c_double),
                ("max_vertices", C.c_uint32), ("status", C.c_uint32), ("incident_cuts", C.c_uint32),
                ("reserved", C.c_uint32)]


STATS_DTYPE = np.dtype([("streamed", "<u4"), ("classified", "<u4"), ("cuts", "<u4"), ("entries", "<u4"),
                        ("flops", "<u8"), ("r2", "<f8"), ("max_vertices", "<u4"), ("status", "<u4"),
                        ("incident_cuts", "<u4"), ("reserved", "<u4")])
assert STATS_DTYPE.itemsize == C.sizeof(CellStats)


class Timings(C.Structure):
    _fields_ = [("bin_ms", C.c_double), ("cells_ms", C.c_double), ("compact_ms", C.c_double),
                ("total_ms", C.c_double), ("table_ms", C.c_double), ("flops", C.c_uint64),
                ("fallback_cells", C.c_uint64), ("capacity_reruns", C.c_uint64), ("plan_ms", C.c_double),
                ("table_reruns", C.c_uint64), ("spill_cells", C.c_uint64), ("incident_cuts", C.c_uint64),
                ("planned_long", C.c_uint64)]


class BatchConfig(C.Structure):
    _fields_ = [("first_capacity", C.c_int32), ("xcd_chunk", C.c_int32), ("redo_chunk", C.c_int32),
                ("learned", C.c_int32), ("eager", C.c_int32)]


_lib = None
_data = None


def load_library(path: str | None = None) -> C.CDLL:
    """Load libtess.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"{p} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    vp, sz, dp = C.c_void_p, C.c_size_t, C.c_void_p
    sig = {
        "tess_opts_default": (None, [vp]),
        "tess_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "tess_ctx_destroy": (None, [vp]),
        "tess_last_error": (C.c_char_p, [vp]),
        "tess_ctx_stream": (vp, [vp]),
        "tess_ctx_synchronize": (C.c_int, [vp]),
        "tess_diagram_create": (C.c_int, [vp, dp, sz, dp, dp, dp, vp, C.POINTER(vp)]),
        "tess_diagram_create_device": (C.c_int, [vp, vp, sz, dp, dp, vp, vp, C.POINTER(vp)]),
        "tess_diagram_destroy": (None, [vp]),
        "tess_diagram_size": (sz, [vp]),
        "tess_diagram_compute": (C.c_int, [vp]),
        "tess_diagram_total_faces": (C.c_int, [vp, C.POINTER(C.c_uint64)]),
        "tess_diagram_copy_cells": (C.c_int, [vp, dp, dp, dp, dp, dp]),
        "tess_diagram_device_cells": (C.c_int, [vp, vp]),
        "tess_diagram_copy_stats": (C.c_int, [vp, dp]),
        "tess_diagram_timings": (C.c_int, [vp, C.POINTER(Timings)]),
        "tess_diagram_grid": (C.c_int, [vp, dp, dp, C.POINTER(C.c_int64)]),
        "tess_diagram_copy_grid": (C.c_int, [vp, dp, dp, dp]),
        "tess_cell_at_index": (C.c_int, [vp, sz, C.c_double, C.c_int64, C.POINTER(vp)]),
        "tess_cell_at_particle": (C.c_int, [vp, dp, C.c_double, C.c_int64, C.POINTER(vp)]),
        "tess_cell_destroy": (None, [vp]),
        "tess_cell_original_index": (C.c_int64, [vp]),
        "tess_cell_volume": (C.c_double, [vp]),
        "tess_cell_neighbors": (sz, [vp, dp, sz]),
        "tess_cell_face_count": (sz, [vp]),
        "tess_cell_face": (C.c_int, [vp, sz, dp, dp]),
        "tess_cell_face_vertices": (sz, [vp, sz, dp, sz]),
        "tess_cell_vertices": (sz, [vp, dp, sz]),
        "tess_model_create": (C.c_int, [vp, C.c_uint64, C.POINTER(vp)]),
        "tess_model_destroy": (None, [vp]),
        "tess_model_get": (C.c_int, [vp, dp, dp, dp]),
        "tess_model_set": (C.c_int, [vp, dp, dp, dp]),
        "tess_model_save": (C.c_int, [vp, C.c_char_p]),
        "tess_model_load": (C.c_int, [vp, C.c_char_p]),
        "tess_model_add_samples": (C.c_int, [vp, vp]),
        "tess_model_sample_count": (sz, [vp]),
        "tess_model_train": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_uint64, dp]),
        "tess_diagram_predict": (C.c_int, [vp, vp, dp]),
        "tess_diagram_compute_learned": (C.c_int, [vp, vp]),
        "tess_diagram_compute_scheduled": (C.c_int, [vp, dp, dp]),
        "tess_diagram_features": (C.c_int, [vp, dp]),
        "tess_mlp_forward": (C.c_int, [vp, dp, sz, dp]),
        "tess_mlp_gradient": (C.c_int, [vp, dp, dp, sz, dp, dp]),
        "tess_mfma_probe": (C.c_int, [vp, dp, dp, dp]),
        "tess_batch_create": (C.c_int, [vp, sz, dp, dp, dp, dp, vp, C.POINTER(vp)]),
        "tess_batch_destroy": (None, [vp]),
        "tess_batch_sets": (sz, [vp]),
        "tess_batch_set_config": (C.c_int, [vp, C.POINTER(BatchConfig)]),
        "tess_batch_get_config": (C.c_int, [vp, C.POINTER(BatchConfig)]),
        "tess_batch_compute": (C.c_int, [vp, vp]),
        "tess_batch_tune": (C.c_int, [vp, vp, C.c_char_p, C.POINTER(BatchConfig)]),
        "tess_batch_total_faces": (C.c_int, [vp, sz, C.POINTER(C.c_uint64)]),
        "tess_batch_copy_cells": (C.c_int, [vp, sz, dp, dp, dp, dp, dp]),
        "tess_batch_timings": (C.c_int, [vp, C.POINTER(Timings)]),
        "tess_diag_arith": (C.c_int, [vp, dp, dp, sz, dp, dp, dp]),
        "tess_compute_noalloc": (C.c_int, [vp, vp, sz, dp, dp, C.c_double, vp, vp, vp, vp, vp, C.c_uint64,
                                           C.POINTER(C.c_uint64)]),
        "tess_diagram_query": (C.c_int, [vp, C.c_int, dp, sz, dp, sz, C.c_int64, C.POINTER(vp)]),
        "tess_diagram_query_device": (C.c_int, [vp, C.c_int, vp, sz, vp, sz, C.c_int64, C.POINTER(vp)]),
        "tess_diagram_cells_at_particles": (C.c_int, [vp, dp, sz, C.c_double, C.c_int64, C.POINTER(vp)]),
        "tess_diagram_compute_geometry": (C.c_int, [vp]),
        "tess_diagram_geometry_sizes": (C.c_int, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "tess_diagram_copy_geometry": (C.c_int, [vp, dp, dp, dp, dp]),
        "tess_diagram_geometry_timings": (C.c_int, [vp, dp]),
        "tess_diagram_copy_index_maps": (C.c_int, [vp, dp, dp]),
        "tess_site_cells_count": (sz, [vp]),
        "tess_site_cells_total_faces": (C.c_int, [vp, C.POINTER(C.c_uint64)]),
        "tess_site_cells_copy": (C.c_int, [vp, dp, dp, dp, dp, dp]),
        "tess_site_cells_copy_stats": (C.c_int, [vp, dp]),
        "tess_site_cells_timings": (C.c_int, [vp, dp, C.POINTER(C.c_uint32)]),
        "tess_site_cells_destroy": (None, [vp]),
        "tess_query_count": (sz, [vp]),
        "tess_query_total": (C.c_uint64, [vp]),
        "tess_query_copy": (C.c_int, [vp, dp, dp]),
        "tess_query_device": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp)]),
        "tess_query_timings": (C.c_int, [vp, dp]),
        "tess_query_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


_hip_rt = None


def _hip():
    """The HIP runtime libtess.so is linked against (already loaded with it)."""
    global _hip_rt
    if _hip_rt is None:
        load_library()
        h = C.CDLL("libamdhip64.so.7")
        h.hipSetDevice.argtypes = [C.c_int]
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        for f in (h.hipSetDevice, h.hipMalloc, h.hipFree, h.hipMemcpy):
            f.restype = C.c_int
        _hip_rt = h
    return _hip_rt


class DeviceArray:
    """A copy of a numpy array in device memory, for the entry points that
    take device-resident inputs (tess_diagram_create_device,
    tess_diagram_query_device).  Allocated with the HIP runtime libtess.so
    uses: a second HIP runtime in the process (such as the one bundled with
    a PyTorch wheel) would not share its device state."""

    def __init__(self, ctx, a):
        a = np.ascontiguousarray(a)
        self.shape, self.dtype, self.nbytes = a.shape, a.dtype, a.nbytes
        hip = _hip()
        hip.hipSetDevice(ctx.device)
        p = C.c_void_p()
        if hip.hipMalloc(C.byref(p), max(1, a.nbytes)) != 0:
            raise TessError(2, "hipMalloc failed")
        self.ptr = p
        if a.nbytes and hip.hipMemcpy(p, _ptr(a), a.nbytes, 1) != 0:  # host to device
            self.close()
            raise TessError(3, "hipMemcpy failed")

    def close(self):
        if getattr(self, "ptr", None):
            _hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One HIP device + stream (tess_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.tess_ctx_create(device, C.byref(h))
        if st != TESS_OK:
            raise TessError(st, f"cannot open HIP device {device}")
        self.h = h
        self.device = device
        self._diagrams = weakref.WeakSet()  # destroyed before the context

    def check(self, st):
        if st != TESS_OK:
            raise TessError(st, self.lib.tess_last_error(self.h).decode())

    def synchronize(self):
        self.check(self.lib.tess_ctx_synchronize(self.h))

    def diag_arith(self, a, b):
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        s, r, q = (np.empty_like(a) for _ in range(3))
        self.check(self.lib.tess_diag_arith(self.h, _ptr(a), _ptr(b), a.size, _ptr(s), _ptr(r), _ptr(q)))
        return s, r, q

    def mfma_probe(self, a_bits, b_bits):
        """One 16x16x32 bf16 MFMA: a [16][32], b [32][16] bf16 bit patterns -> fp32 [16][16]."""
        a = np.ascontiguousarray(a_bits, np.uint16)
        b = np.ascontiguousarray(b_bits, np.uint16)
        assert a.shape == (16, 32) and b.shape == (32, 16)
        d = np.empty((16, 16), np.float32)
        self.check(self.lib.tess_mfma_probe(self.h, _ptr(a), _ptr(b), _ptr(d)))
        return d

    def close(self):
        if self.h:
            for d in list(self._diagrams):
                d.close()
            self.lib.tess_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Cell:
    """A computed cell (interface.rs:236-390); faces are VoronoiFace views."""

    def __init__(self, lib, h):
        self.lib, self.h = lib, h

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.tess_cell_destroy(self.h)
            self.h = None

    def original_index(self):
        return int(self.lib.tess_cell_original_index(self.h))

    def compute_volume(self):
        return float(self.lib.tess_cell_volume(self.h))

    def compute_neighbors(self):
        n = self.lib.tess_cell_neighbors(self.h, None, 0)
        out = np.empty(n, np.int64)
        self.lib.tess_cell_neighbors(self.h, _ptr(out), n)
        return out

    def face_count(self):
        return int(self.lib.tess_cell_face_count(self.h))

    def face(self, f):
        area, nb = C.c_double(), C.c_int64()
        st = self.lib.tess_cell_face(self.h, f, C.byref(area), C.byref(nb))
        if st != TESS_OK:
            raise TessError(st, "bad face index")
        return area.value, nb.value

    def face_vertices(self, f):
        k = self.lib.tess_cell_face_vertices(self.h, f, None, 0)
        out = np.empty((k, 3))
        self.lib.tess_cell_face_vertices(self.h, f, _ptr(out), k)
        return out

    def compute_vertices(self):
        k = self.lib.tess_cell_vertices(self.h, None, 0)
        out = np.empty((k, 3))
        self.lib.tess_cell_vertices(self.h, _ptr(out), k)
        return out


class Diagram:
    """Points + container box (tess_diagram, interface.rs:24-84)."""

    def __init__(self, ctx: Context, xyz, lo, hi, groups=None, density=1.25, mode=MODE_PLAIN,
                 vertex_capacity=0, collect_stats=False, xcd_chunk=0, point_order=ORDER_INPUT):
        self.ctx, self.lib = ctx, ctx.lib
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        o = Opts()
        self.lib.tess_opts_default(C.byref(o))
        o.cell_density, o.mode, o.vertex_capacity, o.collect_stats = density, mode, vertex_capacity, int(collect_stats)
        o.xcd_chunk = xcd_chunk
        o.point_order = point_order
        h = C.c_void_p()
        if isinstance(xyz, DeviceArray):  # device-resident AoS points (and groups)
            assert xyz.dtype == np.float64 and xyz.nbytes % 24 == 0
            self.n = xyz.nbytes // 24
            self.xyz = None
            gp = None if groups is None else groups.ptr
            ctx.check(self.lib.tess_diagram_create_device(ctx.h, xyz.ptr, self.n, _ptr(lo), _ptr(hi), gp,
                                                          C.byref(o), C.byref(h)))
        else:
            xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
            self.n = xyz.shape[0]
            self.xyz = xyz
            g = None if groups is None else np.ascontiguousarray(groups, np.int64)
            ctx.check(self.lib.tess_diagram_create(ctx.h, _ptr(xyz), self.n, _ptr(lo), _ptr(hi), _ptr(g),
                                                   C.byref(o), C.byref(h)))
        self.h = h
        ctx._diagrams.add(self)

    def close(self):
        if self.h and self.ctx.h:
            self.lib.tess_diagram_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compute(self):
        st = self.lib.tess_diagram_compute(self.h)
        if st != TESS_OK:
            raise TessError(st, self.lib.tess_last_error(self.ctx.h).decode())

    def compute_status(self):
        """Run compute and return the status code instead of raising."""
        return self.lib.tess_diagram_compute(self.h)

    def cells(self):
        tf = C.c_uint64()
        self.ctx.check(self.lib.tess_diagram_total_faces(self.h, C.byref(tf)))
        n, f = self.n, tf.value
        vol, fc = np.empty(n), np.empty(n, np.uint32)
        fo, nb, ar = np.empty(n + 1, np.uint64), np.empty(f, np.int32), np.empty(f)
        self.ctx.check(self.lib.tess_diagram_copy_cells(self.h, _ptr(vol), _ptr(fc), _ptr(fo), _ptr(nb), _ptr(ar)))
        return dict(volume=vol, face_count=fc, face_offset=fo, neighbor=nb, area=ar)

    def stats(self):
        out = np.empty(self.n, STATS_DTYPE)
        self.ctx.check(self.lib.tess_diagram_copy_stats(self.h, _ptr(out)))
        return out

    def timings(self):
        t = Timings()
        self.ctx.check(self.lib.tess_diagram_timings(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timings._fields_}

    def grid(self):
        b, s, cpd = np.empty(6), np.empty(3), C.c_int64()
        self.ctx.check(self.lib.tess_diagram_grid(self.h, _ptr(b), _ptr(s), C.byref(cpd)))
        return b, s, cpd.value

    def index_maps(self):
        """(original_of_internal, internal_of_original): the reference's
        Morton internal order (sort_particles, interface.rs:88-148)."""
        a, b = np.empty(self.n, np.uint32), np.empty(self.n, np.uint32)
        self.ctx.check(self.lib.tess_diagram_copy_index_maps(self.h, _ptr(a), _ptr(b)))
        return a, b

    def copy_grid(self):
        _, _, cpd = self.grid()
        cell = np.empty(self.n, np.uint32)
        srt = np.empty(self.n, np.uint32)
        dl = np.empty(cpd ** 3 + 1, np.uint64)
        self.ctx.check(self.lib.tess_diagram_copy_grid(self.h, _ptr(cell), _ptr(srt), _ptr(dl)))
        return cell, srt, dl

    def compute_learned(self, model):
        """All cells with the learned schedule (same results as compute)."""
        self.ctx.check(self.lib.tess_diagram_compute_learned(self.h, model.h))

    def compute_scheduled(self, log1p_c, t1=math.inf, t2=math.inf):
        """All cells scheduled from given per-point log(1 + candidates)."""
        v = np.ascontiguousarray(log1p_c, np.float32)
        assert v.shape == (self.n,)
        cal = np.array([t1, t2], np.float32)
        self.ctx.check(self.lib.tess_diagram_compute_scheduled(self.h, _ptr(v), _ptr(cal)))

    def predict(self, model):
        """[n][2]: predicted final security radius and classified-candidate count."""
        out = np.empty((self.n, 2), np.float32)
        self.ctx.check(self.lib.tess_diagram_predict(self.h, model.h, _ptr(out)))
        return out

    def features(self):
        """bf16 bit patterns [n][32] of the 3x3x3 occupancy features."""
        out = np.empty((self.n, MLP_IN), np.uint16)
        self.ctx.check(self.lib.tess_diagram_features(self.h, _ptr(out)))
        return out

    def query(self, kind, positions, radius, target_group=None):
        """Radius queries (tess_diagram_query): CSR offsets [nq + 1] and ids;
        the ids of query i are ids[offsets[i]:offsets[i + 1]].  radius: one
        value or one per query.  Device times (count pass, fill pass, whole
        call) are kept in last_query_ms."""
        q = np.ascontiguousarray(np.asarray(positions, np.float64).reshape(-1, 3))
        r = np.ascontiguousarray(np.atleast_1d(np.asarray(radius, np.float64)))
        g = -1 if target_group is None else int(target_group)
        h = C.c_void_p()
        self.ctx.check(self.lib.tess_diagram_query(self.h, kind, _ptr(q), len(q), _ptr(r), r.size, g, C.byref(h)))
        res = QueryResult(self.ctx, h)
        try:
            off, ids = res.copy()
            self.last_query_ms = res.timings()
        finally:
            res.close()
        return off, ids

    def query_device(self, kind, xyz_ptr, nq, radius_ptr, nradius, target_group=None):
        """Radius queries over device-resident positions (AoS f64) and radii;
        the results stay on the device in the returned QueryResult."""
        g = -1 if target_group is None else int(target_group)
        h = C.c_void_p()
        self.ctx.check(self.lib.tess_diagram_query_device(self.h, kind, xyz_ptr, nq, radius_ptr, nradius, g,
                                                          C.byref(h)))
        return QueryResult(self.ctx, h)

    def find_cells_in_radius(self, x, y, z, radius):
        """Celery::find_cells_in_radius (celery.rs:753-797): grid cell ids."""
        return self.query(QUERY_CELLS, [x, y, z], radius)[1]

    def find_neighbors_in_cell_radius(self, x, y, z, radius):
        """Celery::find_neighbors_in_cell_radius (celery.rs:802-819)."""
        return self.query(QUERY_CELL_RADIUS, [x, y, z], radius)[1]

    def find_neighbors_in_real_radius(self, x, y, z, radius):
        """Celery::find_neighbors_in_real_radius (celery.rs:825-855)."""
        return self.query(QUERY_REAL_RADIUS, [x, y, z], radius)[1]

    def neighbor_cloud(self, index, radius, target_group=None):
        """Cell::compute_neighbor_cloud (interface.rs:348-365) of the cell of
        point `index` (its position is the point's; host-created diagrams)."""
        return self.query(QUERY_CLOUD, self.xyz[index], radius, target_group)[1]

    def geometry(self):
        """Vertices and face rings of every cell (tess_diagram_compute_geometry):
        vertex_offset [n + 1], vertices [V, 3] (site-relative), ring_offset
        [faces + 1], ring (vertex numbers within the cell)."""
        self.ctx.check(self.lib.tess_diagram_compute_geometry(self.h))
        nv, nr, tf = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.ctx.check(self.lib.tess_diagram_geometry_sizes(self.h, C.byref(nv), C.byref(nr)))
        self.ctx.check(self.lib.tess_diagram_total_faces(self.h, C.byref(tf)))
        vo, v = np.empty(self.n + 1, np.uint64), np.empty((nv.value, 3))
        ro, ring = np.empty(tf.value + 1, np.uint64), np.empty(nr.value, np.uint16)
        self.ctx.check(self.lib.tess_diagram_copy_geometry(self.h, _ptr(vo), _ptr(v), _ptr(ro), _ptr(ring)))
        ms = np.zeros(2)
        self.ctx.check(self.lib.tess_diagram_geometry_timings(self.h, _ptr(ms)))
        return dict(vertex_offset=vo, vertices=v, ring_offset=ro, ring=ring,
                    timings=dict(stage_ms=float(ms[0]), compact_ms=float(ms[1])))

    def cells_at_particles(self, sites, search_radius=None, target_group=None, stats=False):
        """Diagram::get_cell_at_particle + compute_voronoi_cell for sites that
        are not points of the diagram: dict as cells() (+ stats, timings)."""
        q = np.ascontiguousarray(np.asarray(sites, np.float64).reshape(-1, 3))
        r = math.nan if search_radius is None else float(search_radius)
        g = -1 if target_group is None else int(target_group)
        h = C.c_void_p()
        self.ctx.check(self.lib.tess_diagram_cells_at_particles(self.h, _ptr(q), len(q), r, g, C.byref(h)))
        try:
            m = len(q)
            tf = C.c_uint64()
            self.ctx.check(self.lib.tess_site_cells_total_faces(h, C.byref(tf)))
            vol, fc = np.empty(m), np.empty(m, np.uint32)
            fo, nb, ar = np.empty(m + 1, np.uint64), np.empty(tf.value, np.int32), np.empty(tf.value)
            self.ctx.check(self.lib.tess_site_cells_copy(h, _ptr(vol), _ptr(fc), _ptr(fo), _ptr(nb), _ptr(ar)))
            out = dict(volume=vol, face_count=fc, face_offset=fo, neighbor=nb, area=ar)
            if stats:
                st = np.empty(m, STATS_DTYPE)
                self.ctx.check(self.lib.tess_site_cells_copy_stats(h, _ptr(st)))
                out["stats"] = st
            ms, rr = np.zeros(2), C.c_uint32()
            self.ctx.check(self.lib.tess_site_cells_timings(h, _ptr(ms), C.byref(rr)))
            out["timings"] = dict(cells_ms=float(ms[0]), compact_ms=float(ms[1]), capacity_reruns=rr.value)
        finally:
            self.lib.tess_site_cells_destroy(h)
        return out

    def get_cell_at_particle(self, xyz, search_radius=None, target_group=None):
        """interface.rs:211-232 + Cell::compute_voronoi_cell for a site that
        is not a point of the diagram."""
        h = C.c_void_p()
        p = np.ascontiguousarray(xyz, np.float64).reshape(3)
        r = math.nan if search_radius is None else float(search_radius)
        g = -1 if target_group is None else int(target_group)
        self.ctx.check(self.lib.tess_cell_at_particle(self.h, _ptr(p), r, g, C.byref(h)))
        return Cell(self.lib, h)

    def get_cell_at_index(self, index, search_radius=None, target_group=None):
        """interface.rs:186-208 + Cell::compute_voronoi_cell."""
        h = C.c_void_p()
        r = math.nan if search_radius is None else float(search_radius)
        g = -1 if target_group is None else int(target_group)
        self.ctx.check(self.lib.tess_cell_at_index(self.h, index, r, g, C.byref(h)))
        return Cell(self.lib, h)


class QueryResult:
    """CSR results of one tess_diagram_query call, held on the device."""

    def __init__(self, ctx, h):
        self.ctx, self.lib, self.h = ctx, ctx.lib, h
        self.nq = self.lib.tess_query_count(h)
        self.total = self.lib.tess_query_total(h)

    def copy(self):
        off = np.empty(self.nq + 1, np.uint64)
        ids = np.empty(self.total, np.uint32)
        self.ctx.check(self.lib.tess_query_copy(self.h, _ptr(off), _ptr(ids)))
        return off, ids

    def timings(self):
        ms = np.zeros(3)
        self.ctx.check(self.lib.tess_query_timings(self.h, _ptr(ms)))
        return dict(count_ms=float(ms[0]), fill_ms=float(ms[1]), total_ms=float(ms[2]))

    def close(self):
        if self.h:
            self.lib.tess_query_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Model:
    """The learned search-radius model (tess_model): 27 -> 64 -> 64 -> 2, bf16 MFMA."""

    def __init__(self, ctx: Context, seed: int = 0):
        self.ctx, self.lib = ctx, ctx.lib
        h = C.c_void_p()
        ctx.check(self.lib.tess_model_create(ctx.h, seed, C.byref(h)))
        self.h = h
        ctx._diagrams.add(self)

    def close(self):
        if self.h and self.ctx.h:
            self.lib.tess_model_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self):
        p, norm, cal = np.empty(MLP_PARAMS, np.float32), np.empty(4, np.float32), np.empty(4, np.float32)
        self.ctx.check(self.lib.tess_model_get(self.h, _ptr(p), _ptr(norm), _ptr(cal)))
        return p, norm, cal

    def set(self, params=None, norm=None, calib=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        nm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        cl = None if calib is None else np.ascontiguousarray(calib, np.float32)
        if p is not None:
            assert p.shape == (MLP_PARAMS,)
        self.ctx.check(self.lib.tess_model_set(self.h, _ptr(p), _ptr(nm), _ptr(cl)))

    def save(self, path):
        self.ctx.check(self.lib.tess_model_save(self.h, os.fsencode(path)))

    def load(self, path):
        self.ctx.check(self.lib.tess_model_load(self.h, os.fsencode(path)))

    def add_samples(self, diagram):
        self.ctx.check(self.lib.tess_model_add_samples(self.h, diagram.h))

    def sample_count(self):
        return int(self.lib.tess_model_sample_count(self.h))

    def train(self, epochs=4, batch=4096, lr=1e-3, seed=0):
        loss = np.empty(max(epochs, 1), np.float32)
        self.ctx.check(self.lib.tess_model_train(self.h, epochs, batch, lr, seed, _ptr(loss)))
        return loss[:epochs]

    def forward(self, x_bits):
        x = np.ascontiguousarray(x_bits, np.uint16)
        y = np.empty((x.shape[0], 2), np.float32)
        self.ctx.check(self.lib.tess_mlp_forward(self.h, _ptr(x), x.shape[0], _ptr(y)))
        return y

    def gradient(self, x_bits, t):
        x = np.ascontiguousarray(x_bits, np.uint16)
        t = np.ascontiguousarray(t, np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = C.c_float()
        self.ctx.check(self.lib.tess_mlp_gradient(self.h, _ptr(x), _ptr(t), x.shape[0], _ptr(g), C.byref(loss)))
        return loss.value, g


class Batch:
    """Many independent point sets tessellated together (tess_batch).

    sets: list of (xyz [n][3], lo [3], hi [3]).  Per set, results are bitwise
    those of a Diagram of the same points and box."""

    def __init__(self, ctx: Context, sets, density=1.25, collect_stats=False):
        self.ctx, self.lib = ctx, ctx.lib
        self._pts = [np.ascontiguousarray(p, np.float64).reshape(-1, 3) for p, _, _ in sets]
        self.sizes = [len(p) for p in self._pts]
        ptrs = (C.c_void_p * len(sets))(*[p.ctypes.data for p in self._pts])
        n = np.array(self.sizes, np.uint64)
        lo = np.ascontiguousarray([l for _, l, _ in sets], np.float64)
        hi = np.ascontiguousarray([h for _, _, h in sets], np.float64)
        o = Opts()
        self.lib.tess_opts_default(C.byref(o))
        o.cell_density, o.collect_stats = density, int(collect_stats)
        h = C.c_void_p()
        ctx.check(self.lib.tess_batch_create(ctx.h, len(sets), C.cast(ptrs, C.c_void_p), _ptr(n), _ptr(lo), _ptr(hi),
                                             C.byref(o), C.byref(h)))
        self.h = h
        self._pts = None
        ctx._diagrams.add(self)

    def close(self):
        if self.h and self.ctx.h:
            self.lib.tess_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self):
        k = BatchConfig()
        self.ctx.check(self.lib.tess_batch_get_config(self.h, C.byref(k)))
        return {f: getattr(k, f) for f, _ in BatchConfig._fields_}

    def set_config(self, **kw):
        cur = self.config()
        cur.update(kw)
        k = BatchConfig(**cur)
        self.ctx.check(self.lib.tess_batch_set_config(self.h, C.byref(k)))

    def compute(self, model=None):
        self.ctx.check(self.lib.tess_batch_compute(self.h, model.h if model is not None else None))

    def tune(self, model=None, cache_path=None):
        k = BatchConfig()
        self.ctx.check(self.lib.tess_batch_tune(self.h, model.h if model is not None else None,
                                                os.fsencode(cache_path) if cache_path else None, C.byref(k)))
        return {f: getattr(k, f) for f, _ in BatchConfig._fields_}

    def cells(self, s):
        tf = C.c_uint64()
        self.ctx.check(self.lib.tess_batch_total_faces(self.h, s, C.byref(tf)))
        n, f = self.sizes[s], tf.value
        vol, fc = np.empty(n), np.empty(n, np.uint32)
        fo, nb, ar = np.empty(n + 1, np.uint64), np.empty(f, np.int32), np.empty(f)
        self.ctx.check(self.lib.tess_batch_copy_cells(self.h, s, _ptr(vol), _ptr(fc), _ptr(fo), _ptr(nb), _ptr(ar)))
        return dict(volume=vol, face_count=fc, face_offset=fo, neighbor=nb, area=ar)

    def timings(self):
        t = Timings()
        self.ctx.check(self.lib.tess_batch_timings(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timings._fields_}


# ------------------------------------------------------------ workloads
def _data_lib():
    global _data
    if _data is None:
        if not os.path.exists(DATA_PATH):
            raise OSError(f"{DATA_PATH} is missing: build it first")
        d = C.CDLL(DATA_PATH)
        d.tg_uniform.argtypes = [C.c_uint64, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
        d.tg_clustered.argtypes = [C.c_uint64, C.c_size_t, C.c_int, C.c_double, C.c_double, C.c_void_p]
        d.tg_batch_shape.argtypes = [C.c_uint64, C.c_double, C.c_double, C.POINTER(C.c_uint64), C.c_void_p]
        d.tg_batch_points.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_void_p]
        d.tg_fcc_count.argtypes = [C.c_int]
        d.tg_fcc_count.restype = C.c_size_t
        d.tg_fcc.argtypes = [C.c_uint64, C.c_int, C.c_double, C.c_double, C.c_void_p]
        _data = d
    return _data


def uniform(n, seed, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """n points i.i.d. uniform in [lo, hi) (xoshiro256** via SplitMix64)."""
    d = _data_lib()
    out = np.empty((n, 3))
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    d.tg_uniform(seed, n, _ptr(lo), _ptr(hi), _ptr(out))
    return out


def clustered(n, seed, nclusters=64, sigma=0.02, frac=0.8):
    """80% from 64 Gaussian clusters (sigma 0.02, resampled inside the unit cube), 20% uniform."""
    d = _data_lib()
    out = np.empty((n, 3))
    d.tg_clustered(seed, n, nclusters, sigma, frac, _ptr(out))
    return out


def batch_set(seed, nmin=1000.0, nmax=200000.0):
    """One uneven batch set: (points, box side)."""
    d = _data_lib()
    n = C.c_uint64()
    side = np.empty(3)
    d.tg_batch_shape(seed, nmin, nmax, C.byref(n), _ptr(side))
    out = np.empty((n.value, 3))
    d.tg_batch_points(seed, nmin, nmax, _ptr(out))
    return out, side


def fcc(seed, ncell=63, a=1.0, sigma=1e-3):
    """ncell^3 FCC cells (cubic constant a), 4 sites each, Gaussian jitter sigma."""
    d = _data_lib()
    out = np.empty((d.tg_fcc_count(ncell), 3))
    d.tg_fcc(seed, ncell, a, sigma, _ptr(out))
    return out
