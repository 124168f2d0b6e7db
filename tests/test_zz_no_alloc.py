"""tess_compute_noalloc (no counterpart in the reference: for callers that
cannot allocate device memory) against the allocating path it stands for,
tess_diagram_create_device + tess_diagram_compute: bitwise equal outputs on
two calls in a row, the second with a smaller input, through the C ABI.
The two calls share the library's scratch array, so the second also checks
that nothing of the first one's contents leaks into it."""
import ctypes as C

import numpy as np
import pytest

import tess

pytestmark = pytest.mark.gpu
LO, HI = np.zeros(3), np.ones(3)
TESS_E_INVALID, TESS_E_CAPACITY = 1, 4  # include/tess.h


def _lib(ctx):
    lib = ctx.lib
    f = lib.tess_compute_noalloc
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    f.argtypes = [vp, vp, C.c_size_t, dp, dp, C.c_double, vp, vp, vp, vp, vp, C.c_uint64, C.POINTER(C.c_uint64)]
    f.restype = C.c_int
    return lib


def _down(dev, dtype, count):
    out = np.empty(count, dtype)
    if count:
        assert tess._hip().hipMemcpy(out.ctypes.data, dev.ptr, out.nbytes, 2) == 0  # device to host
    return out


def _noalloc(ctx, pts, lo, hi, face_cap=None):
    lib = _lib(ctx)
    n = len(pts)
    cap = 64 * n if face_cap is None else face_cap
    dx = tess.DeviceArray(ctx, np.ascontiguousarray(pts, np.float64))
    vol = tess.DeviceArray(ctx, np.zeros(n))
    fc = tess.DeviceArray(ctx, np.zeros(n, np.uint32))
    fo = tess.DeviceArray(ctx, np.zeros(n + 1, np.uint64))
    nb = tess.DeviceArray(ctx, np.zeros(max(cap, 1), np.int32))
    ar = tess.DeviceArray(ctx, np.zeros(max(cap, 1)))
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    total = C.c_uint64()
    dp = C.POINTER(C.c_double)
    st = lib.tess_compute_noalloc(ctx.h, dx.ptr, n, lo.ctypes.data_as(dp), hi.ctypes.data_as(dp), 1.25, vol.ptr,
                                  fc.ptr, fo.ptr, nb.ptr, ar.ptr, cap, C.byref(total))
    if st != tess.TESS_OK:
        return st, total.value, None
    f = total.value
    return st, f, dict(volume=_down(vol, np.float64, n), face_count=_down(fc, np.uint32, n),
                       face_offset=_down(fo, np.uint64, n + 1), neighbor=_down(nb, np.int32, f),
                       area=_down(ar, np.float64, f))


def _allocating(ctx, pts, lo, hi):
    d = tess.Diagram(ctx, tess.DeviceArray(ctx, np.ascontiguousarray(pts, np.float64)), lo, hi)
    d.compute()
    out = d.cells()
    d.close()
    return out


def _sphere_ring(m=200, seed=7, r=0.3, bg=3000):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(m, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    c = np.full(3, 0.5)
    b = tess.uniform(bg, seed + 1)
    b = b[np.linalg.norm(b - c, axis=1) > 1.5 * r]
    return np.vstack([c[None, :], c + r * v, b])


def test_noalloc_equals_compute_twice(ctx):
    runs = [(tess.uniform(16000, 901), np.full(3, -0.5), np.full(3, 1.5)),  # largest input, long searches
            (_sphere_ring(), LO, HI)]                                       # smaller, with a spilled cell
    assert len(runs[1][0]) < len(runs[0][0])
    for pts, lo, hi in runs:
        print(f"noalloc call: {len(pts)} points, box {lo[0]}..{hi[0]}", flush=True)
        st, _, got = _noalloc(ctx, pts, lo, hi)
        assert st == tess.TESS_OK, ctx.lib.tess_last_error(ctx.h).decode()
        ref = _allocating(ctx, pts, lo, hi)
        for k in ("volume", "face_count", "face_offset", "neighbor", "area"):
            assert np.array_equal(got[k], ref[k]), k


def test_noalloc_limits(ctx):
    pts = tess.uniform(3000, 902)
    st, total, _ = _noalloc(ctx, pts, LO, HI, face_cap=100)
    assert st == TESS_E_CAPACITY and total > 100
    assert b"face_cap" in ctx.lib.tess_last_error(ctx.h)
    big = tess.uniform(tess.NOALLOC_MAX_POINTS + 1, 903)
    st, _, _ = _noalloc(ctx, big, LO, HI)
    assert st == TESS_E_INVALID
    assert b"at most" in ctx.lib.tess_last_error(ctx.h)
    # the refusals leave the scratch usable
    st, _, got = _noalloc(ctx, pts, LO, HI)
    assert st == tess.TESS_OK
    ref = _allocating(ctx, pts, LO, HI)
    assert np.array_equal(got["neighbor"], ref["neighbor"]) and np.array_equal(got["volume"], ref["volume"])
