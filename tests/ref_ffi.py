"""ctypes bindings of oracle/_ref/libtvkref.so: reference code (Renderer/CullingLOD.cpp,
Basics/Vectors.h, VolumeTools.h Filter, Brick.h BrickTable) compiled where it lies by
`make -C oracle ref`.  Test infrastructure: the oracle's restatements are checked
against it.  The library exists only where the reference tree was present at build
time; tests that need it skip otherwise."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_ref", "libtvkref.so")

_lib = None
f32p = C.POINTER(C.c_float)


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        L.ref_look_at.argtypes = [f32p] * 4
        L.ref_perspective.argtypes = [C.c_float] * 4 + [f32p]
        L.ref_mat_mul.argtypes = [f32p] * 3
        cam = [f32p, f32p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32]
        L.ref_cull_planes.argtypes = cam + [C.c_float, f32p, f32p]
        L.ref_cull_visible.argtypes = cam + [f32p, f32p, C.c_int, C.POINTER(C.c_uint8)]
        L.ref_cull_lod_level.argtypes = cam + [C.c_float, f32p, f32p, C.POINTER(C.c_uint32), C.c_int,
                                               C.POINTER(C.c_int)]
        L.ref_filter.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p]
        L.ref_table_order.argtypes = [C.POINTER(C.c_uint64), C.c_int, C.POINTER(C.c_uint64)]
        _lib = L
    return _lib


def _f(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(f32p)


def look_at(eye, at, up):
    out = np.zeros(16, np.float32)
    e, ep = _f(eye)
    a, ap = _f(at)
    u, up_ = _f(up)
    lib().ref_look_at(ep, ap, up_, out.ctypes.data_as(f32p))
    return out.reshape(4, 4)


def perspective(fov, aspect, znear, zfar):
    out = np.zeros(16, np.float32)
    lib().ref_perspective(fov, aspect, znear, zfar, out.ctypes.data_as(f32p))
    return out.reshape(4, 4)


def mat_mul(a, b):
    out = np.zeros(16, np.float32)
    a, ap = _f(a)
    b, bp = _f(b)
    lib().ref_mat_mul(ap, bp, out.ctypes.data_as(f32p))
    return out.reshape(4, 4)


def modelview(camera):
    """GLRenderer.cpp:627 with the reference's MATRIX4 product: rot * trans * view."""
    return mat_mul(mat_mul(camera["rot"], camera["trans"]), camera["view"])


def _cam_args(camera):
    mv, mvp = _f(modelview(camera))
    pj, pjp = _f(camera["proj"])
    keep = (mv, pj)
    aspect = camera["width"] / camera["height"]
    return keep, [mvp, pjp, camera["fov"], aspect, camera["znear"], camera["zfar"], camera["height"]]


def cull_planes(camera, sse=1.0):
    keep, args = _cam_args(camera)
    planes = np.zeros(24, np.float32)
    lf = C.c_float()
    lib().ref_cull_planes(*args, sse, planes.ctypes.data_as(f32p), C.byref(lf))
    return planes.reshape(6, 4), lf.value


def cull_visible(camera, centers, extents):
    keep, args = _cam_args(camera)
    c, cp = _f(centers)
    e, ep = _f(extents)
    n = c.shape[0]
    out = np.zeros(n, np.uint8)
    lib().ref_cull_visible(*args, cp, ep, n, out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out.astype(bool)


def cull_lod_level(camera, centers, extents, voxels, sse=1.0):
    keep, args = _cam_args(camera)
    c, cp = _f(centers)
    e, ep = _f(extents)
    v = np.ascontiguousarray(voxels, np.uint32)
    n = c.shape[0]
    out = np.zeros(n, np.int32)
    lib().ref_cull_lod_level(*args, sse, cp, ep, v.ctypes.data_as(C.POINTER(C.c_uint32)), n,
                             out.ctypes.data_as(C.POINTER(C.c_int)))
    return out


def filter_groups(taps, median, vals):
    vals = np.ascontiguousarray(vals)
    n = vals.shape[0]
    out = np.zeros(n, vals.dtype)
    lib().ref_filter(taps, int(median), vals.dtype.itemsize, vals.ctypes.data, n, out.ctypes.data)
    return out


def table_order(lod_bricks):
    b = np.ascontiguousarray(lod_bricks, np.uint64).reshape(-1, 3)
    total = int(np.prod(b, axis=1).sum())
    out = np.zeros(2 * total, np.uint64)
    lib().ref_table_order(b.ctypes.data_as(C.POINTER(C.c_uint64)), b.shape[0],
                          out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out.reshape(-1, 2)
