"""ctypes bindings of the oracle (oracle/liboracle.so).

The oracle is the CPU restatement of the reference path that every parity test
checks the HIP product against; it is test infrastructure and is only loaded
from tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")


class OrcView(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("rot", C.c_float * 16), ("trans", C.c_float * 16),
                ("proj", C.c_float * 16), ("width", C.c_uint32), ("height", C.c_uint32),
                ("fov", C.c_float), ("znear", C.c_float), ("zfar", C.c_float)]


class OrcParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("lighting", C.c_int), ("srm", C.c_float),
                ("isovalue", C.c_float), ("lod", C.c_int), ("bit_width", C.c_int),
                ("tf1d", C.POINTER(C.c_uint8)), ("tf1d_n", C.c_int),
                ("tf2d", C.POINTER(C.c_uint8)), ("tf2d_w", C.c_int), ("tf2d_h", C.c_int),
                ("tf_scale", C.c_float), ("grad_scale", C.c_float),
                ("domain_scale", C.c_float * 3),
                ("light_ambient", C.c_float * 3), ("light_diffuse", C.c_float * 3),
                ("light_specular", C.c_float * 3), ("light_dir", C.c_float * 3),
                ("iso_color", C.c_float * 3), ("refine_steps", C.c_int),
                ("clamp_to_edge", C.c_int), ("ert_eps", C.c_float), ("clip_on", C.c_int),
                ("clip", C.c_float * 4), ("pow2", C.c_int), ("down8", C.c_int)]


class OrcBrick(C.Structure):
    _fields_ = [("key_index", C.c_uint64), ("toc_index", C.c_uint64),
                ("center", C.c_float * 3), ("extension", C.c_float * 3),
                ("voxel_count", C.c_uint32 * 3), ("tex_min", C.c_float * 3),
                ("tex_max", C.c_float * 3), ("distance", C.c_float), ("is_empty", C.c_int)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    u64p = C.POINTER(C.c_uint64)
    L.orc_octree_meta.restype = C.c_int
    L.orc_octree_meta.argtypes = [u64p, C.POINTER(C.c_uint32), C.c_uint32, u64p, u64p, u64p,
                                  C.POINTER(C.c_double), C.c_int]
    L.orc_ds_create.restype = C.c_void_p
    L.orc_ds_create.argtypes = [C.c_void_p, C.c_int, u64p, C.POINTER(C.c_uint32), C.c_uint32,
                                C.c_int, C.c_int, C.POINTER(C.c_float)]
    L.orc_ds_create_lod0.restype = C.c_void_p
    L.orc_ds_create_lod0.argtypes = [C.c_void_p, C.c_int, u64p, C.POINTER(C.c_uint32), C.c_uint32,
                                     C.c_int, C.POINTER(C.c_float)]
    L.orc_ds_destroy.argtypes = [C.c_void_p]
    L.orc_opacity_pow.restype = None
    L.orc_opacity_pow.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)]
    L.orc_ds_lod_count.restype = C.c_int
    L.orc_ds_lod_count.argtypes = [C.c_void_p]
    L.orc_ds_brick_count.restype = C.c_uint64
    L.orc_ds_brick_count.argtypes = [C.c_void_p]
    L.orc_ds_brick_info.argtypes = [C.c_void_p, C.c_uint64, u64p, C.POINTER(C.c_uint32), u64p,
                                    C.POINTER(C.c_double)]
    L.orc_ds_brick_data.restype = C.c_void_p
    L.orc_ds_brick_data.argtypes = [C.c_void_p, C.c_uint64]
    L.orc_ds_total_bytes.restype = C.c_uint64
    L.orc_ds_total_bytes.argtypes = [C.c_void_p]
    L.orc_ds_all_data.restype = C.c_void_p
    L.orc_ds_all_data.argtypes = [C.c_void_p]
    L.orc_ds_brick_md.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.POINTER(C.c_float),
                                  C.POINTER(C.c_float), C.POINTER(C.c_uint32)]
    L.orc_ds_texcoords.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_int, C.POINTER(C.c_float),
                                   C.POINTER(C.c_float)]
    L.orc_ds_range_max.restype = C.c_double
    L.orc_ds_range_max.argtypes = [C.c_void_p]
    L.orc_tf1d_bytes.argtypes = [C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_uint8), C.c_int]
    L.orc_tf1d_limits.argtypes = [C.POINTER(C.c_uint8), C.c_int, u64p]
    L.orc_tf2d_limits.argtypes = [C.POINTER(C.c_uint8), C.c_int, C.c_int, u64p]
    L.orc_plan.restype = C.c_int
    L.orc_plan.argtypes = [C.c_void_p, C.POINTER(OrcView), C.POINTER(OrcParams), C.POINTER(OrcBrick),
                           C.c_int]
    L.orc_min_lod.restype = C.c_int
    L.orc_min_lod.argtypes = [C.c_void_p, C.POINTER(OrcView), C.c_float]
    L.orc_render.restype = C.c_int
    L.orc_render.argtypes = [C.c_void_p, C.POINTER(OrcView), C.POINTER(OrcParams),
                             C.POINTER(C.c_int), C.POINTER(C.c_float), u64p, C.c_int]
    L.orc_render_fp64.restype = C.c_int
    L.orc_render_fp64.argtypes = [C.c_void_p, C.POINTER(OrcView), C.POINTER(OrcParams),
                                  C.POINTER(C.c_int), C.c_double, C.POINTER(C.c_float), C.c_int]
    L.orc_round_fp16.restype = C.c_float
    L.orc_round_fp16.argtypes = [C.c_float]
    L.orc_brick_shader_vars.argtypes = [C.c_void_p, C.POINTER(OrcParams), C.POINTER(OrcBrick),
                                        C.POINTER(C.c_float), C.POINTER(C.c_float),
                                        C.POINTER(C.c_float)]
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def octree_meta(dims, brick, overlap, maxl=64):
    L = lib()
    d = np.asarray(dims, np.uint64)
    b = np.asarray(brick, np.uint32)
    sz = np.zeros((maxl, 3), np.uint64)
    bc = np.zeros((maxl, 3), np.uint64)
    off = np.zeros(maxl, np.uint64)
    asp = np.zeros((maxl, 3), np.float64)
    n = L.orc_octree_meta(_p(d, C.c_uint64), _p(b, C.c_uint32), overlap, _p(sz, C.c_uint64),
                          _p(bc, C.c_uint64), _p(off, C.c_uint64), _p(asp, C.c_double), maxl)
    return dict(lod_size=sz[:n], lod_bricks=bc[:n], lod_offset=off[:n], lod_aspect=asp[:n])


class Dataset:
    """Oracle dataset: ExtendedOctree conversion of a linear volume."""

    def __init__(self, vol, brick, overlap, clamp=False, median=False, scale=(1, 1, 1), lod0_only=False,
                 dims=None):
        """vol: (z, y, x) uint8/uint16 array; or None with dims (x, y, z) for the
        metadata and brick table alone (no voxels converted)."""
        L = lib()
        if vol is None:
            self.vol = None
            self.bpv = 1
            self.dims = tuple(int(v) for v in dims)
        else:
            vol = np.ascontiguousarray(vol)
            assert vol.dtype in (np.uint8, np.uint16)
            self.vol = vol
            self.bpv = vol.dtype.itemsize
            nz, ny, nx = vol.shape
            self.dims = (nx, ny, nz)
        d = np.asarray(self.dims, np.uint64)
        b = np.asarray(brick, np.uint32)
        sc = np.asarray(scale, np.float32)
        if vol is None:
            self.h = L.orc_ds_create_lod0(None, 1, _p(d, C.c_uint64), _p(b, C.c_uint32), overlap, 0,
                                          _p(sc, C.c_float))
        elif lod0_only:
            self.h = L.orc_ds_create_lod0(vol.ctypes.data, self.bpv, _p(d, C.c_uint64), _p(b, C.c_uint32),
                                          overlap, int(clamp), _p(sc, C.c_float))
        else:
            self.h = L.orc_ds_create(vol.ctypes.data, self.bpv, _p(d, C.c_uint64), _p(b, C.c_uint32),
                                     overlap, int(clamp), int(median), _p(sc, C.c_float))
        if not self.h:
            raise RuntimeError("oracle conversion failed")
        self.brick = tuple(brick)
        self.overlap = overlap

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ds_destroy(self.h)
            self.h = None

    @property
    def lod_count(self):
        return lib().orc_ds_lod_count(self.h)

    @property
    def brick_count(self):
        return int(lib().orc_ds_brick_count(self.h))

    def brick_info(self, i):
        c = np.zeros(4, np.uint64)
        s = np.zeros(3, np.uint32)
        off = np.zeros(1, np.uint64)
        mm = np.zeros(2, np.float64)
        lib().orc_ds_brick_info(self.h, i, _p(c, C.c_uint64), _p(s, C.c_uint32), _p(off, C.c_uint64),
                                _p(mm, C.c_double))
        return dict(coords=tuple(int(v) for v in c), size=tuple(int(v) for v in s),
                    offset=int(off[0]), minmax=(float(mm[0]), float(mm[1])))

    def brick_data(self, i):
        info = self.brick_info(i)
        nx, ny, nz = info["size"]
        n = nx * ny * nz
        ptr = lib().orc_ds_brick_data(self.h, i)
        dt = np.uint8 if self.bpv == 1 else np.uint16
        buf = (C.c_uint8 * (n * self.bpv)).from_address(ptr)
        return np.frombuffer(buf, dtype=dt).copy().reshape(nz, ny, nx)

    def all_data(self):
        n = int(lib().orc_ds_total_bytes(self.h))
        ptr = lib().orc_ds_all_data(self.h)
        buf = (C.c_uint8 * n).from_address(ptr)
        return np.frombuffer(buf, dtype=np.uint8).copy()

    def brick_md(self, lod, lin):
        c = np.zeros(3, np.float32)
        e = np.zeros(3, np.float32)
        n = np.zeros(3, np.uint32)
        lib().orc_ds_brick_md(self.h, lod, lin, _p(c, C.c_float), _p(e, C.c_float), _p(n, C.c_uint32))
        return c, e, n

    def texcoords(self, lod, lin, pow2=False):
        a = np.zeros(3, np.float32)
        b = np.zeros(3, np.float32)
        lib().orc_ds_texcoords(self.h, lod, lin, int(pow2), _p(a, C.c_float), _p(b, C.c_float))
        return a, b

    def range_max(self):
        return float(lib().orc_ds_range_max(self.h))

    def _meta(self):
        if not hasattr(self, "_m"):
            self._m = octree_meta(self.dims, self.brick, self.overlap)
        return self._m

    def lod_bricks(self, lod):
        return tuple(int(v) for v in self._meta()["lod_bricks"][lod])

    def lod_size(self, lod):
        return tuple(int(v) for v in self._meta()["lod_size"][lod])

    def lod_offset(self, lod):
        return int(self._meta()["lod_offset"][lod])


def tf1d_bytes(rgba, used_range=255):
    rgba = np.ascontiguousarray(rgba, np.float32)
    n = rgba.shape[0]
    out = np.zeros((n, 4), np.uint8)
    lib().orc_tf1d_bytes(_p(rgba, C.c_float), n, _p(out, C.c_uint8), used_range)
    return out


def tf1d_limits(rgba8):
    rgba8 = np.ascontiguousarray(rgba8, np.uint8)
    lim = np.zeros(2, np.uint64)
    lib().orc_tf1d_limits(_p(rgba8, C.c_uint8), rgba8.shape[0], _p(lim, C.c_uint64))
    return int(lim[0]), int(lim[1])


def tf2d_limits(rgba8):
    rgba8 = np.ascontiguousarray(rgba8, np.uint8)
    h, w = rgba8.shape[:2]
    lim = np.zeros(4, np.uint64)
    lib().orc_tf2d_limits(_p(rgba8, C.c_uint8), w, h, _p(lim, C.c_uint64))
    return tuple(int(v) for v in lim)


def make_view(camera):
    """camera: dict with view/rot/trans/proj (16 floats each), width, height, fov, znear, zfar."""
    v = OrcView()
    for k in ("view", "rot", "trans", "proj"):
        arr = np.asarray(camera[k], np.float32).reshape(16)
        getattr(v, k)[:] = [float(x) for x in arr]
    v.width, v.height = camera["width"], camera["height"]
    v.fov, v.znear, v.zfar = camera["fov"], camera["znear"], camera["zfar"]
    return v


def make_params(p, keep):
    """p: dict of render parameters; keep: list that keeps TF arrays alive."""
    o = OrcParams()
    o.mode = p["mode"]
    o.lighting = int(p["lighting"])
    o.srm = p.get("srm", 1.0)
    o.isovalue = p.get("isovalue", 0.0)
    o.lod = p.get("lod", 0)
    o.bit_width = p["bit_width"]
    if p.get("tf1d") is not None:
        t = np.ascontiguousarray(p["tf1d"], np.uint8)
        keep.append(t)
        o.tf1d = _p(t, C.c_uint8)
        o.tf1d_n = t.shape[0]
    if p.get("tf2d") is not None:
        t = np.ascontiguousarray(p["tf2d"], np.uint8)
        keep.append(t)
        o.tf2d = _p(t, C.c_uint8)
        o.tf2d_h, o.tf2d_w = t.shape[0], t.shape[1]
    o.tf_scale = p.get("tf_scale", 1.0)
    o.grad_scale = p.get("grad_scale", 1.0)
    o.domain_scale[:] = list(p.get("domain_scale", (1.0, 1.0, 1.0)))
    o.light_ambient[:] = list(p.get("light_ambient", (0.1, 0.1, 0.1)))
    o.light_diffuse[:] = list(p.get("light_diffuse", (1.0, 1.0, 1.0)))
    o.light_specular[:] = list(p.get("light_specular", (1.0, 1.0, 1.0)))
    o.light_dir[:] = list(p.get("light_dir", (0.0, 0.0, -1.0)))
    o.iso_color[:] = list(p.get("iso_color", (0.5, 0.5, 0.5)))
    o.refine_steps = p.get("refine_steps", 5)
    o.clamp_to_edge = int(p.get("clamp_to_edge", 0))
    o.ert_eps = p.get("ert_eps", 0.0)
    if p.get("clip") is not None:
        o.clip_on = 1
        o.clip[:] = [float(v) for v in p["clip"]]
    o.pow2 = int(bool(p.get("pow2", False)))
    o.down8 = int(bool(p.get("down8", False)))
    return o


def plan(ds, camera, params):
    keep = []
    v = make_view(camera)
    prm = make_params(params, keep)
    cap = 1 << 16
    arr = (OrcBrick * cap)()
    n = lib().orc_plan(ds.h, C.byref(v), C.byref(prm), arr, cap)
    assert n >= 0
    return [arr[i] for i in range(n)]


def min_lod(ds, camera, sse=1.0):
    v = make_view(camera)
    return lib().orc_min_lod(ds.h, C.byref(v), sse)


def opacity_pow(y, s):
    """pow(y, s) of the oracle's opacity correction, elementwise (s broadcast)."""
    y = np.ascontiguousarray(y, np.float32)
    s = np.ascontiguousarray(np.broadcast_to(np.asarray(s, np.float32), y.shape))
    out = np.empty_like(y)
    lib().orc_opacity_pow(_p(y, C.c_float), _p(s, C.c_float), int(y.size), _p(out, C.c_float))
    return out


def render(ds, camera, params, roi=None, threads=0):
    keep = []
    v = make_view(camera)
    prm = make_params(params, keep)
    W, H = camera["width"], camera["height"]
    if roi is None:
        roi = (0, 0, W, H)
    r = np.asarray(roi, np.int32)
    out = np.zeros((roi[3] - roi[1], roi[2] - roi[0], 4), np.float32)
    ns = np.zeros(1, np.uint64)
    rc = lib().orc_render(ds.h, C.byref(v), C.byref(prm), _p(r, C.c_int), _p(out, C.c_float),
                          _p(ns, C.c_uint64), threads)
    assert rc == 0
    return out, int(ns[0])


def render_fp64(ds, camera, params, step_model, roi=None, threads=0):
    keep = []
    v = make_view(camera)
    prm = make_params(params, keep)
    W, H = camera["width"], camera["height"]
    if roi is None:
        roi = (0, 0, W, H)
    r = np.asarray(roi, np.int32)
    out = np.zeros((roi[3] - roi[1], roi[2] - roi[0], 4), np.float32)
    rc = lib().orc_render_fp64(ds.h, C.byref(v), C.byref(prm), _p(r, C.c_int), step_model,
                               _p(out, C.c_float), threads)
    assert rc == 0
    return out


def round_fp16(x):
    return float(lib().orc_round_fp16(float(x)))


def look_at(eye, at=(0, 0, 0), up=(0, 1, 0)):
    L = lib()
    L.orc_look_at.argtypes = [C.POINTER(C.c_float)] * 4
    e, a, u = (np.asarray(v, np.float32) for v in (eye, at, up))
    out = np.zeros(16, np.float32)
    L.orc_look_at(_p(e, C.c_float), _p(a, C.c_float), _p(u, C.c_float), _p(out, C.c_float))
    return out.reshape(4, 4)


def perspective(fov, aspect, znear, zfar):
    L = lib()
    L.orc_perspective.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float)]
    out = np.zeros(16, np.float32)
    L.orc_perspective(fov, aspect, znear, zfar, _p(out, C.c_float))
    return out.reshape(4, 4)


def rotation_y(angle):
    L = lib()
    L.orc_rotation_y.argtypes = [C.c_double, C.POINTER(C.c_float)]
    out = np.zeros(16, np.float32)
    L.orc_rotation_y(angle, _p(out, C.c_float))
    return out.reshape(4, 4)


def camera(width, height, eye=(0, 0, 2), fov=45.0, znear=0.01, zfar=1000.0, rot=None):
    """Camera dict of GLRenderer::ComputeViewAndProjection built with the oracle's own math."""
    return dict(view=look_at(eye), proj=perspective(fov, width / height, znear, zfar),
                rot=np.eye(4, dtype=np.float32) if rot is None else np.asarray(rot, np.float32),
                trans=np.eye(4, dtype=np.float32), width=width, height=height, fov=fov, znear=znear, zfar=zfar)


# ---- step-proposal network ---------------------------------------------------------
def _sn_lib():
    L = lib()
    if not getattr(L, "_sn_ready", False):
        L.orc_sn_param_count.restype = C.c_int
        L.orc_sn_forward.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.c_int,
                                     C.POINTER(C.c_double)]
        L.orc_sn_grad.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int,
                                  C.c_int, C.POINTER(C.c_double)]
        L._sn_ready = True
    return L


def sn_param_count():
    return _sn_lib().orc_sn_param_count()


def sn_forward(theta, x, bf16=True):
    theta = np.ascontiguousarray(theta, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros((x.shape[0], 2), np.float64)
    _sn_lib().orc_sn_forward(theta.ctypes.data_as(C.POINTER(C.c_float)), x.ctypes.data_as(C.POINTER(C.c_float)),
                             x.shape[0], int(bf16), y.ctypes.data_as(C.POINTER(C.c_double)))
    return y


def sn_grad(theta, x, tgt, bf16=True):
    """(gradient of the mean weighted loss, summed weighted loss)"""
    theta = np.ascontiguousarray(theta, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    tgt = np.ascontiguousarray(tgt, np.float32)
    g = np.zeros(theta.size + 1, np.float64)
    _sn_lib().orc_sn_grad(theta.ctypes.data_as(C.POINTER(C.c_float)), x.ctypes.data_as(C.POINTER(C.c_float)),
                          tgt.ctypes.data_as(C.POINTER(C.c_float)), x.shape[0], int(bf16),
                          g.ctypes.data_as(C.POINTER(C.c_double)))
    return g[:-1], g[-1]


# ---- GridLeaper traversal --------------------------------------------------------
def leap_pass(ds, camera, params, eye, resident=None, first=True, resume_in=None):
    """One traversal pass of the oracle: (rgba[H,W,4], resume[H,W,8], missing[n], samples)."""
    L = lib()
    if not getattr(L, "_leap_ready", False):
        L.orc_leap_pass.restype = C.c_int
        L.orc_leap_pass.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_uint8),
                                    C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.POINTER(C.c_uint8), C.POINTER(C.c_uint64)]
        L._leap_ready = True
    keep = []
    v = make_view(camera)
    prm = make_params(params, keep)
    W, H = camera["width"], camera["height"]
    out = np.zeros((H, W, 4), np.float32)
    res = np.zeros((H, W, 8), np.float32)
    miss = np.zeros(ds.brick_count, np.uint8)
    e = np.ascontiguousarray(eye, np.float32)
    rin = np.ascontiguousarray(resume_in if resume_in is not None else res, np.float32)
    r8 = None if resident is None else np.ascontiguousarray(resident, np.uint8)
    ns = C.c_uint64()
    rc = L.orc_leap_pass(ds.h, C.byref(v), C.byref(prm), e.ctypes.data_as(C.POINTER(C.c_float)),
                         r8.ctypes.data_as(C.POINTER(C.c_uint8)) if r8 is not None else None, int(first),
                         rin.ctypes.data_as(C.POINTER(C.c_float)), out.ctypes.data_as(C.POINTER(C.c_float)),
                         res.ctypes.data_as(C.POINTER(C.c_float)), miss.ctypes.data_as(C.POINTER(C.c_uint8)),
                         C.byref(ns))
    assert rc == 0, "brick grid of the pool differs from the octree's"
    return out, res, miss, ns.value


# ---- culling, filter and brick-table order (checked against oracle/_ref) ------------
def cull_planes(camera):
    L = lib()
    L.orc_cull_planes.argtypes = [C.POINTER(OrcView), C.POINTER(C.c_float)]
    v = make_view(camera)
    out = np.zeros(24, np.float32)
    L.orc_cull_planes(C.byref(v), _p(out, C.c_float))
    return out.reshape(6, 4)


def cull_visible(camera, centers, extents):
    L = lib()
    L.orc_cull_visible.argtypes = [C.POINTER(OrcView), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_cull_visible.restype = C.c_int
    v = make_view(camera)
    c = np.ascontiguousarray(centers, np.float32)
    e = np.ascontiguousarray(extents, np.float32)
    return np.array([bool(L.orc_cull_visible(C.byref(v), _p(c[i], C.c_float), _p(e[i], C.c_float)))
                     for i in range(c.shape[0])])


def cull_lod_level(camera, centers, extents, voxels, sse=1.0):
    L = lib()
    L.orc_cull_lod_level.argtypes = [C.POINTER(OrcView), C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                     C.POINTER(C.c_uint32)]
    L.orc_cull_lod_level.restype = C.c_int
    v = make_view(camera)
    c = np.ascontiguousarray(centers, np.float32)
    e = np.ascontiguousarray(extents, np.float32)
    n = np.ascontiguousarray(voxels, np.uint32)
    return np.array([L.orc_cull_lod_level(C.byref(v), sse, _p(c[i], C.c_float), _p(e[i], C.c_float),
                                          _p(n[i], C.c_uint32)) for i in range(c.shape[0])], np.int32)


def filter_groups(taps, median, vals):
    L = lib()
    L.orc_filter.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p]
    vals = np.ascontiguousarray(vals)
    out = np.zeros(vals.shape[0], vals.dtype)
    L.orc_filter(taps, int(median), vals.dtype.itemsize, vals.ctypes.data, vals.shape[0], out.ctypes.data)
    return out


def table_order(ds):
    L = lib()
    L.orc_ds_table_order.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.c_uint64]
    L.orc_ds_table_order.restype = C.c_uint64
    n = int(L.orc_ds_table_order(ds.h, None, 0))
    out = np.zeros(2 * n, np.uint64)
    L.orc_ds_table_order(ds.h, _p(out, C.c_uint64), n)
    return out.reshape(-1, 2)
