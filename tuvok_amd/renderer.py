"""HIPRaycaster: the AbstrRenderer-shaped front end of the HIP raycasting path.

Method names follow Tuvok's AbstrRenderer (Renderer/AbstrRenderer.h:112-881)
so code written against GLRaycaster/GLGridLeaper maps one to one:

    ren = HIPRaycaster()
    ren.LoadVolume(vol, brick=(66,)*3, overlap=1)      # RAWConverter + UVF load
    ren.Set1DTrans(tf.GetByteArray())                   # AbstrRenderer::Set1DTrans
    ren.SetRendermode(RM_1DTRANS); ren.SetUseLighting(False)
    ren.SetViewParameters(45, 0.01, 1000); ren.SetViewPosition((0,0,2))
    ren.Resize(512, 512)
    img = ren.Paint()                                   # RGBA32F, premultiplied, GL row order

Every call goes through the C ABI of libtvkhip.so; there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import RM_1DTRANS, RM_2DTRANS, RM_ISOSURFACE, TvkError  # noqa: F401


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def rotation_y(angle):
    """MATRIX4::RotationY (Basics/Vectors.h:994-1002), row-vector layout."""
    c, s = np.float32(np.cos(angle)), np.float32(np.sin(angle))
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, -s, s, c
    return m


class HIPRaycaster:
    def __init__(self, device=0, use_only_power_of_two=False, downsample_to_8bits=False, disable_border=None):
        """AbstrRenderer(pMasterController, bUseOnlyPowerOfTwo, bDownSampleTo8Bits,
        bDisableBorder) (AbstrRenderer.h:160-174).  disable_border (None: the
        clamp_to_edge parameter's default) selects GL_CLAMP_TO_EDGE textures
        without the repeated padding element (GPUMemManDataStructs.cpp:340-360)."""
        self._L = _abi.load()
        h = C.c_void_p()
        rc = self._L.tvk_create(C.byref(h), device)
        if rc != 0:
            raise TvkError(rc, "tvk_create failed (no GPU?)")
        self._h = h
        self._pow2, self._down8 = bool(use_only_power_of_two), bool(downsample_to_8bits)
        self._chk(self._L.tvk_set_compat(self._h, int(self._pow2), int(self._down8)))
        self._p = _abi.Params()
        self._L.tvk_default_params(C.byref(self._p))
        self._p.lod = 0      # RT_CAPTURE renders LOD 0 (AbstrRenderer.cpp:1188-1190)
        if disable_border is not None:
            self._p.clamp_to_edge = int(bool(disable_border))
        self._eye = np.array([0, 0, 1.6], np.float32)   # AbstrRenderer.cpp:63-69
        self._at = np.array([0, 0, 0], np.float32)
        self._up = np.array([0, 1, 0], np.float32)
        self._fov, self._znear, self._zfar = 50.0, 0.01, 1000.0
        self._w, self._h_px = 0, 0
        self._rot = np.eye(4, dtype=np.float32)
        self._user = None
        self.last_stats = _abi.Stats()

    # -- plumbing -----------------------------------------------------------
    def _chk(self, rc):
        if rc != 0:
            raise TvkError(rc, self._L.tvk_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.tvk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- dataset (RAWConverter / UVFDataset / BrickedDataset) ---------------
    def OpenUVF(self, path, budget_bytes=0):
        """A UVF file's ExtendedOctree dataset into the pool (0: every brick in HBM);
        returns the file's description.  The dataset's aspect becomes the
        domain scale (vDomainScale = 1 / aspect)."""
        info = _abi.UvfInfo()
        self._chk(self._L.tvk_open_uvf(self._h, str(path).encode(), int(budget_bytes), C.byref(info)))
        for k in range(3):
            self._p.domain_scale[k] = 1.0 / info.aspect[k]
        self._chk(self._L.tvk_set_params(self._h, C.byref(self._p)))
        return info

    def SetDispatchOrder(self, order):
        """How multi-view renders hand their 16x16 groups to the GPU:
        DISPATCH_RASTER, DISPATCH_GROUPS or DISPATCH_ROWS (most costly first
        within each XCD's eighth), or'ed with DISPATCH_QUEUES (a persistent
        launch whose waves pull 8x8 tiles from per-XCD queues; default
        RASTER | QUEUES).  Frames are identical in every setting."""
        self._chk(self._L.tvk_set_dispatch_order(self._h, int(order)))
        self._dispatch = int(order)

    def SetLODFilter(self, median):
        """bComputeMedian of ExtendedOctreeConverter::Convert for the next bricking."""
        self._chk(self._L.tvk_set_lod_filter(self._h, int(bool(median))))

    def LoadVolume(self, vol, brick, overlap, clamp_to_edge=False, max_lods=0, keep_linear=False):
        vol = np.ascontiguousarray(vol)
        dt = {np.dtype(np.uint8): _abi.TVK_U8, np.dtype(np.uint16): _abi.TVK_U16}[vol.dtype]
        dims = np.array(vol.shape[::-1], np.uint64)
        b = np.array(brick, np.uint32)
        self._chk(self._L.tvk_load_volume(self._h, vol.ctypes.data, dt, dims.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          b.ctypes.data_as(C.POINTER(C.c_uint32)), overlap, int(clamp_to_edge),
                                          max_lods, int(keep_linear)))

    def GenerateVolume(self, kind, seed, dtype, dims, brick, overlap, clamp_to_edge=False, max_lods=0,
                       keep_linear=False):
        dt = {np.dtype(np.uint8): _abi.TVK_U8, np.dtype(np.uint16): _abi.TVK_U16}[np.dtype(dtype)]
        d = np.array(dims, np.uint64)
        b = np.array(brick, np.uint32)
        self._chk(self._L.tvk_generate_volume(self._h, kind, seed, dt, d.ctypes.data_as(C.POINTER(C.c_uint64)),
                                              b.ctypes.data_as(C.POINTER(C.c_uint32)), overlap, int(clamp_to_edge),
                                              max_lods, int(keep_linear)))

    def SetBricks(self, dtype, dims, brick, overlap, keys, lod_count=0, budget_bytes=0):
        """A dataset handed over brick by brick (BrickedDataset): keys is a list of
        (lod, index, (min, max)) covering the ExtendedOctree layout of dims (x, y, z),
        brick (stored size) and overlap; payloads follow with UploadBrick.
        budget_bytes > 0 keeps the bricks in host memory and pages them into
        that much HBM."""
        dt = {np.dtype(np.uint8): _abi.TVK_U8, np.dtype(np.uint16): _abi.TVK_U16}[np.dtype(dtype)]
        arr = (_abi.BrickKey * max(len(keys), 1))()
        for i, (lod, index, mm) in enumerate(keys):
            arr[i].lod, arr[i].index = int(lod), int(index)
            arr[i].minmax[0], arr[i].minmax[1] = float(mm[0]), float(mm[1])
        d = np.array(dims, np.uint64)
        b = np.array(brick, np.uint32)
        self._chk(self._L.tvk_set_bricks(self._h, dt, d.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         b.ctypes.data_as(C.POINTER(C.c_uint32)), int(overlap), int(lod_count), arr,
                                         len(keys), int(budget_bytes)))

    def UploadBrick(self, lod, index, data):
        """Dataset::GetBrick payload: (z, y, x) array of the stored brick."""
        a = np.ascontiguousarray(data)
        self._chk(self._L.tvk_upload_brick(self._h, int(lod), int(index), a.ctypes.data, a.nbytes))

    def SetPoolBudget(self, budget_bytes):
        """Hold the dataset within budget_bytes of HBM (0: every brick resident)."""
        self._chk(self._L.tvk_set_pool_budget(self._h, int(budget_bytes)))

    def PoolResident(self):
        """Per TOC brick: 1 where the brick is resident (in a slot of a budgeted pool)."""
        n = self.GetDatasetDesc().brick_count
        f = np.zeros(n, np.uint8)
        self._chk(self._L.tvk_pool_resident(self._h, f.ctypes.data_as(C.POINTER(C.c_uint8)), n))
        return f

    def LeapSetRequestTable(self, size=509, rehash=10):
        """GLHashTable size and rehash count of the traversal's brick requests."""
        self._chk(self._L.tvk_leap_set_request_table(self._h, int(size), int(rehash)))

    def LeapRequests(self):
        """The last pass's request table (entry = 4D brick index + 1, 0 = free)."""
        n = C.c_size_t()
        self._chk(self._L.tvk_leap_requests(self._h, None, 0, C.byref(n)))
        t = np.zeros(n.value, np.uint32)
        self._chk(self._L.tvk_leap_requests(self._h, t.ctypes.data_as(C.POINTER(C.c_uint32)), n.value, C.byref(n)))
        return t

    def GetPoolStats(self):
        s = _abi.PoolStats()
        self._chk(self._L.tvk_get_pool_stats(self._h, C.byref(s)))
        return s

    def Rebrick(self, brick, overlap=1, max_lods=0):
        """Re-brick the dataset from its kept linear volume (keep_linear=True)."""
        b = np.array(brick, np.uint32)
        self._chk(self._L.tvk_rebrick(self._h, b.ctypes.data_as(C.POINTER(C.c_uint32)), int(overlap), int(max_lods)))

    def ShareDataset(self, other):
        """Render `other`'s dataset (its HBM brick pool is shared, not copied)."""
        self._chk(self._L.tvk_share_dataset(self._h, other._h))

    def DownloadVolume(self):
        d = self.GetDatasetDesc()
        dt = np.uint8 if d.bit_width == 8 else np.uint16
        out = np.empty((d.dims[2], d.dims[1], d.dims[0]), dt)
        self._chk(self._L.tvk_download_volume(self._h, out.ctypes.data, out.nbytes))
        return out

    def GetDatasetDesc(self):
        d = _abi.DatasetDesc()
        self._chk(self._L.tvk_dataset_info(self._h, C.byref(d)))
        return d

    def GetLODInfo(self, lod):
        size = np.zeros(3, np.uint64)
        bricks = np.zeros(3, np.uint64)
        off = C.c_uint64()
        asp = np.zeros(3, np.float64)
        self._chk(self._L.tvk_lod_info(self._h, lod, size.ctypes.data_as(C.POINTER(C.c_uint64)),
                                       bricks.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(off),
                                       asp.ctypes.data_as(C.POINTER(C.c_double))))
        return dict(size=tuple(int(v) for v in size), bricks=tuple(int(v) for v in bricks), offset=off.value,
                    aspect=tuple(float(v) for v in asp))

    def GetBrickInfo(self, toc):
        c = np.zeros(4, np.uint64)
        s = np.zeros(3, np.uint32)
        off = C.c_uint64()
        self._chk(self._L.tvk_brick_info(self._h, toc, c.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         s.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(off)))
        return dict(coords=tuple(int(v) for v in c), size=tuple(int(v) for v in s), offset=off.value)

    def GetBrickMetadata(self, lod, index):
        c = np.zeros(3, np.float32)
        e = np.zeros(3, np.float32)
        n = np.zeros(3, np.uint32)
        self._chk(self._L.tvk_brick_md(self._h, lod, index, _fp(c), _fp(e), n.ctypes.data_as(C.POINTER(C.c_uint32))))
        return c, e, n

    def MaxMinForKey(self, toc):
        mm = np.zeros(2, np.float64)
        self._chk(self._L.tvk_brick_minmax(self._h, toc, mm.ctypes.data_as(C.POINTER(C.c_double))))
        return float(mm[0]), float(mm[1])

    def GetBrick(self, toc):
        info = self.GetBrickInfo(toc)
        nx, ny, nz = info["size"]
        dt = np.uint8 if self.GetDatasetDesc().bit_width == 8 else np.uint16
        out = np.empty((nz, ny, nx), dt)
        self._chk(self._L.tvk_get_brick(self._h, toc, out.ctypes.data, out.nbytes))
        return out

    def GetPool(self):
        d = self.GetDatasetDesc()
        out = np.empty(d.brick_bytes, np.uint8)
        self._chk(self._L.tvk_get_pool(self._h, out.ctypes.data, out.nbytes))
        return out

    def MaxGradientMagnitude(self):
        v = C.c_float()
        self._chk(self._L.tvk_max_gradient(self._h, C.byref(v)))
        return v.value

    # -- transfer functions ---------------------------------------------------
    def Set1DTrans(self, rgba8):
        t = np.ascontiguousarray(rgba8, np.uint8).reshape(-1, 4)
        self._chk(self._L.tvk_set_tf1d(self._h, t.ctypes.data_as(C.POINTER(C.c_uint8)), t.shape[0]))

    def Set2DTrans(self, rgba8):
        t = np.ascontiguousarray(rgba8, np.uint8)
        h, w = t.shape[:2]
        self._chk(self._L.tvk_set_tf2d(self._h, t.ctypes.data_as(C.POINTER(C.c_uint8)), w, h))

    def Changed2DTrans(self, tf2d):
        """AbstrRenderer::Changed2DTrans: a TransferFunction2D with swatches (and the
        1D TF under them) is painted by the library's rasteriser and set; a plain
        float table is set as its bytes."""
        from .transfer import _swatch_arrays
        if not tf2d.swatches and tf2d.trans1d is None:
            self.Set2DTrans(tf2d.GetByteArray())
            return
        w, h = tf2d.GetSize()
        sw, pa, st, npts, nst = _swatch_arrays(tf2d.swatches)
        t1 = tf2d.trans1d if tf2d.trans1d is not None else np.zeros((1, 4), np.uint8)
        self._chk(self._L.tvk_set_tf2d_swatches(self._h, w, h, t1.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                0 if tf2d.trans1d is None else t1.shape[0], sw, len(tf2d.swatches),
                                                _fp(pa), npts, st, nst))

    # -- state (AbstrRenderer setters) -----------------------------------------
    def _push(self):
        self._chk(self._L.tvk_set_params(self._h, C.byref(self._p)))

    def SetRendermode(self, mode):
        self._p.mode = mode
        self._push()

    def SetUseLighting(self, on):
        self._p.lighting = int(bool(on))
        self._push()

    def SetSampleRateModifier(self, srm):
        self._p.srm = float(srm)
        self._push()

    def SetIsoValue(self, v):
        self._p.isovalue = float(v)
        self._push()

    def SetLOD(self, lod):
        """Fixed LOD (>= 0) or ComputeMinLODForCurrentView (-1)."""
        self._p.lod = int(lod)
        self._push()

    def SetGradientScale(self, s):
        self._p.grad_scale = float(s)
        self._push()

    def SetTFScale(self, s):
        self._p.tf_scale = float(s)
        self._push()

    def SetRefineSteps(self, n):
        self._p.refine_steps = int(n)
        self._push()

    def SetEarlyRayTermination(self, eps):
        self._p.ert_eps = float(eps)
        self._push()

    def SetClampToEdge(self, on):
        self._p.clamp_to_edge = int(bool(on))
        self._push()

    def SetParams(self, **kw):
        for k, v in kw.items():
            f = getattr(self._p, k)
            if hasattr(f, "__len__"):
                f[:] = list(v)
            else:
                setattr(self._p, k, v)
        self._push()

    def GetParams(self):
        return self._p

    def SetCountSamples(self, on):
        self._chk(self._L.tvk_set_count_samples(self._h, int(bool(on))))

    # -- view -----------------------------------------------------------------
    def SetViewParameters(self, fov, znear, zfar):
        self._fov, self._znear, self._zfar = float(fov), float(znear), float(zfar)
        self._camera()

    def SetViewPosition(self, eye, at=(0, 0, 0), up=(0, 1, 0)):
        self._eye = np.asarray(eye, np.float32)
        self._at = np.asarray(at, np.float32)
        self._up = np.asarray(up, np.float32)
        self._camera()

    def Resize(self, w, h):
        self._w, self._h_px = int(w), int(h)
        self._camera()

    def SetRotation(self, rot):
        self._rot = np.ascontiguousarray(rot, np.float32).reshape(4, 4)
        if self._w:
            self._chk(self._L.tvk_set_rotation(self._h, _fp(self._rot)))

    def SetClipPlane(self, plane):
        """AbstrRenderer::SetClipPlane: ExtendedPlane::Plane() in world space (x, y, z, w)."""
        self._clip = np.ascontiguousarray(plane, np.float32).reshape(4)
        self._chk(self._L.tvk_set_clip_plane(self._h, _fp(self._clip), int(getattr(self, "_clip_on", False))))

    def EnableClipPlane(self, on=True):
        """AbstrRenderer::EnableClipPlane / DisableClipPlane."""
        self._clip_on = bool(on)
        self._chk(self._L.tvk_set_clip_plane(self._h, None, int(self._clip_on)))

    def DisableClipPlane(self):
        self.EnableClipPlane(False)

    def CropDataset(self, keep_old=False):
        """AbstrRenderer::CropDataset: zero the voxels on the clipped side of the clip
        plane and brick the volume again; returns the normalised-coordinate plane used."""
        pl = np.zeros(4, np.float32)
        self._chk(self._L.tvk_crop_dataset(self._h, int(bool(keep_old)), _fp(pl)))
        return pl

    def Pick(self, x, y):
        """GLRenderer::Pick: eye-space isosurface hit at window position (x, y from the top)."""
        v = np.zeros(3, np.float32)
        self._chk(self._L.tvk_pick(self._h, int(x), int(y), _fp(v)))
        return v

    def SetUserMatrices(self, view, proj, rot=None, trans=None):
        """AbstrRenderer::SetUserMatrices with explicit row-vector matrices."""
        v = np.ascontiguousarray(view, np.float32)
        p = np.ascontiguousarray(proj, np.float32)
        r = np.ascontiguousarray(np.eye(4) if rot is None else rot, np.float32)
        t = np.ascontiguousarray(np.eye(4) if trans is None else trans, np.float32)
        self._user = (v, p, r, t)
        self._chk(self._L.tvk_set_view_matrices(self._h, _fp(v), _fp(r), _fp(t), _fp(p), self._w, self._h_px,
                                                self._fov, self._znear, self._zfar))

    def _camera(self):
        if not self._w:
            return
        self._chk(self._L.tvk_set_camera(self._h, _fp(self._eye), _fp(self._at), _fp(self._up), self._fov,
                                         self._znear, self._zfar, self._w, self._h_px))
        if not np.array_equal(self._rot, np.eye(4, dtype=np.float32)):
            self._chk(self._L.tvk_set_rotation(self._h, _fp(self._rot)))

    def GetViewMatrices(self):
        v = np.zeros(16, np.float32)
        p = np.zeros(16, np.float32)
        self._chk(self._L.tvk_get_view_matrices(self._h, _fp(v), _fp(p)))
        return v.reshape(4, 4), p.reshape(4, 4)

    # -- planning / rendering -------------------------------------------------------
    def CurrentLOD(self):
        v = C.c_int()
        self._chk(self._L.tvk_current_lod(self._h, C.byref(v)))
        return v.value

    def PlanFrame(self):
        """The visible-brick list of BuildSubFrameBrickList, in compositing order."""
        n = C.c_int()
        self._chk(self._L.tvk_plan(self._h, None, 0, C.byref(n)))
        arr = (_abi.BrickRec * max(n.value, 1))()
        self._chk(self._L.tvk_plan(self._h, arr, n.value, C.byref(n)))
        return [arr[i] for i in range(n.value)]

    def Paint(self, download=True):
        out = np.empty((self._h_px, self._w, 4), np.float32) if download else None
        self._chk(self._L.tvk_render(self._h, _fp(out) if download else None, C.byref(self.last_stats)))
        self._last_views = 1
        return out

    def PaintViews(self, rotations, download=True, stats=True):
        """One launch over every rotation; returns the frames (download) or None.
        With download and stats both False the call returns once the frame is
        queued (last_stats is left as it was): the next call plans while this
        frame runs, and Synchronize() waits for it."""
        rots = np.ascontiguousarray(np.stack([np.asarray(r, np.float32).reshape(4, 4) for r in rotations]))
        n = rots.shape[0]
        out = np.empty((n, self._h_px, self._w, 4), np.float32) if download else None
        self._chk(self._L.tvk_render_views(self._h, n, _fp(rots), _fp(out) if download else None,
                                           C.byref(self.last_stats) if (stats or download) else None))
        self._last_views = n
        return out

    # -- ray-guided traversal (GLGridLeaper) ----------------------------------------
    def LeapSetResident(self, flags=None):
        """Residency per TOC brick (None: every brick resident)."""
        if flags is None:
            self._chk(self._L.tvk_leap_set_resident(self._h, None, 0))
            return
        f = np.ascontiguousarray(flags, np.uint8)
        self._chk(self._L.tvk_leap_set_resident(self._h, f.ctypes.data_as(C.POINTER(C.c_uint8)), f.size))

    def LeapRender(self, first=True):
        """One traversal pass; returns (RGBA32F frame, LeapStats)."""
        out = np.empty((self._h_px, self._w, 4), np.float32)
        st = _abi.LeapStats()
        self._chk(self._L.tvk_leap_render(self._h, int(bool(first)), _fp(out), C.byref(st)))
        self._last_views = 1
        return out, st

    def LeapMissing(self):
        n = self.GetDatasetDesc().brick_count
        f = np.zeros(n, np.uint8)
        cnt = C.c_uint64()
        self._chk(self._L.tvk_leap_missing(self._h, f.ctypes.data_as(C.POINTER(C.c_uint8)), n, C.byref(cnt)))
        return f

    def LeapPageIn(self, max_bricks=(1 << 62)):
        k = C.c_uint64()
        self._chk(self._L.tvk_leap_page_in(self._h, int(max_bricks), C.byref(k)))
        return k.value

    def LeapResumeState(self):
        out = np.empty((self._h_px, self._w, 8), np.float32)
        self._chk(self._L.tvk_leap_resume_state(self._h, _fp(out), out.size))
        return out

    def LeapConverge(self, max_passes=32, page_per_pass=(1 << 62)):
        """Passes until no brick is reported missing (GLGridLeaper's Continue3DDraw
        loop), paging the reported bricks in between.  Returns (frame, [stats])."""
        img, st = self.LeapRender(first=True)
        stats = [st]
        while st.missing and len(stats) < max_passes:
            self.LeapPageIn(page_per_pass)
            img, st = self.LeapRender(first=False)
            stats.append(st)
        return img, stats

    # -- step proposal ------------------------------------------------------------
    def SetStepProposal(self, net, margin=0.0, skip_threshold=1.0, max_log2=3, fallback_bound=0.0, skip_bound=0.0,
                        fixed_log2=-1, fallback_tv=0.0):
        """Attach a StepNet (tuvok_amd.stepnet) to march in chunks with proposed
        step multipliers; net=None restores the fixed step, unless fixed_log2 >= 0
        (the chunked march without a network: multiplier 2^fixed_log2).
        fallback_bound / fallback_tv: a chunk whose contribution bound / variation
        risk exceeds them marches the fixed step (0: never)."""
        if net is None and fixed_log2 < 0:
            self._chk(self._L.tvk_set_step_proposal(self._h, None, None))
            self._net = None
            return
        p = _abi.ProposalParams(float(margin), float(skip_threshold), int(max_log2), float(fallback_bound),
                                float(skip_bound), int(fixed_log2), float(fallback_tv))
        self._chk(self._L.tvk_set_step_proposal(self._h, net._h if net is not None else None, C.byref(p)))
        self._net = net   # keep it alive while attached

    def GenerateSegments(self, rotations, cap, keep=1.0, seed=0, fit=False):
        """Chunk samples of the fixed-step march over the given views (records:
        features, transmittance before, step count, RGBA at x1/x2/x4/x8).  With
        fit=True and more selected chunks than cap, the views are marched again
        with the keep fraction that selects about cap of them (an unbiased
        sample instead of the first cap).  The device appends records in
        whatever order its waves finish; they are returned sorted by a 64-bit
        FNV-1a hash of their 48 words, so the same views give the same records
        in the same order (and training on them is repeatable)."""
        rots = np.ascontiguousarray(np.stack([np.asarray(r, np.float32).reshape(4, 4) for r in rotations]))
        out = np.zeros((int(cap), 48), np.float32)
        n, total = C.c_uint32(), C.c_uint32()
        self._chk(self._L.tvk_generate_segments(self._h, rots.shape[0], _fp(rots), int(cap), float(keep), int(seed),
                                                _fp(out), C.byref(n), C.byref(total)))
        # When more chunks are selected than cap, the device keeps the first cap in
        # the order its waves finish, which varies from run to run.  fit=True
        # lowers the keep fraction until the whole selection fits; without fit
        # the kept set is not repeatable, and that is an error.
        tries = 0
        while fit and total.value > cap and tries < 8:
            keep = float(keep) * 0.97 * cap / total.value
            self._chk(self._L.tvk_generate_segments(self._h, rots.shape[0], _fp(rots), int(cap), keep, int(seed),
                                                    _fp(out), C.byref(n), C.byref(total)))
            tries += 1
        if total.value > cap:
            raise ValueError(f"{total.value} chunks selected for {cap} records: lower keep or pass fit=True "
                             f"(the first {cap} in wave-finish order would not be repeatable)")
        self.last_segments_total = int(total.value)
        out = out[:n.value]
        w = out.view(np.uint32)
        h = np.full(out.shape[0], 0xcbf29ce484222325, np.uint64)
        for k in range(w.shape[1]):
            h ^= w[:, k]
            h *= np.uint64(0x100000001b3)
        return out[np.argsort(h, kind="stable")]

    def Display(self, bg_top=(0.0, 0.0, 0.0), bg_bottom=None):
        """The last rendered frame(s) as the window shows them: over the
        background ramp (flat when bg_bottom is None), 8 bits per channel,
        GL row order; (views, H, W, 4) uint8, or (H, W, 4) for one view."""
        top = np.ascontiguousarray(bg_top, np.float32)
        bot = np.ascontiguousarray(bg_top if bg_bottom is None else bg_bottom, np.float32)
        nv = max(getattr(self, "_last_views", 1), 1)
        out = np.zeros((nv, self._h_px, self._w, 4), np.uint8)
        self._chk(self._L.tvk_display(self._h, _fp(top), _fp(bot), out.ctypes.data_as(C.POINTER(C.c_uint8)),
                                      out.nbytes))
        return out[0] if nv == 1 else out

    def CellGrid(self):
        """The dataset's cell grid (uint32 words (max << 16) | min), TOC brick
        after TOC brick, each x-fastest with ceil(n/8) cells per axis."""
        n = C.c_uint64()
        self._chk(self._L.tvk_cell_grid(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.uint32)
        self._chk(self._L.tvk_cell_grid(self._h, out.ctypes.data_as(C.POINTER(C.c_uint32)), out.size, C.byref(n)))
        return out

    def TouchedBytes(self, unique=False):
        """Stored bytes and count of the bricks the last counted render sampled:
        summed over its views, or (unique=True) over their union."""
        b, n, ub, un = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._chk(self._L.tvk_touched_bytes(self._h, C.byref(b), C.byref(n), C.byref(ub), C.byref(un)))
        return (ub.value, un.value) if unique else (b.value, n.value)

    def TouchedMask(self):
        """Per TOC brick: 1 where the last counted render took a sample."""
        n = self.GetDatasetDesc().brick_count
        m = np.zeros(n, np.uint8)
        self._chk(self._L.tvk_touched_mask(self._h, m.ctypes.data_as(C.POINTER(C.c_uint8)), n))
        return m

    def BrickStoredBytes(self):
        """Per TOC brick: stored voxels (incl. overlap) x bytes per voxel."""
        d = self.GetDatasetDesc()
        out = np.zeros(d.brick_count, np.uint64)
        bpv = d.bit_width // 8
        for i in range(d.brick_count):
            out[i] = int(np.prod(self.GetBrickInfo(i)["size"])) * bpv
        return out

    # -- timing ---------------------------------------------------------------------
    def EventRecord(self, slot):
        self._chk(self._L.tvk_event_record(self._h, slot))

    def EventElapsed(self, a, b):
        ms = C.c_float()
        self._chk(self._L.tvk_event_elapsed(self._h, a, b, C.byref(ms)))
        return ms.value

    def Synchronize(self):
        self._chk(self._L.tvk_synchronize(self._h))
