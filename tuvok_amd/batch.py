"""RenderBatch: many render requests in one persistent-kernel frame.

A request is an HIPRaycaster (its dataset, transfer function, parameters and
camera) plus a view rotation and an image size.  ``Render()`` plans every
request, uploads the brick lists and the request table, and replays the frame
as a hipGraph: one persistent launch per kernel variant whose waves pull ray
tiles from per-XCD device queues.  Each request's frame equals
``HIPRaycaster.Paint()`` of its renderer at that rotation and size.

    b = RenderBatch()
    i = b.Add(ren_a, rotation_y(0.3), 1024, 768)
    j = b.Add(ren_b, None, 256, 256)
    b.Autotune()             # picks tile shape and grid size on these requests
    b.Render()
    img = b.Frame(i)         # RGBA32F, premultiplied, GL row order
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import TvkError


class RenderBatch:
    def __init__(self, device=0):
        self._L = _abi.load()
        h = C.c_void_p()
        rc = self._L.tvk_batch_create(C.byref(h), device)
        if rc != 0:
            raise TvkError(rc, "tvk_batch_create failed (no GPU?)")
        self._h = h
        self._sizes = []
        self._keep = []     # renderers referenced by the requests stay alive with the batch
        self.last_stats = _abi.Stats()

    def _chk(self, rc):
        if rc != 0:
            raise TvkError(rc, self._L.tvk_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.tvk_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def Clear(self):
        self._chk(self._L.tvk_batch_clear(self._h))
        self._sizes, self._keep = [], []

    def Add(self, renderer, rotation, width, height):
        rot = None
        if rotation is not None:
            rot = np.ascontiguousarray(np.asarray(rotation, np.float32).reshape(4, 4))
        i = C.c_int()
        self._chk(self._L.tvk_batch_add(self._h, renderer._h,
                                        rot.ctypes.data_as(C.POINTER(C.c_float)) if rot is not None else None,
                                        int(width), int(height), C.byref(i)))
        self._sizes.append((int(width), int(height)))
        self._keep.append(renderer)
        return i.value

    def SetTuning(self, tile_w=8, tile_h=8, groups_per_cu=0, use_graph=True):
        t = _abi.Tuning(tile_w, tile_h, groups_per_cu, int(bool(use_graph)))
        self._chk(self._L.tvk_batch_set_tuning(self._h, C.byref(t)))

    def GetTuning(self):
        t = _abi.Tuning()
        self._chk(self._L.tvk_batch_get_tuning(self._h, C.byref(t)))
        return dict(tile_w=t.tile_w, tile_h=t.tile_h, groups_per_cu=t.groups_per_cu, use_graph=bool(t.use_graph))

    def SetStepProposal(self, net, margin=0.0, skip_threshold=1.0, max_log2=3, fallback_bound=0.0, skip_bound=0.0,
                        fallback_tv=0.0):
        """Learned step for the 1D/2D-TF requests (tuvok_amd.stepnet.StepNet); None: fixed step."""
        if net is None:
            self._chk(self._L.tvk_batch_set_step_proposal(self._h, None, None))
            self._net = None
            return
        p = _abi.ProposalParams(float(margin), float(skip_threshold), int(max_log2), float(fallback_bound),
                                float(skip_bound), -1, float(fallback_tv))
        self._chk(self._L.tvk_batch_set_step_proposal(self._h, net._h, C.byref(p)))
        self._net = net

    def SetCountSamples(self, on):
        self._chk(self._L.tvk_batch_set_count_samples(self._h, int(bool(on))))

    def Render(self):
        self._chk(self._L.tvk_batch_render(self._h, C.byref(self.last_stats)))
        return self.last_stats

    def Frame(self, i):
        w, h = self._sizes[i]
        out = np.empty((h, w, 4), np.float32)
        self._chk(self._L.tvk_batch_frame(self._h, int(i), out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def Autotune(self, reps=2):
        t = _abi.Tuning()
        ms = C.c_float()
        self._chk(self._L.tvk_batch_autotune(self._h, int(reps), C.byref(t), C.byref(ms)))
        return dict(tile_w=t.tile_w, tile_h=t.tile_h, groups_per_cu=t.groups_per_cu,
                    use_graph=bool(t.use_graph)), ms.value
