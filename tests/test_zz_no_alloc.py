"""tvk_render_views_noalloc (the allocation-free render) against tvk_render_views.

Called through the C ABI with a device output buffer from hipMalloc; two calls
in a row, the second with fewer views and a smaller image, each equal to the
allocating entry point's frames bit for bit (same kernels, same inputs), and
the same composited-sample counts.  Runs alone in a fresh process and after
the other GPU tests.
"""
import ctypes as C

import numpy as np
import pytest

from tuvok_amd import _abi
from tuvok_amd import GEN_BLOBS, RM_1DTRANS, RM_2DTRANS, HIPRaycaster, rotation_y
from tuvok_amd.transfer import tf1d_blue_white_red, tf2d_gaussian_lobes

pytestmark = pytest.mark.gpu

TVK_ERR_ARG = -1


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _noalloc(ren, rots, w, h, dbuf, cap_bytes):
    L = ren._L
    rots = np.ascontiguousarray(np.stack([np.asarray(r, np.float32).reshape(4, 4) for r in rots]))
    n = rots.shape[0]
    st = _abi.Stats()
    rc = L.tvk_render_views_noalloc(ren._h, n, rots.ctypes.data_as(C.POINTER(C.c_float)), dbuf, cap_bytes,
                                    C.byref(st))
    return rc, st, n


def _download(hip, dbuf, n, w, h):
    out = np.empty((n, h, w, 4), np.float32)
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(out.ctypes.data, dbuf, out.nbytes, 2) == 0   # hipMemcpyDeviceToHost
    return out


@pytest.mark.parametrize("mode", [RM_1DTRANS, RM_2DTRANS])
def test_noalloc_equals_render_views(mode):
    print(f"[noalloc] mode {mode}: 128^3 uint8 blobs, 34^3 bricks", flush=True)
    hip = _hip()
    ren = HIPRaycaster(0)
    try:
        ren.GenerateVolume(GEN_BLOBS, 3, np.uint8, (128, 128, 128), (34, 34, 34), 1, max_lods=1, keep_linear=True)
        ren.SetGradientScale(1.0 / ren.MaxGradientMagnitude())
        ren.Set1DTrans(tf1d_blue_white_red().GetByteArray())
        ren.Set2DTrans(tf2d_gaussian_lobes().GetByteArray())
        ren.SetRendermode(mode)
        ren.SetUseLighting(True)
        ren.SetViewParameters(45.0, 0.01, 1000.0)
        ren.SetViewPosition((0.0, 0.0, 2.0))
        ren.SetCountSamples(True)
        cap = 4 * 96 * 80 * 16
        dbuf = C.c_void_p()
        assert hip.hipMalloc(C.byref(dbuf), cap) == 0
        try:
            for n, (w, h) in ((4, (96, 80)), (2, (48, 40))):
                ren.Resize(w, h)
                rots = [rotation_y(0.4 + 1.3 * k) for k in range(n)]
                want = ren.PaintViews(rots)
                want_samples = ren.last_stats.samples
                rc, st, _ = _noalloc(ren, rots, w, h, dbuf, cap)
                assert rc == 0, ren._L.tvk_last_error(ren._h).decode()
                got = _download(hip, dbuf, n, w, h)
                assert want[..., 3].max() > 0.05, "empty frames"
                assert np.array_equal(got, want), float(np.abs(got - want).max())
                assert st.samples == want_samples and st.samples > 0
                assert st.bricks_nonempty > 0
            # a too small output buffer and too many views are refused with an error
            ren.Resize(96, 80)
            rc, _, _ = _noalloc(ren, [np.eye(4)], 96, 80, dbuf, 96 * 80 * 16 - 4)
            assert rc == TVK_ERR_ARG
            ren.Resize(4, 4)
            rc, _, _ = _noalloc(ren, [np.eye(4)] * 282, 4, 4, dbuf, cap)
            assert rc == TVK_ERR_ARG
            assert b"limit" in ren._L.tvk_last_error(ren._h)
        finally:
            hip.hipFree(dbuf)
    finally:
        ren.close()
