"""Transfer functions with the byte semantics of the reference.

TransferFunction1D / TransferFunction2D keep float RGBA tables and hand the
renderer RGBA8 tables exactly as TransferFunction1D::GetByteArray
(IO/TransferFunction1D.cpp:311-330, truncating '(uchar)(clamp(x,0,1)*255)')
and TransferFunction2D::GetByteArray (IO/TransferFunction2D.cpp:222-264).
Non-zero limits follow ComputeNonZeroLimits (TransferFunction1D.cpp:362-371,
TransferFunction2D.cpp:378-397).
"""
import numpy as np


def _to_bytes(rgba, used_range=255):
    # (unsigned char)(max(0, min(v, 1)) * cUsedRange), evaluated in float32
    v = np.clip(np.asarray(rgba, np.float32), np.float32(0), np.float32(1)) * np.float32(used_range)
    return v.astype(np.uint8)


class TransferFunction1D:
    def __init__(self, size=256):
        self.color = np.zeros((size, 4), np.float32)

    def GetSize(self):
        return self.color.shape[0]

    def SetStdFunction(self, center=0.3, inv_gradient=0.3):
        """TransferFunction1D::SetStdFunction (TransferFunction1D.cpp:81-113): smoothstep ramp."""
        n = self.GetSize()
        ic = int((n - 1) * float(min(max(0.0, center), 1.0)))
        ig = int((n - 1) * float(min(max(0.0, inv_gradient), 1.0)))
        rs = 0 if ig // 2 > ic else ic - ig // 2
        re = n if ig // 2 + ic > n else ic + ig // 2
        for comp in range(4):
            self.color[:rs, comp] = 0
            for i in range(rs, re):
                t = np.float32(i - ic + ig // 2) / np.float32(ig)
                self.color[i, comp] = t * t * (3 - 2 * t)   # MathTools smoothstep
            self.color[re:, comp] = 1
        return self

    def Set(self, rgba8):
        """TransferFunction1D::Set(rgba8) (TransferFunction1D.cpp:146-177)."""
        b = np.asarray(rgba8, np.uint8).reshape(-1, 4)
        lo = b.min(axis=0).astype(np.float32)
        hi = b.max(axis=0).astype(np.float32)
        out = np.zeros(b.shape, np.float32)
        for c in range(3):
            span = hi[c] - lo[c]
            out[:, c] = 0 if span == 0 else (b[:, c].astype(np.float32) - lo[c]) / span
        out[:, 3] = b[:, 3].astype(np.float32) / np.float32(255)
        self.color = out
        return self

    def GetByteArray(self, used_range=255):
        return _to_bytes(self.color, used_range)

    def GetNonZeroLimits(self):
        nz = np.nonzero(self.color[:, 3] != 0)[0]
        if len(nz) == 0:
            return (self.GetSize(), 0)
        return (int(nz[0]), int(nz[-1]))


class TFPolygon:
    """A swatch of the 2D TF (IO/TransferFunction2D.h, TFPolygon): a polygon in
    normalised table coordinates filled with a linear (point 0 -> point 1) or
    radial (centre point 0, radius |point 1 - point 0|) gradient; stops are
    (position, (r, g, b, a)) with floats in 0..1."""

    def __init__(self, points, gradient=((0.0, 0.5), (1.0, 0.5)), stops=((0.0, (0, 0, 0, 1)), (1.0, (1, 1, 1, 1))),
                 radial=False):
        self.pPoints = [tuple(map(float, p)) for p in points]
        self.pGradientCoords = [tuple(map(float, gradient[0])), tuple(map(float, gradient[1]))]
        self.pGradientStops = [(float(t), tuple(map(float, c))) for t, c in stops]
        self.bRadial = bool(radial)


def _swatch_arrays(swatches):
    """The C ABI's swatch, point and stop arrays of a list of TFPolygon."""
    import ctypes as C
    from . import _abi
    sw = (_abi.Swatch * max(1, len(swatches)))()
    pts, stops = [], []
    for i, s in enumerate(swatches):
        sw[i].radial = int(s.bRadial)
        sw[i].first_point, sw[i].n_points = len(pts) // 2, len(s.pPoints)
        for p in s.pPoints:
            pts.extend(p)
        g = list(s.pGradientCoords[0]) + list(s.pGradientCoords[1])
        for k in range(4):
            sw[i].gradient[k] = g[k]
        sw[i].first_stop, sw[i].n_stops = len(stops), len(s.pGradientStops)
        stops.extend(s.pGradientStops)
    pa = np.ascontiguousarray(pts if pts else [0.0], np.float32)
    st = (_abi.GradientStop * max(1, len(stops)))()
    for i, (t, c) in enumerate(stops):
        st[i].pos = t
        for k in range(4):
            st[i].rgba[k] = c[k]
    return sw, pa, st, len(pts) // 2, len(stops)


class TransferFunction2D:
    """2D TF over (value, 1 - |gradient| * scale); rows index the gradient axis.

    Either a float table (`color`) or, once swatches or a 1D TF are set, the
    table painted from them (TransferFunction2D::RenderTransferFunction8Bit,
    rasterised by libtvkhip.so's tvk_rasterize_tf2d without Qt)."""

    def __init__(self, width=256, height=256):
        self.color = np.zeros((height, width, 4), np.float32)
        self.swatches = []
        self.trans1d = None   # RGBA8 bytes of the 1D TF painted under the swatches

    def SwatchPushBack(self, swatch):
        self.swatches.append(swatch)

    def Update1DTrans(self, tf1d):
        """TransferFunction2D::Update1DTrans: the 1D TF (TransferFunction1D or RGBA8 bytes)."""
        b = tf1d.GetByteArray() if hasattr(tf1d, "GetByteArray") else tf1d
        self.trans1d = np.ascontiguousarray(b, np.uint8).reshape(-1, 4)

    def GetSize(self):
        return (self.color.shape[1], self.color.shape[0])

    def GetByteArray(self, used_range=255):
        if not self.swatches and self.trans1d is None:
            return _to_bytes(self.color, used_range)
        import ctypes as C
        from . import _abi
        L = _abi.load()
        w, h = self.GetSize()
        out = np.zeros((h, w, 4), np.uint8)
        sw, pa, st, npts, nst = _swatch_arrays(self.swatches)
        t1 = self.trans1d if self.trans1d is not None else np.zeros((1, 4), np.uint8)
        rc = L.tvk_rasterize_tf2d(w, h, t1.ctypes.data_as(C.POINTER(C.c_uint8)),
                                  0 if self.trans1d is None else t1.shape[0], sw, len(self.swatches),
                                  pa.ctypes.data_as(C.POINTER(C.c_float)), npts, st, nst,
                                  out.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != 0:
            raise _abi.TvkError(rc, "tvk_rasterize_tf2d failed")
        if used_range != 255:   # GetByteArray(pcData, cUsedRange): byte * 255 / range, truncated
            out = (out.astype(np.float32) * np.float32(255.0 / used_range)).astype(np.uint8)
        return out

    def GetNonZeroLimits(self):
        a = self.GetByteArray()[:, :, 3]
        ys, xs = np.nonzero(a)
        h, w = a.shape
        if len(xs) == 0:
            return (w, 0, h, 0)
        return (int(xs.min()), int(xs.max()), int(ys.min()), int(ys.max()))


def tf1d_blue_white_red(size=256, alpha_start=40, alpha_max=0.6):
    """1D TF: opacity 0 below `alpha_start`, linear to alpha_max at size-1;
    colour is a 3-stop blue-white-red map."""
    tf = TransferFunction1D(size)
    i = np.arange(size, dtype=np.float64)
    t = i / (size - 1)
    blue = np.array([0.0, 0.0, 1.0])
    white = np.array([1.0, 1.0, 1.0])
    red = np.array([1.0, 0.0, 0.0])
    col = np.where((t < 0.5)[:, None], blue + (white - blue) * (2 * t)[:, None],
                   white + (red - white) * (2 * t - 1)[:, None])
    alpha = np.where(i < alpha_start, 0.0, alpha_max * (i - alpha_start) / (size - 1 - alpha_start))
    tf.color[:, :3] = col.astype(np.float32)
    tf.color[:, 3] = alpha.astype(np.float32)
    return tf


def tf1d_random(seed, size=256):
    """A random 1D TF: opacity 0 below a random threshold, then a linear ramp to a
    random peak; colour interpolated between two random stops."""
    rng = np.random.default_rng(seed)
    lo = int(rng.integers(20, 160))
    amax = float(rng.uniform(0.2, 0.9))
    c0, c1 = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3)
    tf = TransferFunction1D(size)
    i = np.arange(size, dtype=np.float64)
    t = (i / (size - 1))[:, None]
    tf.color[:, :3] = (c0 + (c1 - c0) * t).astype(np.float32)
    tf.color[:, 3] = np.where(i < lo, 0.0, amax * (i - lo) / (size - 1 - lo)).astype(np.float32)
    return tf


def tf2d_gaussian_lobes(width=256, height=256, lobes=None):
    """2D TF made of Gaussian lobes in (value, gradient-row) space.

    Each lobe: (value centre, gradient centre, value sigma, gradient sigma,
    peak alpha, rgb), coordinates in [0,1] of the table axes.  Row y of the
    table is looked up with 1 - |g| * gradScale, so y = 1 is zero gradient.
    """
    if lobes is None:
        lobes = [
            (0.30, 0.95, 0.05, 0.05, 0.05, (0.9, 0.7, 0.5)),
            (0.55, 0.85, 0.06, 0.10, 0.10, (0.3, 0.6, 1.0)),
            (0.75, 0.70, 0.05, 0.15, 0.30, (1.0, 0.3, 0.2)),
            (0.915, 0.90, 0.02, 0.12, 0.90, (1.0, 1.0, 0.8)),
        ]
    tf = TransferFunction2D(width, height)
    x = (np.arange(width) + 0.5) / width
    y = (np.arange(height) + 0.5) / height
    X, Y = np.meshgrid(x, y)
    rgb = np.zeros((height, width, 3))
    a = np.zeros((height, width))
    for (vc, gc, vs, gs, amax, col) in lobes:
        w = amax * np.exp(-0.5 * (((X - vc) / vs) ** 2 + ((Y - gc) / gs) ** 2))
        rgb += w[..., None] * np.asarray(col)
        a += w
    rgb = np.where(a[..., None] > 0, rgb / np.maximum(a[..., None], 1e-12), 0)
    tf.color[..., :3] = rgb.astype(np.float32)
    tf.color[..., 3] = np.clip(a, 0, 1).astype(np.float32)
    return tf


def tf1d_phantom(size=256, lo=0.08):
    """1D TF for the nested-ellipsoid phantom (intensities 0-4000 plus noise):
    transparent background below `lo` of the range, then faint, rising opacity
    over the soft classes and denser opacity over the bright ones; colour goes
    from dark red through orange to near white."""
    tf = TransferFunction1D(size)
    t = np.arange(size, dtype=np.float64) / (size - 1)
    u = np.clip((t - lo) / (1.0 - lo), 0.0, 1.0)
    alpha = np.where(t < lo, 0.0, 0.01 + 0.05 * u + 0.25 * u ** 4)
    tf.color[:, 0] = (0.5 + 0.5 * u).astype(np.float32)
    tf.color[:, 1] = (0.15 + 0.8 * u).astype(np.float32)
    tf.color[:, 2] = (0.1 + 0.8 * u ** 2).astype(np.float32)
    tf.color[:, 3] = alpha.astype(np.float32)
    return tf
