"""Shared helpers of the parity tests."""
import numpy as np

import oracle_ffi as O


def blobs_volume(n=64, seed=0, nblobs=12, dtype=np.uint8):
    """CPU version of the Gaussian-blob workload (for oracle-only tests)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.2, 0.8, (nblobs, 3))
    s = rng.uniform(0.03, 0.12, nblobs)
    a = rng.uniform(0.3, 1.0, nblobs)
    g = (np.arange(n) + 0.5) / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    f = np.zeros((n, n, n))
    for k in range(nblobs):
        f += a[k] * np.exp(-((x - c[k, 0]) ** 2 + (y - c[k, 1]) ** 2 + (z - c[k, 2]) ** 2) / (2 * s[k] ** 2))
    full = 255 if dtype == np.uint8 else 65535
    v = np.rint(full * f / f.max()) + rng.integers(-4, 5, f.shape) * (full // 255)
    return np.clip(v, 0, full).astype(dtype)


def oracle_camera_from(ren, rot=None):
    """Camera dict for the oracle holding exactly the product's matrices."""
    view, proj = ren.GetViewMatrices()
    return dict(view=view, proj=proj, rot=np.eye(4, dtype=np.float32) if rot is None else np.asarray(rot, np.float32),
                trans=np.eye(4, dtype=np.float32), width=ren._w, height=ren._h_px, fov=ren._fov,
                znear=ren._znear, zfar=ren._zfar)


def product_tf_scale(ren):
    """fTransScale as GLRenderer::CalculateScaling computes it: (2^bits-1)/range max,
    1 for 16-bit data downsampled to 8-bit textures."""
    d = ren.GetDatasetDesc()
    if getattr(ren, "_down8", False) and d.bit_width != 8:
        return 1.0
    return float(np.float32(((1 << d.bit_width) - 1) / d.range_max))


def oracle_params_from(ren, tf1d=None, tf2d=None, bit_width=None, tf_scale=None):
    p = ren.GetParams()
    if bit_width is None:
        bit_width = ren.GetDatasetDesc().bit_width
    if tf_scale is None:
        tf_scale = product_tf_scale(ren)
    return dict(mode=p.mode, lighting=p.lighting, srm=p.srm, isovalue=p.isovalue, lod=p.lod, bit_width=bit_width,
                tf1d=tf1d, tf2d=tf2d, tf_scale=tf_scale, grad_scale=p.grad_scale,
                domain_scale=tuple(p.domain_scale), light_ambient=tuple(p.light_ambient),
                light_diffuse=tuple(p.light_diffuse), light_specular=tuple(p.light_specular),
                light_dir=tuple(p.light_dir), iso_color=tuple(p.iso_color), refine_steps=p.refine_steps,
                clamp_to_edge=p.clamp_to_edge, ert_eps=p.ert_eps,
                clip=getattr(ren, "_clip", None) if getattr(ren, "_clip_on", False) else None,
                pow2=getattr(ren, "_pow2", False), down8=getattr(ren, "_down8", False))


def err_stats(a, b):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    return float(d.max()), float(d.mean())
