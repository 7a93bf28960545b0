"""Pins the oracle's compatibility textures (AbstrRenderer's bDownSampleTo8Bits
and bUseOnlyPowerOfTwo) on the CPU.

  * 8-bit downsampling quantizes each brick at upload
    ((unsigned char)((255 v - fMin) / (fMax - fMin)), GPUMemManDataStructs.cpp:
    410-425) and renders with a TF scale of 1 (GLRenderer.cpp:1813-1819): at LOD 0,
    where the bricks hold the volume's own voxels, a 16-bit dataset rendered this
    way is bit for bit the 8-bit dataset of the quantized volume rendered plainly
    -- including the wrap of out-of-range results (fMin > 0), whose low byte x86
    keeps;
  * power-of-two padding: bricks whose stored size already is a power of two
    render exactly as without the flag (the texture coordinates and voxel step
    reduce to the plain ones); other bricks move into a larger texture, and an
    isosurface -- fixed by the data, not by the step -- lands on the same pixels
    with nearly the same shading.
"""
import numpy as np

import oracle_ffi as O
from helpers import blobs_volume, err_stats
from tuvok_amd.transfer import tf1d_blue_white_red, tf2d_gaussian_lobes


def params(mode=0, lighting=False, tf1d=None, tf2d=None, bit_width=8, **kw):
    p = dict(mode=mode, lighting=lighting, bit_width=bit_width, tf1d=tf1d, tf2d=tf2d, tf_scale=1.0)
    p.update(kw)
    return p


def quantize_x86(v, fmin, fmax):
    """The reference's (unsigned char)((255.0*v - fMin) / (fMax - fMin)) as x86 evaluates it."""
    q = (255.0 * v.astype(np.float64) - fmin) / (fmax - fmin)
    return (np.trunc(q).astype(np.int64) & 0xFF).astype(np.uint8)


def sphere_u8(n=48, rad=0.3):
    g = (np.arange(n) + 0.5) / n - 0.5
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    return np.clip(128 + 20 * (rad - np.sqrt(x * x + y * y + z * z)) * n, 0, 255).astype(np.uint8)


def test_downsample_renders_the_quantized_volume():
    vol = blobs_volume(48, seed=3, dtype=np.uint16)
    vol[0, 0, 0], vol[-1, -1, -1] = 0, 65535          # range [0, 65535]: q(0) = 0, scale 1
    q = quantize_x86(vol, 0.0, 65535.0)
    d16 = O.Dataset(vol, (18, 18, 18), 1, lod0_only=True)
    d8 = O.Dataset(q, (18, 18, 18), 1, lod0_only=True)
    tf = tf1d_blue_white_red().GetByteArray()
    tf2 = tf2d_gaussian_lobes().GetByteArray()
    cam = O.camera(56, 48, eye=(0.2, 0.1, 1.9), rot=O.rotation_y(0.4))
    for mode, lit in ((0, True), (1, False), (1, True)):
        a, na = O.render(d16, cam, params(mode, lit, tf, tf2, bit_width=16, down8=True, grad_scale=0.5))
        b, nb = O.render(d8, cam, params(mode, lit, tf, tf2, bit_width=8, grad_scale=0.5))
        assert a[..., 3].max() > 0.1
        assert np.array_equal(a, b) and na == nb, (mode, lit)
        c, _ = O.render(d16, cam, params(mode, lit, tf, tf2, bit_width=16, grad_scale=0.5,
                                         tf_scale=float(np.float32(65535 / 65535))))
        assert not np.array_equal(a, c)     # the 16-bit textures interpolate finer values


def test_downsample_wraps_like_x86_when_the_range_starts_above_zero():
    vol = blobs_volume(40, seed=5, dtype=np.uint16)
    vol = (vol // 2 + 20000).astype(np.uint16)          # fMin > 0: low values quantize below 0
    fmin, fmax = float(vol.min()), float(vol.max())
    q = quantize_x86(vol, fmin, fmax)
    assert (q > 200).sum() > 0 and ((255.0 * vol.astype(np.float64) - fmin) / (fmax - fmin) > 256).any()
    # clamp-to-edge bricking: overlap voxels are copies, so they quantize alike
    d16 = O.Dataset(vol, (18, 18, 18), 1, clamp=True, lod0_only=True)
    d8 = O.Dataset(q, (18, 18, 18), 1, clamp=True, lod0_only=True)
    tf = tf1d_blue_white_red().GetByteArray()
    cam = O.camera(48, 40, eye=(0.0, 0.0, 1.9))
    a, na = O.render(d16, cam, params(0, False, tf, None, bit_width=16, down8=True))
    b, nb = O.render(d8, cam, params(0, False, tf, None, bit_width=8))
    assert np.array_equal(a, b) and na == nb


def test_downsample_isovalue_normalisation():
    """GetNormalizedIsovalue with 8-bit textures: iso / range max (not / 2^16)."""
    vol = sphere_u8(48).astype(np.uint16) * 200          # range [0, 51000]
    d16 = O.Dataset(vol, (18, 18, 18), 1, lod0_only=True)
    d8 = O.Dataset(quantize_x86(vol, 0.0, float(vol.max())), (18, 18, 18), 1, lod0_only=True)
    cam = O.camera(48, 48, eye=(0.0, 0.0, 1.8))
    # 16-bit iso 25500 -> 25500 / 51000 = 0.5; the 8-bit dataset: iso 128 -> 128 / 256 = 0.5
    a, _ = O.render(d16, cam, params(2, False, None, None, bit_width=16, isovalue=25500.0, down8=True,
                                     refine_steps=5))
    b, _ = O.render(d8, cam, params(2, False, None, None, bit_width=8, isovalue=128.0, refine_steps=5))
    assert (a[..., 3] > 0).sum() > 100
    assert np.array_equal(a, b)


def test_pow2_bricks_already_powers_of_two_render_unchanged():
    vol = blobs_volume(60, seed=1)
    ds = O.Dataset(vol, (32, 32, 32), 1, lod0_only=True)       # stored 32 = 2^5 (cores 30)
    tf = tf1d_blue_white_red().GetByteArray()
    cam = O.camera(48, 40, eye=(0.1, 0.0, 1.9))
    a, na = O.render(ds, cam, params(0, True, tf, pow2=True))
    b, nb = O.render(ds, cam, params(0, True, tf))
    assert np.array_equal(a, b) and na == nb


def test_pow2_padding_keeps_the_isosurface():
    vol = sphere_u8(48)
    ds = O.Dataset(vol, (18, 18, 18), 1, lod0_only=True)      # stored 18 -> textures of 32
    cam = O.camera(64, 64, eye=(0.15, 0.1, 1.8))
    for border in (False, True):
        a, na = O.render(ds, cam, params(2, False, None, None, isovalue=128.0, refine_steps=5, pow2=True,
                                         clamp_to_edge=border))
        b, nb = O.render(ds, cam, params(2, False, None, None, isovalue=128.0, refine_steps=5,
                                         clamp_to_edge=border))
        hit_a, hit_b = a[..., 3] > 0, b[..., 3] > 0
        assert hit_b.sum() > 500
        assert (hit_a != hit_b).sum() <= 0.01 * hit_b.sum()
        both = hit_a & hit_b
        mx, mean = err_stats(a[both], b[both])
        assert mean < 0.02, mean
        assert na > nb          # the padded texture's texel is smaller: more, shorter steps
        assert not np.array_equal(a, b)
