"""The renderer's compatibility textures on the GPU (AbstrRenderer's
bUseOnlyPowerOfTwo, bDownSampleTo8Bits and bDisableBorder): frames and sample
counts equal the oracle's, a downsampled 16-bit dataset renders bit for bit as
the 8-bit dataset of its quantized volume, and the paths that read the pool's
own bricks refuse the flags."""
import numpy as np
import pytest

import oracle_ffi as O
from helpers import err_stats, oracle_camera_from, oracle_params_from
from test_oracle_compat import quantize_x86
from tuvok_amd import (GEN_BLOBS, RM_1DTRANS, RM_2DTRANS, RM_ISOSURFACE, HIPRaycaster, TvkError, rotation_y,
                       tf1d_blue_white_red, tf2d_gaussian_lobes)

pytestmark = pytest.mark.gpu

TOL_MAX = 1.0 / 255.0
TOL_MEAN = 1e-3


def setup(ren, dtype, n=48, brick=18, w=72, h=64):
    ren.GenerateVolume(GEN_BLOBS, 2, dtype, (n, n, n), (brick,) * 3, 1, max_lods=1, keep_linear=True)
    vol = ren.DownloadVolume()
    tf = tf1d_blue_white_red().GetByteArray()
    tf2 = tf2d_gaussian_lobes().GetByteArray()
    ren.Set1DTrans(tf)
    ren.Set2DTrans(tf2)
    ren.SetViewParameters(45.0, 0.01, 1000.0)
    ren.SetViewPosition((0.15, -0.1, 1.9))
    ren.Resize(w, h)
    ren.SetGradientScale(0.75)
    ren.SetEarlyRayTermination(0.0)
    return vol, tf, tf2


def check(ren, ds, tf, tf2, rot, label):
    cam = oracle_camera_from(ren, rot)
    ref, ns = O.render(ds, cam, oracle_params_from(ren, tf1d=tf, tf2d=tf2))
    ren.SetRotation(rot)
    ren.SetCountSamples(True)
    img = ren.Paint()
    mx, mean = err_stats(img, ref)
    print(f"{label}: max {mx:.3g} mean {mean:.3g} samples oracle {ns} gpu {ren.last_stats.samples}", flush=True)
    assert ref[..., 3].max() > 0.05
    assert mx <= TOL_MAX and mean < TOL_MEAN
    assert ren.last_stats.samples == ns
    return img


@pytest.mark.parametrize("dtype,border", [(np.uint8, False), (np.uint16, False), (np.uint8, True)])
@pytest.mark.parametrize("mode,lit", [(RM_1DTRANS, True), (RM_2DTRANS, False), (RM_ISOSURFACE, False)])
def test_power_of_two_textures_match_oracle(mode, lit, dtype, border):
    print(f"pow2 mode={mode} lit={lit} dtype={np.dtype(dtype).name} disable_border={border}", flush=True)
    ren = HIPRaycaster(0, use_only_power_of_two=True, disable_border=border)
    plain = HIPRaycaster(0, disable_border=border)
    try:
        vol, tf, tf2 = setup(ren, dtype)
        setup(plain, dtype)
        for r in (ren, plain):
            r.SetRendermode(mode)
            r.SetUseLighting(lit)
            r.SetIsoValue(0.45 * float(vol.max()))
        ds = O.Dataset(vol, (18, 18, 18), 1)
        rot = rotation_y(0.6)
        img = check(ren, ds, tf, tf2, rot, "pow2")
        plain.SetRotation(rot)
        plain.SetCountSamples(True)
        ref_plain = plain.Paint()
        assert ren.last_stats.samples > plain.last_stats.samples     # texels of the 32^3 texture
        assert not np.array_equal(img, ref_plain)
    finally:
        ren.close()
        plain.close()


@pytest.mark.parametrize("mode,lit", [(RM_1DTRANS, False), (RM_1DTRANS, True), (RM_2DTRANS, True),
                                      (RM_ISOSURFACE, False)])
def test_downsampled_textures_match_oracle(mode, lit):
    print(f"down8 mode={mode} lit={lit}", flush=True)
    ren = HIPRaycaster(0, downsample_to_8bits=True)
    try:
        vol, tf, tf2 = setup(ren, np.uint16)
        ren.SetRendermode(mode)
        ren.SetUseLighting(lit)
        ren.SetIsoValue(0.45 * float(vol.max()))
        check(ren, O.Dataset(vol, (18, 18, 18), 1), tf, tf2, rotation_y(-0.4), "down8")
    finally:
        ren.close()


def test_power_of_two_and_downsampled_together_match_oracle():
    ren = HIPRaycaster(0, use_only_power_of_two=True, downsample_to_8bits=True)
    try:
        vol, tf, tf2 = setup(ren, np.uint16)
        ren.SetRendermode(RM_1DTRANS)
        ren.SetUseLighting(True)
        check(ren, O.Dataset(vol, (18, 18, 18), 1), tf, tf2, rotation_y(0.3), "pow2+down8")
    finally:
        ren.close()


def test_downsampled_frames_equal_the_quantized_volume():
    """LOD 0 holds the volume's voxels, so quantizing bricks equals quantizing the
    volume: the 16-bit dataset with 8-bit textures renders exactly as the 8-bit
    dataset of (255 v - fMin) / (fMax - fMin) (range [0, 65535]: scale 1)."""
    n = 48
    g = np.random.default_rng(7)
    vol = np.clip(g.normal(30000, 15000, (n, n, n)), 0, 65535).astype(np.uint16)
    vol[0, 0, 0], vol[-1, -1, -1] = 0, 65535
    q = quantize_x86(vol, 0.0, 65535.0)
    ren = HIPRaycaster(0, downsample_to_8bits=True)
    plain = HIPRaycaster(0)
    try:
        tf = tf1d_blue_white_red().GetByteArray()
        tf2 = tf2d_gaussian_lobes().GetByteArray()
        for r, v in ((ren, vol), (plain, q)):
            r.LoadVolume(v, (18, 18, 18), 1, max_lods=1)
            r.Set1DTrans(tf)
            r.Set2DTrans(tf2)
            r.SetViewParameters(45.0, 0.01, 1000.0)
            r.SetViewPosition((0.1, 0.2, 1.9))
            r.Resize(64, 56)
            r.SetGradientScale(0.5)
            r.SetRotation(rotation_y(0.8))
        for mode, lit in ((RM_1DTRANS, True), (RM_2DTRANS, True)):
            for r in (ren, plain):
                r.SetRendermode(mode)
                r.SetUseLighting(lit)
            a, b = ren.Paint(), plain.Paint()
            assert a[..., 3].max() > 0.05
            assert np.array_equal(a, b), (mode, lit)
    finally:
        ren.close()
        plain.close()


def test_paths_that_read_the_pool_refuse_the_flags():
    ren = HIPRaycaster(0, downsample_to_8bits=True)
    try:
        setup(ren, np.uint16)
        ren.SetRendermode(RM_1DTRANS)
        with pytest.raises(TvkError):
            ren.LeapRender(first=True)
        with pytest.raises(TvkError):
            ren.SetPoolBudget(1 << 20)
            ren.Paint()
    finally:
        ren.close()
