"""The UVF/ExtendedOctree reader feeding the brick pool: files laid out as the
reference's writers lay them out (uvf_file.py) from the oracle's conversion --
uncompressed and zlib bricks, bricks stored out of TOC order, a block to step
over before the TOC block, a big-endian container header, a non-unit aspect --
load into a pool byte-equal to the oracle's, with its min/max per brick, and
render the oracle's frames, resident and under an HBM budget."""
import numpy as np
import pytest

import oracle_ffi as O
import uvf_file as U
from helpers import err_stats, oracle_camera_from, oracle_params_from
from tuvok_amd import RM_1DTRANS, RM_2DTRANS, HIPRaycaster, TvkError, rotation_y, tf1d_blue_white_red
from tuvok_amd.transfer import tf2d_gaussian_lobes

pytestmark = pytest.mark.gpu


def volume(dt, dims=(40, 36, 44), seed=3):
    g = np.random.default_rng(seed)
    full = 255 if dt == np.uint8 else 65535
    z, y, x = np.meshgrid(*[np.linspace(-1, 1, n) for n in dims[::-1]], indexing="ij")
    smooth = np.exp(-3 * (x * x + y * y + z * z))
    v = full * smooth + g.integers(0, full // 16, dims[::-1])
    return np.clip(v, 0, full).astype(dt)


def render_pair(ren, ds, mode, lit):
    tf = tf1d_blue_white_red().GetByteArray()
    tf2 = tf2d_gaussian_lobes().GetByteArray()
    ren.Set1DTrans(tf)
    ren.Set2DTrans(tf2)
    ren.SetRendermode(mode)
    ren.SetUseLighting(lit)
    ren.SetGradientScale(0.8)
    ren.SetViewParameters(45.0, 0.01, 1000.0)
    ren.SetViewPosition((0.1, 0.15, 1.9))
    ren.Resize(64, 56)
    rot = rotation_y(0.5)
    ren.SetRotation(rot)
    ref, ns = O.render(ds, oracle_camera_from(ren, rot), oracle_params_from(ren, tf1d=tf, tf2d=tf2))
    ren.SetCountSamples(True)
    img = ren.Paint()
    return img, ref, ns


@pytest.mark.parametrize("dt,big_endian,aspect", [(np.uint8, False, (1.0, 1.0, 1.0)),
                                                  (np.uint16, True, (1.0, 1.0, 1.0)),
                                                  (np.uint16, False, (1.0, 2.0, 0.5))])
def test_uvf_file_loads_the_oracle_pool_and_frames(tmp_path, dt, big_endian, aspect):
    print(f"uvf dtype={np.dtype(dt).name} big_endian={big_endian} aspect={aspect}", flush=True)
    vol = volume(dt)
    ds = O.Dataset(vol, (12, 12, 12), 1, scale=aspect)
    n = ds.brick_count
    path = tmp_path / "vol.uvf"
    U.write_uvf(path, ds, big_endian=big_endian, aspect=aspect,
                codec=lambda i: U.CT_ZLIB if i % 3 else U.CT_NONE, order=list(range(n))[::-1])
    ren = HIPRaycaster(0)
    try:
        info = ren.OpenUVF(path)
        assert info.brick_count == n and info.lod_count == ds.lod_count
        assert info.zlib_bricks == sum(1 for i in range(n) if i % 3)
        assert tuple(info.dims) == ds.dims and tuple(info.aspect) == aspect
        assert np.array_equal(ren.GetPool(), ds.all_data())
        for i in range(n):
            assert ren.MaxMinForKey(i) == ds.brick_info(i)["minmax"]
        for mode, lit in ((RM_1DTRANS, True), (RM_2DTRANS, False)):
            img, ref, ns = render_pair(ren, ds, mode, lit)
            mx, mean = err_stats(img, ref)
            print(f"  mode={mode} lit={lit}: max {mx:.3g} mean {mean:.3g} samples {ren.last_stats.samples}/{ns}",
                  flush=True)
            assert ref[..., 3].max() > 0.05
            assert mx <= 1.0 / 255 and mean < 1e-3 and ren.last_stats.samples == ns
    finally:
        ren.close()


def test_uvf_file_under_a_budget_renders_the_resident_frames(tmp_path):
    vol = volume(np.uint16, seed=5)
    ds = O.Dataset(vol, (12, 12, 12), 1)
    path = tmp_path / "vol.uvf"
    U.write_uvf(path, ds, codec=lambda i: U.CT_ZLIB)
    full, part = HIPRaycaster(0), HIPRaycaster(0)
    try:
        full.OpenUVF(path)
        slot = 12 * 12 * 12 * 2 * 2 + 4096
        part.OpenUVF(path, budget_bytes=8 * slot)
        assert part.GetPoolStats().slots < ds.brick_count
        for mode, lit in ((RM_1DTRANS, True), (RM_2DTRANS, True)):
            a, _, _ = render_pair(full, ds, mode, lit)
            b, _, _ = render_pair(part, ds, mode, lit)
            assert a[..., 3].max() > 0.05
            assert np.array_equal(a, b)
    finally:
        full.close()
        part.close()


def test_uvf_files_the_reader_refuses(tmp_path):
    ds = O.Dataset(volume(np.uint8, dims=(20, 20, 20)), (12, 12, 12), 1)
    ren = HIPRaycaster(0)
    try:
        cases = {
            "lzma": dict(codec=lambda i: U.CT_LZMA if i == 1 else U.CT_NONE),
            "atlas": dict(atlas=lambda i: (64, 64) if i == 0 else (0, 0)),
        }
        for name, kw in cases.items():
            p = tmp_path / f"{name}.uvf"
            U.write_uvf(p, ds, **kw)
            with pytest.raises(TvkError):
                ren.OpenUVF(p)
        good = tmp_path / "good.uvf"
        size = U.write_uvf(good, ds)
        cut = tmp_path / "cut.uvf"
        cut.write_bytes(good.read_bytes()[:size - 100])
        with pytest.raises(TvkError):
            ren.OpenUVF(cut)
        bad = tmp_path / "bad.uvf"
        bad.write_bytes(b"NOT-UVF!" + good.read_bytes()[8:])
        with pytest.raises(TvkError):
            ren.OpenUVF(bad)
        ren.OpenUVF(good)                       # the context still works
        assert np.array_equal(ren.GetPool(), ds.all_data())
    finally:
        ren.close()
