"""Parity of the HIP path (through the C ABI) with the oracle.

Bit-exact: brick payloads, overlap voxels, per-brick min/max, brick table,
visible-brick list.  Rendered RGBA: max |diff| <= 1/255 per channel and
mean < 1e-3 (the HIP march repeats the oracle's float operations, so the
observed difference is ~1e-7; the tolerance is the acceptance gate).
"""
import numpy as np
import pytest

import oracle_ffi as O
from helpers import blobs_volume, err_stats, oracle_camera_from, oracle_params_from
from tuvok_amd import (DISPATCH_GROUPS, DISPATCH_QUEUES, DISPATCH_RASTER, DISPATCH_ROWS, GEN_BLOBS, GEN_FBM_SHELLS,
                       RM_1DTRANS,
                       RM_2DTRANS, RM_ISOSURFACE, HIPRaycaster, TvkError, rotation_y, tf1d_blue_white_red,
                       tf2d_gaussian_lobes)

pytestmark = pytest.mark.gpu

TOL_MAX = 1.0 / 255.0
TOL_MEAN = 1e-3


@pytest.fixture(scope="module")
def ren():
    r = HIPRaycaster(0)
    yield r
    r.close()


def lod0_bricks(ds):
    return [i for i in range(ds.brick_count) if ds.brick_info(i)["coords"][3] == 0]


BRICK_CASES = [
    ((64, 64, 64), (18, 18, 18), 1, False, np.uint8),
    ((37, 21, 50), (12, 10, 16), 1, False, np.uint16),
    ((40, 36, 44), (12, 10, 10), 2, True, np.uint16),
    ((33, 5, 1), (12, 6, 6), 1, False, np.uint8),
    # rows longer than one gather chunk (64 blocks), the middle brick inside the volume
    ((600, 12, 10), (258, 6, 6), 1, False, np.uint16),
    ((1100, 8, 6), (514, 6, 6), 1, True, np.uint8),
    ((300, 19, 15), (66, 10, 10), 2, False, np.uint16),
]


@pytest.mark.parametrize("dims,brick,ov,clamp,dt", BRICK_CASES)
def test_device_bricking_bit_exact(ren, dims, brick, ov, clamp, dt):
    print(f"bricking dims={dims} brick={brick} ov={ov} clamp={clamp} dtype={np.dtype(dt).name}", flush=True)
    rng = np.random.default_rng(11)
    full = 255 if dt == np.uint8 else 65535
    vol = rng.integers(0, full + 1, size=dims[::-1]).astype(dt)
    ds = O.Dataset(vol, brick, ov, clamp=clamp)
    ren.LoadVolume(vol, brick, ov, clamp_to_edge=clamp, max_lods=1)
    d = ren.GetDatasetDesc()
    assert d.brick_count == len(lod0_bricks(ds))
    pool = ren.GetPool()
    ref = ds.all_data()
    n0 = sum(int(np.prod(ds.brick_info(i)["size"])) for i in lod0_bricks(ds)) * vol.itemsize
    assert pool.nbytes == n0
    assert np.array_equal(pool, ref[:n0])
    for i in lod0_bricks(ds):
        assert ren.MaxMinForKey(i) == ds.brick_info(i)["minmax"]
        assert ren.GetBrickInfo(i)["size"] == ds.brick_info(i)["size"]
        c = ds.brick_info(i)["coords"]
        bc = ren.GetLODInfo(0)["bricks"]
        lin = c[0] + c[1] * bc[0] + c[2] * bc[0] * bc[1]
        gc, ge, gn = ren.GetBrickMetadata(0, lin)
        oc, oe, on = ds.brick_md(0, lin)
        assert np.array_equal(gc, oc) and np.array_equal(ge, oe) and np.array_equal(gn, on)
    assert d.range_max == ds.range_max()


def test_device_bricking_thin_cores():
    """Cores thinner than the overlap (5-voxel bricks, overlap 2): every brick is
    the zero-padded volume cut at coords * core - overlap.  The reference's
    converter fills some of those overlap voxels from other voxels (its copy
    order, not reproduced), so the oracle is compared on the cores only."""
    dims, brick, ov = (12, 9, 7), (6, 5, 5), 2
    vol = np.random.default_rng(12).integers(1, 256, size=dims[::-1]).astype(np.uint8)
    ds = O.Dataset(vol, brick, ov)
    r = HIPRaycaster(0)
    try:
        r.LoadVolume(vol, brick, ov, max_lods=1)
        pool = r.GetPool()
    finally:
        r.close()
    core = [b - 2 * ov for b in brick]
    pad = np.zeros([d + 2 * ov + b for d, b in zip(dims[::-1], brick[::-1])], np.uint8)
    pad[ov:ov + dims[2], ov:ov + dims[1], ov:ov + dims[0]] = vol
    for i in lod0_bricks(ds):
        info = ds.brick_info(i)
        (nx, ny, nz), c, o = info["size"], info["coords"], info["offset"]
        g = pool[o:o + nx * ny * nz].reshape(nz, ny, nx)
        want = pad[c[2] * core[2]:c[2] * core[2] + nz, c[1] * core[1]:c[1] * core[1] + ny,
                   c[0] * core[0]:c[0] * core[0] + nx]
        assert np.array_equal(g, want), info
        ref = np.asarray(ds.brick_data(i)).reshape(nz, ny, nx)
        assert np.array_equal(g[ov:nz - ov, ov:ny - ov, ov:nx - ov], ref[ov:nz - ov, ov:ny - ov, ov:nx - ov]), info


@pytest.mark.parametrize("dims,brick,dt", [((44, 36, 40), (10, 10, 10), np.uint8),
                                           ((70, 50, 90), (12, 12, 16), np.uint16)])
def test_device_lod_pyramid_bit_exact(ren, dims, brick, dt):
    """All levels, overlap voxels included (FillOverlap's copy order reproduced)."""
    print(f"lod pyramid dims={dims} brick={brick}", flush=True)
    full = 255 if dt == np.uint8 else 65535
    vol = np.random.default_rng(5).integers(0, full + 1, size=dims[::-1]).astype(dt)
    ds = O.Dataset(vol, brick, 1)
    ren.LoadVolume(vol, brick, 1, max_lods=0)
    pool = ren.GetPool()
    ref = ds.all_data()
    assert pool.nbytes == ref.nbytes
    assert np.array_equal(pool, ref)
    for i in range(ds.brick_count):
        assert ren.MaxMinForKey(i) == ds.brick_info(i)["minmax"]


def test_device_lod_pyramid_clamped_cores(ren):
    """With clamp-to-edge the coarse levels' cores match; a few overlap corners of the
    reference hold stale buffer contents (not reproduced)."""
    vol = np.random.default_rng(6).integers(0, 256, size=(44, 36, 40)).astype(np.uint8)
    ds = O.Dataset(vol, (10, 10, 10), 1, clamp=True)
    ren.LoadVolume(vol, (10, 10, 10), 1, clamp_to_edge=True, max_lods=0)
    pool = ren.GetPool()
    ref = ds.all_data()
    for i in range(ds.brick_count):
        info = ds.brick_info(i)
        nx, ny, nz = info["size"]
        g = pool[info["offset"]:info["offset"] + nx * ny * nz].reshape(nz, ny, nx)
        r = ref[info["offset"]:info["offset"] + nx * ny * nz].reshape(nz, ny, nx)
        assert np.array_equal(g[1:-1, 1:-1, 1:-1], r[1:-1, 1:-1, 1:-1]), info
        if info["coords"][3] == 0:
            assert np.array_equal(g, r), info


def setup_blobs(ren, n=64, brick=18, w=96, h=80, eye=(0.2, -0.1, 2.0), dtype=np.uint8):
    ren.GenerateVolume(GEN_BLOBS, 0, dtype, (n, n, n), (brick,) * 3, 1, max_lods=1, keep_linear=True)
    vol = ren.DownloadVolume()
    tf = tf1d_blue_white_red().GetByteArray()
    ren.Set1DTrans(tf)
    ren.SetViewParameters(45.0, 0.01, 1000.0)
    ren.SetViewPosition(eye)
    ren.Resize(w, h)
    return vol, tf


def compare(ren, ds, tf1d=None, tf2d=None, rot=None):
    cam = oracle_camera_from(ren, rot)
    p = oracle_params_from(ren, tf1d=tf1d, tf2d=tf2d)
    ref, ns = O.render(ds, cam, p)
    ren.SetCountSamples(True)
    if rot is not None:
        ren.SetRotation(rot)
    img = ren.Paint()
    mx, mean = err_stats(img, ref)
    return img, ref, mx, mean, ns, ren.last_stats.samples


def test_plan_matches_oracle(ren):
    vol, tf = setup_blobs(ren)
    ren.SetRendermode(RM_1DTRANS)
    ds = O.Dataset(vol, (18, 18, 18), 1)
    for rot in (None, rotation_y(0.7)):
        if rot is not None:
            ren.SetRotation(rot)
        got = ren.PlanFrame()
        cam = oracle_camera_from(ren, rot)
        want = O.plan(ds, cam, oracle_params_from(ren, tf1d=tf))
        assert len(got) == len(want)
        for g, w in zip(got, want):
            assert (g.key_index, g.toc_index, g.is_empty) == (w.key_index, w.toc_index, w.is_empty)
            assert g.distance == w.distance
            assert list(g.center) == list(w.center) and list(g.extension) == list(w.extension)
            assert list(g.tex_min) == list(w.tex_min) and list(g.tex_max) == list(w.tex_max)
    ren.SetRotation(np.eye(4, dtype=np.float32))


def test_plan_of_a_culling_view_matches_oracle(ren):
    """A close, off-axis view whose frustum cuts the volume: the culled list (fewer
    bricks than the level holds) equals the oracle's item by item, order included."""
    vol, tf = setup_blobs(ren)
    ren.SetRendermode(RM_1DTRANS)
    ds = O.Dataset(vol, (18, 18, 18), 1)
    nlod0 = len(lod0_bricks(ds))
    seen_culled = 0
    for eye, at in (((0.35, 0.25, 0.45), (0.3, 0.2, 0.0)), ((-0.6, 0.1, 0.3), (-0.2, -0.3, 0.1)),
                    ((0.1, 0.7, 0.2), (0.25, 0.0, -0.3))):
        ren.SetViewPosition(eye, at=at)
        got = ren.PlanFrame()
        want = O.plan(ds, oracle_camera_from(ren), oracle_params_from(ren, tf1d=tf))
        assert len(got) == len(want)
        seen_culled += nlod0 - len(got)
        for g, w in zip(got, want):
            assert (g.key_index, g.toc_index, g.is_empty, g.distance) == (w.key_index, w.toc_index, w.is_empty,
                                                                         w.distance)
            assert list(g.center) == list(w.center) and list(g.tex_min) == list(w.tex_min)
    assert seen_culled > 0
    ren.SetViewPosition((0.2, -0.1, 2.0))


def test_current_lod_matches_oracle_over_distances(ren):
    """ComputeMinLODForCurrentView of the product (tvk_current_lod with lod = -1)
    equals the oracle's as the eye moves from 1.5 to 12 volume widths, at two
    viewport heights and screen-space errors; several levels are chosen."""
    dims = (256, 256, 256)
    ren.LoadVolume(np.zeros(dims[::-1], np.uint8), (66, 66, 66), 1, max_lods=0)
    ds = O.Dataset(np.zeros(dims[::-1], np.uint8), (66, 66, 66), 1, lod0_only=True)
    ren.Set1DTrans(tf1d_blue_white_red().GetByteArray())
    ren.SetViewParameters(45.0, 0.01, 1000.0)
    ren.SetLOD(-1)
    seen = set()
    try:
        for sse in (1.0, 2.5):
            ren.SetParams(screen_space_error=sse)
            for dist in np.linspace(1.5, 12.0, 22):
                for h in (270, 1080):
                    ren.Resize(h * 16 // 9, h)
                    ren.SetViewPosition((0.0, 0.0, float(dist)))
                    got = ren.CurrentLOD()
                    assert got == O.min_lod(ds, oracle_camera_from(ren), sse), (sse, dist, h)
                    seen.add(got)
    finally:
        ren.SetLOD(0)
        ren.SetParams(screen_space_error=1.0)
    assert len(seen) >= 3


@pytest.mark.parametrize("mode,lit", [(RM_1DTRANS, False), (RM_1DTRANS, True), (RM_2DTRANS, False),
                                      (RM_2DTRANS, True), (RM_ISOSURFACE, False)])
def test_render_modes_match_oracle(ren, mode, lit):
    print(f"render mode={mode} lit={lit}", flush=True)
    vol, tf = setup_blobs(ren)
    ds = O.Dataset(vol, (18, 18, 18), 1)
    tf2 = tf2d_gaussian_lobes().GetByteArray()
    ren.Set2DTrans(tf2)
    ren.SetRendermode(mode)
    ren.SetUseLighting(lit)
    ren.SetGradientScale(1.0 / max(ren.MaxGradientMagnitude(), 1e-6))
    ren.SetIsoValue(128.0)
    ren.SetEarlyRayTermination(0.0)
    img, ref, mx, mean, ns_ref, ns_gpu = compare(ren, ds, tf1d=tf, tf2d=tf2)
    print(f"  max {mx:.3g} mean {mean:.3g} samples oracle {ns_ref} gpu {ns_gpu}", flush=True)
    assert ref[..., 3].max() > 0.05
    assert mx <= TOL_MAX and mean < TOL_MEAN
    assert ns_gpu == ns_ref


@pytest.mark.parametrize("mode,lit", [(RM_1DTRANS, False), (RM_1DTRANS, True), (RM_ISOSURFACE, False)])
def test_uint16_modes_match_oracle(ren, mode, lit):
    """uint16 volumes through the 1D (unlit: the 2x2x2 fast form with 4-byte
    x-pair loads) and isosurface paths: frames and sample counts equal the
    oracle's."""
    print(f"uint16 mode={mode} lit={lit}", flush=True)
    vol, tf = setup_blobs(ren, dtype=np.uint16)
    ds = O.Dataset(vol, (18, 18, 18), 1)
    ren.SetRendermode(mode)
    ren.SetUseLighting(lit)
    ren.SetIsoValue(0.4 * float(vol.max()))
    ren.SetEarlyRayTermination(0.0)
    img, ref, mx, mean, ns_ref, ns_gpu = compare(ren, ds, tf1d=tf)
    print(f"  max {mx:.3g} mean {mean:.3g} samples oracle {ns_ref} gpu {ns_gpu}", flush=True)
    assert ref[..., 3].max() > 0.05
    assert mx <= TOL_MAX and mean < TOL_MEAN
    assert ns_gpu == ns_ref


@pytest.mark.parametrize("mode,lit,srm,lod", [(RM_1DTRANS, True, 2.0, 0), (RM_2DTRANS, True, 2.0, 0),
                                               (RM_1DTRANS, False, 0.5, 0), (RM_1DTRANS, True, 4.0, 0),
                                               (RM_1DTRANS, True, 0.7, 0), (RM_1DTRANS, True, 1.0, 1),
                                               (RM_2DTRANS, True, 1.0, 2), (RM_2DTRANS, False, 2.0, 1)])
def test_sample_rate_and_lod_match_oracle(ren, mode, lit, srm, lod):
    """Opacity correction 1 - (1 - a)^s with s = (1/srm) x the level's domain
    ratio (0.25 .. 4 here, and 1/0.7): frames and sample counts equal the
    oracle's at other sample rates and at coarser levels."""
    print(f"srm={srm} lod={lod} mode={mode} lit={lit}", flush=True)
    n = 64
    ren.GenerateVolume(GEN_BLOBS, 0, np.uint8, (n, n, n), (18, 18, 18), 1, max_lods=0, keep_linear=True)
    vol = ren.DownloadVolume()
    tf = tf1d_blue_white_red().GetByteArray()
    ren.Set1DTrans(tf)
    tf2 = tf2d_gaussian_lobes().GetByteArray()
    ren.Set2DTrans(tf2)
    ren.SetViewParameters(45.0, 0.01, 1000.0)
    ren.SetViewPosition((0.2, -0.1, 2.0))
    ren.Resize(96, 80)
    ren.SetRendermode(mode)
    ren.SetUseLighting(lit)
    ren.SetGradientScale(1.0 / max(ren.MaxGradientMagnitude(), 1e-6))
    ren.SetEarlyRayTermination(0.0)
    ren.SetSampleRateModifier(srm)
    ren.SetLOD(lod)
    ds = O.Dataset(vol, (18, 18, 18), 1)
    img, ref, mx, mean, ns_ref, ns_gpu = compare(ren, ds, tf1d=tf, tf2d=tf2)
    print(f"  max {mx:.3g} mean {mean:.3g} samples oracle {ns_ref} gpu {ns_gpu}", flush=True)
    ren.SetSampleRateModifier(1.0)
    ren.SetLOD(0)
    assert ref[..., 3].max() > 0.05
    assert mx <= TOL_MAX and mean < TOL_MEAN
    assert ns_gpu == ns_ref


def test_blobs_256_full_frame(ren):
    """The 256^3 uint8 1D-TF workload at 512x512, whole frame against the oracle."""
    print("blobs 256^3, 512x512", flush=True)
    vol, tf = setup_blobs(ren, n=256, brick=66, w=512, h=512, eye=(0.0, 0.0, 2.0))
    ren.SetRendermode(RM_1DTRANS)
    ren.SetUseLighting(False)
    ren.SetEarlyRayTermination(0.0)
    ds = O.Dataset(vol, (66, 66, 66), 1)
    img, ref, mx, mean, ns_ref, ns_gpu = compare(ren, ds, tf1d=tf)
    print(f"  max {mx:.3g} mean {mean:.3g} samples oracle {ns_ref} gpu {ns_gpu}", flush=True)
    assert mx <= TOL_MAX and mean < TOL_MEAN
    assert ns_gpu == ns_ref
    fp64 = O.render_fp64(ds, oracle_camera_from(ren), oracle_params_from(ren, tf1d=tf), 64 / 66 / 256 * 0.5)
    m64, a64 = err_stats(img, fp64)
    print(f"  vs FP64 integral: max {m64:.3g} mean {a64:.3g}", flush=True)


def test_early_termination_bound(ren):
    vol, tf = setup_blobs(ren, eye=(0.0, 0.0, 1.7))
    ren.SetRendermode(RM_1DTRANS)
    ren.SetUseLighting(False)
    ds = O.Dataset(vol, (18, 18, 18), 1)
    ren.SetEarlyRayTermination(0.0)
    exact = ren.Paint()
    for eps in (1 / 2048, 1 / 512):
        ren.SetEarlyRayTermination(eps)
        cut = ren.Paint()
        assert np.abs(cut - exact).max() <= eps + 1e-6
        cam = oracle_camera_from(ren)
        ref, _ = O.render(ds, cam, oracle_params_from(ren, tf1d=tf))
        assert np.abs(cut - ref).max() <= 1e-5
    ren.SetEarlyRayTermination(0.0)


def test_views_batch_equals_single(ren):
    setup_blobs(ren)
    ren.SetRendermode(RM_1DTRANS)
    ren.SetUseLighting(True)
    rots = [rotation_y(a) for a in (0.0, 0.9, 2.1, 4.0)]
    batch = ren.PaintViews(rots)
    for k, r in enumerate(rots):
        ren.SetRotation(r)
        assert np.array_equal(ren.Paint(), batch[k])
    ren.SetRotation(np.eye(4, dtype=np.float32))


@pytest.mark.parametrize("mode,lit,learned", [(RM_1DTRANS, True, False), (RM_2DTRANS, False, False),
                                              (RM_ISOSURFACE, True, False), (RM_1DTRANS, True, True)])
def test_dispatch_orders_give_the_same_frames(ren, mode, lit, learned):
    """Every dispatch setting (raster, cost-ordered groups or rows, each with
    one workgroup per group or with per-XCD work queues) renders the same
    frames bit for bit with the same sample count: views of a size with
    partial 16x16 groups, the fixed march and the chunked march with its exact
    skip."""
    print(f"dispatch orders mode={mode} lit={lit} learned={learned}", flush=True)
    setup_blobs(ren, w=200, h=136)
    ren.Set2DTrans(tf2d_gaussian_lobes().GetByteArray())
    ren.SetIsoValue(128.0)
    ren.SetRendermode(mode)
    ren.SetUseLighting(lit)
    if learned:
        ren.SetStepProposal(None, fixed_log2=0, skip_bound=0.0)
    rots = [rotation_y(a) for a in (0.0, 0.9, 2.1, 4.0, 5.5)]
    settings = [o | q for q in (0, DISPATCH_QUEUES) for o in (DISPATCH_RASTER, DISPATCH_GROUPS, DISPATCH_ROWS)]
    ren.SetCountSamples(True)
    got = {}
    try:
        for o in settings:
            ren.SetDispatchOrder(o)
            got[o] = (ren.PaintViews(rots), ren.last_stats.samples)
        for bad in (3, 7, 8, -1):
            with pytest.raises(TvkError):
                ren.SetDispatchOrder(bad)
    finally:
        ren.SetDispatchOrder(DISPATCH_RASTER | DISPATCH_QUEUES)
        ren.SetCountSamples(False)
        ren.SetStepProposal(None)
    ref, ns = got[DISPATCH_RASTER]
    assert ref.max() > 0.05 and ns > 0
    for o in settings[1:]:
        assert np.array_equal(got[o][0], ref) and got[o][1] == ns, o
    if not learned:
        # the tuner keeps one of the settings, and the frames stay the same
        from tuvok_amd.autotune import DISPATCH_SETTINGS, tune_dispatch_order
        best, times = tune_dispatch_order(ren, rots, reps=1)
        try:
            assert best in DISPATCH_SETTINGS and set(times) == set(DISPATCH_SETTINGS)
            assert np.array_equal(ren.PaintViews(rots), ref)
        finally:
            ren.SetDispatchOrder(DISPATCH_RASTER | DISPATCH_QUEUES)


def test_isosurface_early_termination_skips_bricks_behind_the_hit(ren):
    """With early termination on, the isosurface march skips the bricks whose
    segment starts beyond the ray's hit so far: the same frames and hit
    targets (Pick) as the exact mode, fewer samples."""
    from tuvok_amd import GEN_ELLIPSOID_SDF
    ren.GenerateVolume(GEN_ELLIPSOID_SDF, 2, np.uint8, (96, 96, 96), (18, 18, 18), 1, max_lods=1)
    ren.SetRendermode(RM_ISOSURFACE)
    ren.SetUseLighting(True)
    ren.SetIsoValue(128.0)
    ren.SetViewParameters(45.0, 0.01, 1000.0)
    ren.SetViewPosition((0.1, -0.2, 1.8))
    ren.Resize(120, 96)
    rots = [rotation_y(a) for a in (0.0, 0.7, 2.5, 4.4)]
    ren.SetCountSamples(True)
    try:
        got = {}
        for eps in (0.0, 2.0 ** -11):
            ren.SetEarlyRayTermination(eps)
            frames = ren.PaintViews(rots)
            ns = ren.last_stats.samples
            ren.SetRotation(rots[1])
            ren.Paint()
            picks = []
            for (x, y) in ((60, 48), (30, 40), (90, 70)):
                try:
                    picks.append(ren.Pick(x, y).tolist())
                except TvkError:   # no hit at that pixel
                    picks.append(None)
            got[eps] = (frames, ns, picks)
    finally:
        ren.SetCountSamples(False)
        ren.SetEarlyRayTermination(0.0)
        ren.SetRotation(np.eye(4, dtype=np.float32))
    (f0, n0, p0), (f1, n1, p1) = got[0.0], got[2.0 ** -11]
    print(f"iso samples exact {n0}, with the skip {n1}", flush=True)
    assert f0[..., 3].max() > 0.5
    assert np.array_equal(f0, f1) and p0 == p1
    assert n1 < n0


def test_fbm_shells_u16_lit_2d_crop(ren):
    """uint16 fBm+shells, lit 2D TF, 2 views; a crop of each frame against the oracle."""
    n = 128
    ren.GenerateVolume(GEN_FBM_SHELLS, 1, np.uint16, (n, n, n), (34, 34, 34), 1, max_lods=1, keep_linear=True)
    vol = ren.DownloadVolume()
    tf2 = tf2d_gaussian_lobes().GetByteArray()
    ren.Set2DTrans(tf2)
    ren.SetRendermode(RM_2DTRANS)
    ren.SetUseLighting(True)
    g = ren.MaxGradientMagnitude()
    ren.SetGradientScale(1.0 / g)
    ren.SetEarlyRayTermination(0.0)
    ren.SetViewParameters(45.0, 0.01, 1000.0)
    ren.SetViewPosition((0.0, 0.0, 2.0))
    ren.Resize(192, 108)
    ds = O.Dataset(vol, (34, 34, 34), 1)
    for rot in (rotation_y(0.0), rotation_y(1.3)):
        img, ref, mx, mean, ns_ref, ns_gpu = compare(ren, ds, tf2d=tf2, rot=rot)
        print(f"  fbm rot: max {mx:.3g} mean {mean:.3g} samples {ns_ref}/{ns_gpu}", flush=True)
        assert ref[..., 3].max() > 0.05
        assert mx <= TOL_MAX and mean < TOL_MEAN
    ren.SetRotation(np.eye(4, dtype=np.float32))


def test_full_size_bricking_properties(ren):
    """1024^3 uint16 with 128^3 cores: every brick's core equals the linear volume,
    overlaps equal the neighbours (zero outside), min/max match the payload."""
    n, b = 1024, 130
    print("full-size bricking 1024^3 u16", flush=True)
    ren.GenerateVolume(GEN_FBM_SHELLS, 1, np.uint16, (n, n, n), (b, b, b), 1, max_lods=1, keep_linear=True)
    vol = ren.DownloadVolume()
    pool = ren.GetPool().view(np.uint16)
    pad = np.zeros((n + 2, n + 2, n + 2), np.uint16)
    pad[1:-1, 1:-1, 1:-1] = vol
    off = 0
    for z in range(8):
        for y in range(8):
            for x in range(8):
                br = pool[off:off + b ** 3].reshape(b, b, b)
                want = pad[z * 128:z * 128 + b, y * 128:y * 128 + b, x * 128:x * 128 + b]
                assert np.array_equal(br, want), (x, y, z)
                toc = x + 8 * y + 64 * z
                assert ren.MaxMinForKey(toc) == (float(want.min()), float(want.max()))
                off += b ** 3
    assert off == pool.size


def test_display_over_background_closed_form():
    """CopyImageToDisplayBuffer over ClearColorBuffer / DrawBackGradient: the
    premultiplied frame blended with (SRC_ALPHA, ONE_MINUS_SRC_ALPHA) over a
    vertical ramp into 8 bits, recomputed in float32 with numpy."""
    from tuvok_amd import workloads
    r = workloads.setup(workloads.BLOBS_256, width=96, height=64)
    rots = [rotation_y(0.2), rotation_y(1.3)]
    frames = r.PaintViews(rots)
    top = np.array([0.2, 0.3, 0.9], np.float32)
    bot = np.array([0.9, 0.6, 0.1], np.float32)
    got = r.Display(top, bot)
    H, W = frames.shape[1:3]
    t = ((np.arange(H, dtype=np.float32) + np.float32(0.5)) / np.float32(H))[None, :, None, None]
    bg = bot + t * (top - bot)
    a = frames[..., 3:4]
    rgb = frames[..., :3] * a + bg * (np.float32(1) - a)
    q = lambda x: np.floor(np.clip(x, 0, 1).astype(np.float32) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    want = np.concatenate([q(rgb), q(a * a)], axis=-1)
    assert got.shape == want.shape and np.array_equal(got, want)
    assert (frames[..., 3] > 0.1).any() and (frames[..., 3] == 0).any()
    # one colour: a flat clear
    flat = r.Display((0.5, 0.5, 0.5))
    assert np.array_equal(flat[..., :3][frames[..., 3] == 0], np.full(((frames[..., 3] == 0).sum(), 3), 128, np.uint8))
    r.close()


def test_swatch_2d_tf_renders_like_its_table(ren):
    """Changed2DTrans with swatches (rasterised by the library) renders exactly as
    the same table set as bytes, and as the oracle with that table."""
    from tuvok_amd import TFPolygon, TransferFunction2D
    vol, _ = setup_blobs(ren)
    t = TransferFunction2D(256, 256)
    t.Update1DTrans(tf1d_blue_white_red())
    t.SwatchPushBack(TFPolygon([(0.45, 0.0), (0.95, 0.0), (0.95, 1.0), (0.45, 1.0)], gradient=((0.45, 0.5), (0.95, 0.5)),
                               stops=((0.0, (0.2, 0.4, 1.0, 0.1)), (1.0, (1.0, 0.9, 0.2, 0.8)))))
    t.SwatchPushBack(TFPolygon([(0.6, 0.2), (0.8, 0.5), (0.6, 0.8)], gradient=((0.7, 0.5), (0.8, 0.5)),
                               stops=((0.0, (1, 0, 0, 1)), (1.0, (1, 0, 0, 0))), radial=True))
    table = t.GetByteArray()
    ren.SetRendermode(RM_2DTRANS)
    ren.SetUseLighting(True)
    ren.SetGradientScale(1.0 / ren.MaxGradientMagnitude())
    ren.Changed2DTrans(t)
    a = ren.Paint()
    ren.Set2DTrans(table)
    b = ren.Paint()
    assert np.array_equal(a, b) and a[..., 3].max() > 0.05
    ds = O.Dataset(vol, (18, 18, 18), 1)
    ref, _ = O.render(ds, oracle_camera_from(ren), oracle_params_from(ren, tf2d=table))
    mx, mean = err_stats(a, ref)
    assert mx <= TOL_MAX and mean < TOL_MEAN


@pytest.mark.parametrize("mode,lit", [(RM_1DTRANS, True), (RM_2DTRANS, False), (RM_ISOSURFACE, False)])
def test_clip_plane_matches_oracle(ren, mode, lit):
    """SetClipPlane + EnableClipPlane: the culled list (AbstrRenderer::Clipped) and
    the clipped frames equal the oracle's; bricks and samples are removed."""
    vol, tf = setup_blobs(ren, eye=(0.1, 0.2, 1.9))
    ren.SetRotation(rotation_y(0.5))
    ren.SetRendermode(mode)
    ren.SetUseLighting(lit)
    ren.SetIsoValue(100.0)
    tf2 = tf2d_gaussian_lobes().GetByteArray()
    ren.Set2DTrans(tf2)
    ren.SetGradientScale(1.0 / ren.MaxGradientMagnitude())
    ds = O.Dataset(vol, (18, 18, 18), 1)
    ren.SetCountSamples(True)
    full = ren.Paint()
    full_n, full_list = ren.last_stats.samples, len(ren.PlanFrame())
    try:
        ren.SetClipPlane((0.6, -0.3, 0.74, 0.08))
        ren.EnableClipPlane(True)
        rot = rotation_y(0.5)
        got = ren.PlanFrame()
        want = O.plan(ds, oracle_camera_from(ren, rot), oracle_params_from(ren, tf1d=tf, tf2d=tf2))
        assert [(g.toc_index, g.is_empty) for g in got] == [(w.toc_index, w.is_empty) for w in want]
        assert len(got) < full_list
        img, ref, mx, mean, ns, hs = compare(ren, ds, tf1d=tf, tf2d=tf2, rot=rot)
        print(f"clip mode={mode} lit={lit}: max {mx:.3g} mean {mean:.3g} samples {hs}/{full_n}", flush=True)
        assert mx <= TOL_MAX and mean < TOL_MEAN
        assert hs == ns and hs < full_n
        assert not np.array_equal(img, full)
    finally:
        ren.DisableClipPlane()
        ren.SetRotation(np.eye(4, dtype=np.float32))


def test_pick_returns_the_isosurface_hit(ren):
    """GLRenderer::Pick on a sphere SDF: a window position covered by the sphere gives
    an eye-space point on the sphere (within a voxel of its radius) that projects
    back onto that pixel; the background raises "No intersection"; outside
    isosurface mode Pick is refused."""
    n = 64
    g = (np.arange(n) + 0.5) / n - 0.5
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    rad = 0.3
    sdf = np.clip(128 + 20 * (rad - np.sqrt(x * x + y * y + z * z)) * n, 0, 255).astype(np.uint8)
    ren.LoadVolume(sdf, (18, 18, 18), 1, max_lods=1)
    ren.SetRendermode(RM_ISOSURFACE)
    ren.SetIsoValue(128.0)
    ren.SetViewParameters(45.0, 0.01, 1000.0)
    ren.SetViewPosition((0.0, 0.0, 2.0))
    ren.SetRotation(np.eye(4, dtype=np.float32))
    W, H = 96, 80
    ren.Resize(W, H)
    img = ren.Paint()
    view, proj = ren.GetViewMatrices()
    inv_view = np.linalg.inv(view.astype(np.float64))
    checked = 0
    for (px, py) in [(48, 40), (40, 35), (55, 46), (47, 30)]:
        row = H - py
        assert img[row, px, 3] == 1.0
        p = ren.Pick(px, py).astype(np.float64)
        m = np.append(p, 1.0) @ inv_view                 # eye -> model (rotation is identity)
        assert abs(np.linalg.norm(m[:3]) - rad) < 1.5 / n, (px, py, np.linalg.norm(m[:3]))
        c = np.append(p, 1.0) @ proj.astype(np.float64)
        sx, sy = (c[0] / c[3] + 1) * 0.5 * W, (c[1] / c[3] + 1) * 0.5 * H
        assert abs(sx - (px + 0.5)) < 0.01 and abs(sy - (row + 0.5)) < 0.01, (sx, sy, px, row)
        checked += 1
    assert checked == 4
    with pytest.raises(Exception, match="No intersection"):
        ren.Pick(2, 2)
    ren.SetRendermode(RM_1DTRANS)
    with pytest.raises(Exception, match="isosurface"):
        ren.Pick(48, 40)


def test_crop_dataset_matches_closed_form(ren):
    """CropDataset: the LOD-0 voxels on the clipped side of the plane (taken out of
    the viewing transformation and made scale invariant) become 0 and the volume
    is bricked again: the kept volume equals the numpy crop with the plane used,
    the plane is the float64 restatement's, and the pool is the oracle's bricking
    of the cropped volume -- for a loaded volume and for bricks handed over key
    by key (flattened on the device)."""
    dims = (40, 32, 36)
    vol = np.random.default_rng(71).integers(1, 256, size=dims[::-1]).astype(np.uint8)
    plane = np.array([0.4, -0.5, 0.77, 0.05], np.float32)
    rot = rotation_y(0.8)
    for ingest_bricks in (False, True):
        if ingest_bricks:
            ds0 = O.Dataset(vol, (12, 12, 12), 1)
            keys = []
            for i in range(ds0.brick_count):
                info = ds0.brick_info(i)
                x, y, z, lod = info["coords"]
                bx, by, _ = ds0.lod_bricks(lod)
                keys.append((lod, x + bx * (y + by * z), info["minmax"]))
            ren.SetBricks(np.uint8, dims, (12, 12, 12), 1, keys)
            for i, (lod, index, _) in enumerate(keys):
                ren.UploadBrick(lod, index, ds0.brick_data(i))
        else:
            ren.LoadVolume(vol, (12, 12, 12), 1, max_lods=0, keep_linear=True)
        ren.SetViewParameters(45.0, 0.01, 1000.0)
        ren.SetViewPosition((0.0, 0.0, 2.0))
        ren.Resize(32, 32)
        ren.SetRotation(rot)
        ren.SetClipPlane(plane)
        used = ren.CropDataset()
        # the plane in float64: world plane * (world^-1)^-T, normalised; then times
        # the scale-invariance scaling, normalised
        world = rot.astype(np.float64)
        pm = plane.astype(np.float64) @ world.T
        pm /= np.linalg.norm(pm[:3])
        sc = np.array(dims, np.float64) / max(dims)
        pm = np.append(pm[:3] * sc, pm[3])
        pm /= np.linalg.norm(pm[:3])
        assert np.abs(used - pm).max() < 1e-5
        X, Y, Z = dims
        zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
        f32 = np.float32
        vx = (xx.astype(f32) / f32(X)) - f32(0.5)
        vy = (yy.astype(f32) / f32(Y)) - f32(0.5)
        vz = (zz.astype(f32) / f32(Z)) - f32(0.5)
        d = ((used[0] * vx + used[1] * vy) + used[2] * vz) + used[3]
        want = np.where(d >= 0, 0, vol).astype(np.uint8)
        assert 0.2 < (want == 0).mean() < 0.8
        ds = O.Dataset(want, (12, 12, 12), 1)
        assert np.array_equal(ren.GetPool(), ds.all_data())
        if not ingest_bricks:
            assert np.array_equal(ren.DownloadVolume(), want)
    ren.SetRotation(np.eye(4, dtype=np.float32))
    ren.DisableClipPlane()


@pytest.mark.parametrize("dims,brick,ov,dt,median", [
    ((44, 36, 40), (10, 10, 10), 1, np.uint8, True),        # even cores, median filter
    ((37, 21, 50), (12, 10, 16), 1, np.uint16, True),
    ((29, 16, 9), (17, 17, 17), 1, np.uint8, False),        # odd cores: brick-to-brick filtering
    ((45, 23, 30), (19, 19, 19), 2, np.uint16, False),
    ((33, 40, 22), (13, 15, 13), 1, np.uint8, True),
    ((61, 7, 33), (11, 9, 13), 1, np.uint16, True),
])
def test_lod_filter_and_odd_cores_bit_exact(ren, dims, brick, ov, dt, median):
    """ExtendedOctreeConverter's bComputeMedian and odd brick cores: every level,
    overlaps included, and every brick's min/max equal the oracle's."""
    print(f"lod filter dims={dims} brick={brick} ov={ov} median={median}", flush=True)
    full = 255 if dt == np.uint8 else 65535
    vol = np.random.default_rng(9).integers(0, full + 1, size=dims[::-1]).astype(dt)
    ds = O.Dataset(vol, brick, ov, median=median)
    ren.SetLODFilter(median)
    try:
        ren.LoadVolume(vol, brick, ov, max_lods=0)
    finally:
        ren.SetLODFilter(False)
    pool = ren.GetPool()
    ref = ds.all_data()
    assert ds.lod_count > 1 and pool.nbytes == ref.nbytes
    assert np.array_equal(pool, ref)
    for i in range(ds.brick_count):
        assert ren.MaxMinForKey(i) == ds.brick_info(i)["minmax"]


def test_odd_cores_with_clamp_are_refused(ren):
    vol = np.zeros((20, 20, 20), np.uint8)
    with pytest.raises(Exception):
        ren.LoadVolume(vol, (11, 11, 11), 1, clamp_to_edge=True, max_lods=0)


@pytest.mark.gpu
def test_queued_renders_give_the_same_frames(ren):
    """Renders that return once queued (no download, no stats) leave the frames
    a waiting render gives: the next call's host planning and staging copies do
    not disturb the frame still on the GPU."""
    setup_blobs(ren)
    views = [rotation_y(a) for a in (0.0, 0.7, 1.9, 3.1)]
    other = [rotation_y(a) for a in (0.3, 1.1, 2.5, 4.0)]
    ren.PaintViews(views)
    ref = ren.Display()
    for _ in range(3):
        ren.PaintViews(other, download=False, stats=False)
        ren.PaintViews(views, download=False, stats=False)
    assert np.array_equal(ren.Display(), ref)
    ren.PaintViews(other, download=False, stats=False)
    ren.Synchronize()
    assert np.array_equal(ren.PaintViews(other), ren.PaintViews(other))
