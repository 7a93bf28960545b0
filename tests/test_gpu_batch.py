"""Batched rendering (persistent kernel + device work queues + hipGraph) and the
isosurface workload, through the C ABI.

A batch request's frame must equal the single render of its renderer at the
same rotation and size bit for bit: both run the same planner and the same
per-ray march, only the tile scheduling differs.  The single render is itself
checked against the oracle by test_gpu_parity.py; the isosurface SDF case is
checked against the oracle here.
"""
import numpy as np
import pytest

import oracle_ffi as O
from helpers import err_stats, oracle_camera_from, oracle_params_from
from tuvok_amd import (GEN_ELLIPSOID_SDF, GEN_FBM_SHELLS, RM_2DTRANS, RM_ISOSURFACE, HIPRaycaster, RenderBatch,
                       rotation_y, tf2d_gaussian_lobes, workloads)

pytestmark = pytest.mark.gpu

TOL_MAX = 1.0 / 255.0
TOL_MEAN = 1e-3


def small_requests():
    """8 requests over the three batch volumes at 1/8 scale, uneven sizes."""
    rng = np.random.default_rng(3)
    out = []
    for k in range(8):
        w, h = int(rng.integers(40, 200)), int(rng.integers(40, 200))
        out.append((k % 3, 2000 + k, rotation_y(float(rng.uniform(0, 6.28))), w, h))
    return out


def extra_renderers():
    """A lit 2D-TF uint16 request and an isosurface request (other kernel variants)."""
    a = HIPRaycaster(0)
    a.GenerateVolume(GEN_FBM_SHELLS, 1, np.uint16, (96, 96, 96), (34, 34, 34), 1, max_lods=1, keep_linear=True)
    a.Set2DTrans(tf2d_gaussian_lobes().GetByteArray())
    a.SetRendermode(RM_2DTRANS)
    a.SetUseLighting(True)
    a.SetGradientScale(1.0 / a.MaxGradientMagnitude())
    a.SetViewParameters(45.0, 0.01, 1000.0)
    a.SetViewPosition((0.0, 0.0, 2.0))
    a.Resize(120, 90)
    b = HIPRaycaster(0)
    b.GenerateVolume(GEN_ELLIPSOID_SDF, 2, np.uint8, (128, 128, 128), (34, 34, 34), 1, max_lods=1)
    b.SetRendermode(RM_ISOSURFACE)
    b.SetIsoValue(128.0)
    b.SetRefineSteps(4)
    b.SetViewParameters(45.0, 0.01, 1000.0)
    b.SetViewPosition((0.0, 0.0, 2.0))
    b.Resize(100, 100)
    return [(a, rotation_y(0.4), 120, 90), (b, rotation_y(1.1), 100, 100)]


@pytest.fixture(scope="module")
def mixed():
    batch, rens, owners = workloads.setup_batch(small_requests(), scale=8)
    sizes = [(r[3], r[4]) for r in small_requests()]
    rots = [r[2] for r in small_requests()]
    for (r, rot, w, h) in extra_renderers():
        batch.Add(r, rot, w, h)
        r.SetRotation(rot)
        rens.append(r)
        sizes.append((w, h))
        rots.append(rot)
    yield batch, rens, sizes, rots, owners
    batch.close()


def singles(rens, eps):
    out = []
    for r in rens:
        r.SetEarlyRayTermination(eps)
        out.append(r.Paint())
    return out


@pytest.mark.parametrize("tile,graph", [((8, 8), True), ((16, 4), True), ((32, 2), False), ((4, 16), True)])
@pytest.mark.parametrize("eps", [0.0, 1.0 / 2048])
def test_batch_equals_single_renders(mixed, tile, graph, eps):
    batch, rens, sizes, rots, _ = mixed
    print(f"batch tile={tile} graph={graph} eps={eps}", flush=True)
    want = singles(rens, eps)
    batch.SetTuning(tile_w=tile[0], tile_h=tile[1], use_graph=graph)
    batch.Render()
    for i, w in enumerate(want):
        got = batch.Frame(i)
        assert got.shape == w.shape
        assert np.array_equal(got, w), (i, float(np.abs(got - w).max()))
    assert any(w[..., 3].max() > 0.05 for w in want)


def test_batch_graph_replays_with_eager_work_between(mixed):
    """Every replay of the captured frame equals the eager (uncaptured) result."""
    batch, rens, _, _, _ = mixed
    for r in rens:
        r.SetEarlyRayTermination(0.0)
    batch.SetTuning(use_graph=False)
    batch.Render()
    eager = [batch.Frame(i) for i in range(len(rens))]
    batch.SetTuning(use_graph=True)
    for rep in range(4):
        batch.Render()
        for i in range(len(rens)):
            assert np.array_equal(batch.Frame(i), eager[i]), (rep, i)
        rens[rep % len(rens)].Paint()   # eager work of the program between replays


def test_batch_sample_count_is_the_sum_of_single_counts(mixed):
    batch, rens, _, _, _ = mixed
    total = 0
    for r in rens:
        r.SetEarlyRayTermination(0.0)
        r.SetCountSamples(True)
        r.Paint()
        total += r.last_stats.samples
        r.SetCountSamples(False)
    batch.SetTuning()
    batch.SetCountSamples(True)
    st = batch.Render()
    batch.SetCountSamples(False)
    assert st.samples == total


def test_batch_autotune_keeps_results(mixed):
    batch, rens, _, _, _ = mixed
    want = singles(rens, 0.0)
    best, ms = batch.Autotune(reps=1)
    print(f"autotune: {best} {ms:.3f} ms", flush=True)
    assert best["tile_w"] * best["tile_h"] == 64 and ms > 0
    assert batch.GetTuning() == best
    batch.Render()
    for i, w in enumerate(want):
        assert np.array_equal(batch.Frame(i), w)


def test_batch_rejects_bad_tuning_and_empty(mixed):
    b = RenderBatch(0)
    with pytest.raises(Exception):
        b.Render()                      # empty batch
    with pytest.raises(Exception):
        b.SetTuning(tile_w=8, tile_h=4)  # not 64 pixels
    b.close()


def test_iso_sdf_refine4_matches_oracle():
    """Ellipsoid SDF, iso 128 with 4 bisection steps, gradient shading, vs the oracle."""
    r = HIPRaycaster(0)
    n = 128
    r.GenerateVolume(GEN_ELLIPSOID_SDF, 2, np.uint8, (n, n, n), (34, 34, 34), 1, max_lods=1, keep_linear=True)
    vol = r.DownloadVolume()
    r.SetRendermode(RM_ISOSURFACE)
    r.SetIsoValue(128.0)
    r.SetRefineSteps(4)
    r.SetEarlyRayTermination(0.0)
    r.SetViewParameters(45.0, 0.01, 1000.0)
    r.SetViewPosition((0.1, 0.2, 2.0))
    r.Resize(128, 96)
    ds = O.Dataset(vol, (34, 34, 34), 1)
    for rot in (None, rotation_y(0.8)):
        if rot is not None:
            r.SetRotation(rot)
        ref, _ = O.render(ds, oracle_camera_from(r, rot), oracle_params_from(r))
        img = r.Paint()
        mx, mean = err_stats(img, ref)
        print(f"  iso sdf: max {mx:.3g} mean {mean:.3g} hits {int((ref[..., 3] > 0).sum())}", flush=True)
        assert (ref[..., 3] > 0).sum() > 100
        assert mx <= TOL_MAX and mean < TOL_MEAN
    r.close()


def test_sdf_iso_2048_full_size():
    """The isosurface workload at full size: 2048^3 uint8 signed distance of 500
    ellipsoids (128^3 cores + 1 overlap), iso 128, 4 bisection steps, gradient
    shading, 1920x1080.  The frame is opaque-or-empty everywhere and a 96x64
    window at the centre equals the oracle (LOD0 bricking of the downloaded
    volume) within the gate."""
    print("sdf iso 2048: generating", flush=True)
    r = workloads.setup(workloads.SDF_ISO_2048, keep_linear=True)
    r.SetEarlyRayTermination(0.0)
    img = r.Paint()
    ms = []
    for _ in range(3):
        r.Paint(download=False)
        ms.append(r.last_stats.ms)
    print(f"sdf iso 2048: {min(ms):.2f} ms per 1080p frame", flush=True)
    a = img[..., 3]
    assert np.isfinite(img).all()
    assert set(np.unique(a)).issubset({0.0, 1.0})
    hits = int((a > 0).sum())
    print(f"sdf iso 2048: {hits} hit pixels", flush=True)
    assert hits > 0.05 * a.size
    assert img[..., :3].min() >= 0.0 and img[..., :3].max() <= 1.0
    print("sdf iso 2048: oracle bricking (LOD0)", flush=True)
    vol = r.DownloadVolume()
    ds = O.Dataset(vol, workloads.SDF_ISO_2048["brick"], 1, lod0_only=True)
    del vol
    x0, y0 = 960 - 48, 540 - 32
    ref, _ = O.render(ds, oracle_camera_from(r), oracle_params_from(r), roi=(x0, y0, x0 + 96, y0 + 64))
    got = img[y0:y0 + 64, x0:x0 + 96]
    mx, mean = err_stats(got, ref)
    print(f"sdf iso 2048 window: max {mx:.3g} mean {mean:.3g} hits {int((ref[..., 3] > 0).sum())}", flush=True)
    assert (ref[..., 3] > 0).sum() > 100
    assert mx <= TOL_MAX and mean < TOL_MEAN
    r.close()


def test_batch48_full_size():
    """The 48-request batch over 256^3 / 1024^3 / 2048^3 uint8 volumes in one frame:
    premultiplied colour bounds hold everywhere and three requests equal their
    single renders."""
    print("batch48 full size: generating volumes", flush=True)
    reqs = workloads.batch48()
    batch, rens, owners = workloads.setup_batch(reqs, ert_eps=1.0 / 2048)
    print("batch48: autotune", flush=True)
    best, ms = batch.Autotune(reps=1)
    print(f"batch48: best {best} {ms:.2f} ms", flush=True)
    batch.Render()
    px = 0
    for i in range(len(reqs)):
        f = batch.Frame(i)
        px += f.shape[0] * f.shape[1]
        assert np.isfinite(f).all()
        a = f[..., 3]
        assert a.min() >= 0.0 and a.max() <= 1.0 + 1e-6
        assert (f[..., :3] <= a[..., None] + 1e-5).all()
    for i in (0, 17, 47):
        assert np.array_equal(batch.Frame(i), rens[i].Paint()), i
    print(f"batch48: {px} pixels", flush=True)
    batch.close()


def test_lod64g_batch_full_size():
    """4096x4096x2048 uint16 (64 GiB) with a 5-level pyramid in the HBM pool, 32
    requests mixing 3840x2160 and 1920x1080 at different distances: each request
    takes its screen-space-error level, the batch renders them in one frame and
    two of them equal their single renders."""
    print("lod64g: generating 64 GiB volume + pyramid", flush=True)
    reqs = workloads.lod_requests()
    batch, rens, owner = workloads.setup_lod_batch(reqs)
    d = owner.GetDatasetDesc()
    print(f"lod64g: pool {d.pool_bytes / 2**30:.1f} GiB, {d.brick_count} bricks, {d.lod_count} levels", flush=True)
    assert d.lod_count == 5
    lods = [r.CurrentLOD() for r in rens]
    print(f"lod64g: levels {sorted(set(lods))}", flush=True)
    assert len(set(lods)) >= 2 and max(lods) <= 4
    st = batch.Render()
    st = batch.Render()
    px = sum(w * h for (_, _, w, h) in reqs)
    print(f"lod64g: {px / 1e6:.1f} Mpixels in {st.ms:.1f} ms ({px / st.ms / 1e3:.1f} Mpixels/s)", flush=True)
    for i in (0, 1, 30, 31):
        f = batch.Frame(i)
        a = f[..., 3]
        assert np.isfinite(f).all() and a.min() >= 0 and a.max() <= 1 + 1e-6
        assert (f[..., :3] <= a[..., None] + 1e-5).all()
        if i in (0, 31):
            assert np.array_equal(f, rens[i].Paint()), i
    batch.close()


def test_lod64g_learned_batch():
    """The 32-request batch over the 64 GiB volume with the learned step
    proposal: a network trained on chunk samples of 12 other views (their own
    levels), calibrated on 4 further requests within 0.5/255 of the fixed
    batch, renders the 32 requests within 1/255 of the fixed batch."""
    print("lod64g learned: generating 64 GiB volume + pyramid", flush=True)
    owner = workloads.lod_owner()
    reqs = workloads.lod_requests()
    cal = workloads.lod_requests(n=4, seed=19)
    train = [(d, rot, 960, 540) for (d, rot, _, _) in workloads.lod_requests(n=12, seed=29)]
    res = workloads.lod_learned(owner, reqs, cal, train)
    print(f"lod64g learned: fixed {res['fixed_ms']} ms, learned {res['learned_ms']} ms ({res['speedup']}x), "
          f"max err {res['max_err_x255']}/255, proposal {res['proposal']}, chunks {res['chunks']}", flush=True)
    for t in sorted(res["calibration"], key=lambda t: t[2])[:4]:
        print(f"  calibration: {t}", flush=True)
    assert res["chunks"] > 1000 and res["loss_last"] < res["loss_first"]
    assert res["max_err"] <= TOL_MAX
    owner.close()


def test_batch_frames_follow_the_last_successful_render():
    """A frame is returned only for a request of the last successful render: after
    Clear and re-adding a larger request (no render yet) Frame is refused instead
    of reading past the output buffer; after the render it is the new frame."""
    a = HIPRaycaster(0)
    a.GenerateVolume(GEN_FBM_SHELLS, 4, np.uint16, (64, 64, 64), (34, 34, 34), 1, max_lods=1, keep_linear=True)
    a.Set2DTrans(tf2d_gaussian_lobes().GetByteArray())
    a.SetRendermode(RM_2DTRANS)
    a.SetGradientScale(1.0 / a.MaxGradientMagnitude())
    a.SetViewParameters(45.0, 0.01, 1000.0)
    a.SetViewPosition((0.0, 0.0, 2.0))
    a.Resize(32, 24)
    b = RenderBatch(0)
    try:
        b.Add(a, np.eye(4, dtype=np.float32), 32, 24)
        b.Render()
        assert b.Frame(0).shape == (24, 32, 4)
        b.Clear()
        b.Add(a, np.eye(4, dtype=np.float32), 300, 200)
        with pytest.raises(Exception, match="no such request"):
            b.Frame(0)
        b.Render()
        a.Resize(300, 200)
        assert np.array_equal(b.Frame(0), a.Paint())
    finally:
        b.close()
        a.close()
