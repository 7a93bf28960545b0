"""Brick-level ingest (tvk_set_bricks / tvk_upload_brick) and the budgeted HBM
pool, through the C ABI.

Ingest: the oracle's bricks, handed over key by key with their min/max the
way a BrickedDataset holds them, give the oracle's pool byte for byte and its
frames.  Budget: a dataset held in fewer slots than a frame needs renders in
parts, paging bricks in least recently used first; every frame equals the
all-resident frame bit for bit (same planner, same march, the parts continue
each other's pixels), sample counts included.
"""
import numpy as np
import pytest

import oracle_ffi as O
from helpers import err_stats, oracle_camera_from, oracle_params_from
from tuvok_amd import (GEN_BLOBS, GEN_ELLIPSOID_SDF, RM_1DTRANS, RM_2DTRANS, RM_ISOSURFACE, HIPRaycaster, rotation_y,
                       tf1d_blue_white_red, tf2d_gaussian_lobes)

pytestmark = pytest.mark.gpu


def keys_of(ds):
    keys = []
    for i in range(ds.brick_count):
        info = ds.brick_info(i)
        x, y, z, lod = info["coords"]
        bx, by, _ = ds.lod_bricks(lod)
        keys.append((lod, x + bx * (y + by * z), info["minmax"]))
    return keys


def ingest(ren, ds, vol, budget=0, order=None):
    keys = keys_of(ds)
    ren.SetBricks(vol.dtype, vol.shape[::-1], ds.brick, ds.overlap, keys, budget_bytes=budget)
    idx = range(ds.brick_count) if order is None else order
    for i in idx:
        lod, index, _ = keys[i]
        ren.UploadBrick(lod, index, ds.brick_data(i))


def view(ren, w=96, h=80, eye=(0.3, -0.2, 1.8)):
    ren.SetViewParameters(45.0, 0.01, 1000.0)
    ren.SetViewPosition(eye)
    ren.Resize(w, h)


def test_ingest_equals_oracle():
    """Bricks uploaded in a shuffled order give the oracle's pool and frames; a
    render before the last payload is refused; an unknown key is a range error."""
    vol = np.random.default_rng(31).integers(0, 256, size=(40, 44, 36)).astype(np.uint8)
    vol = (vol // 4 + np.linspace(0, 190, 36)[None, None, :]).astype(np.uint8)
    ds = O.Dataset(vol, (12, 12, 12), 1)
    r = HIPRaycaster(0)
    try:
        order = list(np.random.default_rng(2).permutation(ds.brick_count))
        keys = keys_of(ds)
        r.SetBricks(vol.dtype, vol.shape[::-1], ds.brick, ds.overlap, keys)
        for i in order[:-1]:
            r.UploadBrick(keys[i][0], keys[i][1], ds.brick_data(i))
        tf = tf1d_blue_white_red().GetByteArray()
        r.Set1DTrans(tf)
        r.SetRendermode(RM_1DTRANS)
        view(r)
        with pytest.raises(Exception, match="not uploaded"):
            r.Paint()
        with pytest.raises(Exception, match="unknown brick"):
            r.UploadBrick(0, 10 ** 6, ds.brick_data(0))
        last = order[-1]
        r.UploadBrick(keys[last][0], keys[last][1], ds.brick_data(last))
        assert np.array_equal(r.GetPool(), ds.all_data())
        for i in range(ds.brick_count):
            assert r.MaxMinForKey(i) == ds.brick_info(i)["minmax"]
        for lit in (False, True):
            r.SetUseLighting(lit)
            img = r.Paint()
            ref, ns = O.render(ds, oracle_camera_from(r), oracle_params_from(r, tf1d=tf))
            mx, mean = err_stats(img, ref)
            assert img[..., 3].max() > 0.05
            assert mx <= 1.0 / 255 and mean < 1e-3, (lit, mx, mean)
    finally:
        r.close()


def resident_and_budgeted(setup, budgets, frames):
    """Frames of a resident renderer and of the same dataset at each budget."""
    a = HIPRaycaster(0)
    setup(a)
    want = frames(a)
    out = []
    for b in budgets:
        a.SetPoolBudget(b)
        got = frames(a)
        out.append((b, got, a.GetPoolStats()))
    a.close()
    return want, out


@pytest.mark.parametrize("mode,lit,eps", [(RM_1DTRANS, False, 0.0), (RM_1DTRANS, True, 1 / 512),
                                          (RM_2DTRANS, True, 0.0), (RM_ISOSURFACE, False, 0.0),
                                          (RM_ISOSURFACE, True, 1 / 512)])
def test_budgeted_frames_equal_resident(mode, lit, eps):
    print(f"budget mode={mode} lit={lit} eps={eps}", flush=True)

    def setup(r):
        kind = GEN_ELLIPSOID_SDF if mode == RM_ISOSURFACE else GEN_BLOBS
        r.GenerateVolume(kind, 7, np.uint8, (128, 128, 128), (34, 34, 34), 1, max_lods=2, keep_linear=True)
        r.Set1DTrans(tf1d_blue_white_red().GetByteArray())
        r.Set2DTrans(tf2d_gaussian_lobes().GetByteArray())
        r.SetGradientScale(1.0 / r.MaxGradientMagnitude())
        r.SetRendermode(mode)
        r.SetIsoValue(128.0)
        r.SetUseLighting(lit)
        r.SetEarlyRayTermination(eps)
        r.SetCountSamples(True)
        view(r, 120, 90, eye=(0.2, 0.3, 1.6))

    rots = [rotation_y(a) for a in (0.0, 1.2, 2.9)]

    def frames(r):
        imgs = r.PaintViews(rots)
        return imgs, r.last_stats.samples, r.Paint(), r.last_stats.samples

    slot = 40 * 36 * 36 * 2   # blocked 34^3 uint8 brick: 40 x 36 x 36 voxels padded, doubled rows
    want, got = resident_and_budgeted(setup, [3 * slot, 11 * slot, 200 * slot], frames)
    assert want[0][..., 3].max() > 0.05 or mode == RM_ISOSURFACE
    for b, g, st in got:
        assert np.array_equal(g[0], want[0]), (b, float(np.abs(g[0] - want[0]).max()))
        assert np.array_equal(g[2], want[2]), b
        if not (mode == RM_ISOSURFACE and eps > 0):
            # the isosurface's brick skip behind the hit knows only the hits of
            # its own part, so a budgeted frame may take more samples
            assert g[1] == want[1] and g[3] == want[3], (b, g[1], want[1])
        if b == 3 * slot:
            assert st.slots == 3 and st.parts > 1 and st.evictions > 0, (st.slots, st.parts, st.evictions)


def test_budgeted_ingest_and_lru():
    """A dataset ingested straight into a budget: frames equal the resident
    ingest; a second identical frame that fits the slots uploads nothing; the
    payloads come back from the host store."""
    vol = np.random.default_rng(41).integers(0, 65536, size=(48, 40, 56)).astype(np.uint16)
    ds = O.Dataset(vol, (18, 18, 18), 1)
    tf = tf1d_blue_white_red().GetByteArray()
    a, b = HIPRaycaster(0), HIPRaycaster(0)
    try:
        ingest(a, ds, vol)
        slot = 20 * 20 * 20 * 2 * 2   # blocked 18^3 uint16 brick, doubled rows
        ingest(b, ds, vol, budget=12 * slot)
        for r in (a, b):
            r.Set1DTrans(tf)
            r.SetRendermode(RM_1DTRANS)
            r.SetUseLighting(True)
            r.SetLOD(1)               # 8 bricks: the frame fits the 12 slots
            view(r, 64, 48)
        want1 = a.Paint()
        assert np.array_equal(b.Paint(), want1)
        st0 = b.GetPoolStats()
        assert st0.slots == 12 and st0.parts == 1
        assert np.array_equal(b.Paint(), want1)
        st1 = b.GetPoolStats()
        assert st1.uploads == st0.uploads   # every brick of the frame stayed resident
        for r in (a, b):
            r.SetLOD(0)               # 36 bricks: rendered in parts
        want = a.Paint()
        assert np.array_equal(b.Paint(), want)
        assert b.GetPoolStats().parts > 1
        assert np.array_equal(b.GetPool(), ds.all_data())
        b.SetPoolBudget(3 * slot)
        assert np.array_equal(b.Paint(), want)
        assert b.GetPoolStats().parts > 1
        b.SetPoolBudget(0)
        assert np.array_equal(b.Paint(), want)
        assert np.array_equal(b.GetPool(), ds.all_data())
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("budget_slots", [0, 5])
def test_cpp_client_ingests_and_renders(tmp_path, budget_slots):
    """tests/cpp/ingest_render, a C++ program built against include/tvk_raycaster.hpp
    and linked to libtvkhip.so, hands the oracle's bricks over key by key (with
    a budget of a few slots, or all resident) and paints: the frame matches the
    oracle's and the Python front end's bit for bit."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "ingest_render")
    assert os.path.exists(exe), "tests/cpp/ingest_render is built by __graft_entry__.build()"
    vol = np.random.default_rng(51).integers(0, 256, size=(36, 30, 42)).astype(np.uint8)
    vol = (vol // 3 + np.linspace(0, 170, 42)[None, None, :]).astype(np.uint8)
    ds = O.Dataset(vol, (12, 12, 12), 1)
    keys = keys_of(ds)
    w, h, eye, fov = 80, 64, (0.25, 0.1, 1.9), 45.0
    tf = tf1d_blue_white_red().GetByteArray()
    with open(tmp_path / "meta.txt", "w") as f:
        f.write(f"1 42 30 36 12 12 12 1 {len(keys)} {w} {h} {eye[0]} {eye[1]} {eye[2]} {fov} 1\n")
    rec = np.zeros(len(keys), dtype=[("lod", "<u4"), ("index", "<u8"), ("mn", "<f8"), ("mx", "<f8")])
    for i, (lod, index, mm) in enumerate(keys):
        rec[i] = (lod, index, mm[0], mm[1])
    rec.tofile(tmp_path / "keys.bin")
    with open(tmp_path / "bricks.bin", "wb") as f:
        for i in range(ds.brick_count):
            f.write(ds.brick_data(i).tobytes())
    np.ascontiguousarray(tf, np.uint8).tofile(tmp_path / "tf1d.bin")
    slot = 16 * 12 * 12 * 2
    r = subprocess.run([exe, str(tmp_path), str(budget_slots * slot)], capture_output=True, text=True, timeout=60)
    print(r.stdout.strip(), r.stderr.strip(), flush=True)
    assert r.returncode == 0, r.stderr
    img = np.fromfile(tmp_path / "frame.f32", np.float32).reshape(h, w, 4)
    py = HIPRaycaster(0)
    try:
        ingest(py, ds, vol)
        py.Set1DTrans(tf)
        py.SetRendermode(RM_1DTRANS)
        py.SetUseLighting(True)
        py.SetViewParameters(fov, 0.01, 1000.0)
        py.SetViewPosition(eye)
        py.Resize(w, h)
        want = py.Paint()
        ref, _ = O.render(ds, oracle_camera_from(py), oracle_params_from(py, tf1d=tf))
    finally:
        py.close()
    assert img[..., 3].max() > 0.05
    assert np.array_equal(img, want)
    mx, mean = err_stats(img, ref)
    assert mx <= 1.0 / 255 and mean < 1e-3


def numpy_max_gradient(vol, clamp=False):
    """MaxGradientMagnitude of the level-0 2D histogram, closed form: the largest
    |(f(x-1)-f(x+1))/(2 norm), ...| over every voxel, zero (or the edge value)
    outside the volume, in float64 (Histogram2DDataBlock.cpp:326-395)."""
    norm = 255.0 if vol.dtype == np.uint8 else 65535.0
    p = np.pad(vol.astype(np.float64), 1, mode="edge" if clamp else "constant")
    gx = (p[1:-1, 1:-1, :-2] - p[1:-1, 1:-1, 2:]) / (norm * 2)
    gy = (p[1:-1, :-2, 1:-1] - p[1:-1, 2:, 1:-1]) / (norm * 2)
    gz = (p[:-2, 1:-1, 1:-1] - p[2:, 1:-1, 1:-1]) / (norm * 2)
    return np.float32(np.sqrt(gx * gx + gy * gy + gz * gz).max())


@pytest.mark.parametrize("dt,brick,ov,clamp", [(np.uint8, (12, 12, 12), 1, False), (np.uint16, (18, 14, 16), 2, False),
                                               (np.uint16, (10, 10, 10), 1, True)])
def test_max_gradient_closed_form(dt, brick, ov, clamp):
    """tvk_max_gradient of a loaded volume, of the same bricks ingested key by key
    (no linear volume), and of a budgeted pool all equal the numpy maximum."""
    full = 255 if dt == np.uint8 else 65535
    rng = np.random.default_rng(61)
    vol = rng.integers(0, full + 1, size=(30, 26, 34)).astype(dt)
    vol[5:9, 4:20, 3:30] = full        # a slab with sharp faces
    want = numpy_max_gradient(vol, clamp)
    a = HIPRaycaster(0)
    try:
        a.LoadVolume(vol, brick, ov, clamp_to_edge=clamp, max_lods=0, keep_linear=True)
        assert a.MaxGradientMagnitude() == want
        ds = O.Dataset(vol, brick, ov, clamp=clamp)
        if not clamp:   # the oracle's coarse clamped overlaps are not the device's (level 0 is)
            ingest(a, ds, vol)
            assert a.MaxGradientMagnitude() == want
        # four slots (a generous bound of one brick's blocked bytes each)
        a.SetPoolBudget(4 * int(np.prod([((s + 7) // 8) * 8 for s in brick])) * 2 * vol.itemsize)
        assert a.GetPoolStats().slots >= 2
        assert a.MaxGradientMagnitude() == want
    finally:
        a.close()
