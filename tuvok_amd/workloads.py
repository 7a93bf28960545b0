"""Synthetic workloads of the raycasting path (volume, bricking, TF, camera, views).

BLOBS_256 is the CPU-runnable 1D-TF case (256^3 uint8 blobs, 64^3 cores + 1
voxel overlap, 512x512, one view).  FBM_SHELLS_1024 is the lit 2D-TF case at
scale (1024^3 uint16 fBm + shells, 128^3 cores + 1 voxel overlap, 1920x1080,
16 views on an orbit).  SDF_ISO_2048 is the isosurface case (2048^3 uint8
signed distance of 500 ellipsoids, iso 128, 4 bisection steps).  PHANTOM_2048
is the second volume of the learned-step case (2048^3 uint16 phantom of 40
nested ellipsoids, intensities 0-4000, Gaussian noise sigma 20).  batch48()
lists the 48 uneven requests over three uint8 volumes rendered as one batch.
"""
import numpy as np

from . import _abi
from .renderer import HIPRaycaster, rotation_y
from .transfer import tf1d_blue_white_red, tf1d_phantom, tf1d_random, tf2d_gaussian_lobes

BLOBS_256 = dict(
    name="blobs256_u8_1dtf_512x512",
    gen=_abi.GEN_BLOBS, seed=0, dtype=np.uint8, dims=(256, 256, 256), brick=(66, 66, 66), overlap=1,
    mode=_abi.RM_1DTRANS, lighting=False, width=512, height=512, fov=45.0, eye=(0.0, 0.0, 2.0), views=1,
)

FBM_SHELLS_1024 = dict(
    name="fbm_shells1024_u16_2dtf_lit_1080p_x16",
    gen=_abi.GEN_FBM_SHELLS, seed=1, dtype=np.uint16, dims=(1024, 1024, 1024), brick=(130, 130, 130), overlap=1,
    mode=_abi.RM_2DTRANS, lighting=True, width=1920, height=1080, fov=45.0, eye=(0.0, 0.0, 2.0), views=16,
)


SDF_ISO_2048 = dict(
    name="sdf2048_u8_iso128_refine4_1080p",
    gen=_abi.GEN_ELLIPSOID_SDF, seed=2, dtype=np.uint8, dims=(2048, 2048, 2048), brick=(130, 130, 130), overlap=1,
    mode=_abi.RM_ISOSURFACE, lighting=True, isovalue=128.0, refine_steps=4, width=1920, height=1080, fov=45.0,
    eye=(0.0, 0.0, 2.0), views=1,
)
PHANTOM_2048 = dict(
    name="phantom2048_u16_1dtf_lit_1080p",
    gen=_abi.GEN_PHANTOM, seed=4, dtype=np.uint16, dims=(2048, 2048, 2048), brick=(130, 130, 130), overlap=1,
    mode=_abi.RM_1DTRANS, lighting=True, width=1920, height=1080, fov=45.0, eye=(0.0, 0.0, 2.0), views=16,
    tf=tf1d_phantom,
)

LOD_64G = dict(
    name="fbm_shells4096x4096x2048_u16_5lod_2dtf_lit_32req",
    gen=_abi.GEN_FBM_SHELLS, seed=3, dtype=np.uint16, dims=(4096, 4096, 2048), brick=(258, 258, 258), overlap=1,
    mode=_abi.RM_2DTRANS, lighting=True, max_lods=5, fov=45.0,
)


def lod_requests(n=32, seed=9, near=2.0, far=8.0):
    """Requests of the large-volume workload: (eye distance, rotation, width,
    height); half 3840x2160 and half 1920x1080, distances log-uniform."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        w, h = (3840, 2160) if k % 2 == 0 else (1920, 1080)
        out.append((float(np.exp(rng.uniform(np.log(near), np.log(far)))), rotation_y(rng.uniform(0, 2 * np.pi)),
                    w, h))
    return out


def lod_owner(cfg=LOD_64G, device=0, dims=None):
    """The renderer that holds the large volume (every kept level in the HBM
    pool, linear volume kept for the gradient maximum)."""
    owner = HIPRaycaster(device)
    owner.GenerateVolume(cfg["gen"], cfg["seed"], cfg["dtype"], dims or cfg["dims"], cfg["brick"], cfg["overlap"],
                         max_lods=cfg["max_lods"], keep_linear=True)
    return owner


def lod_renderer(owner, dist, rot, w, h, cfg=LOD_64G, device=0, ert_eps=2.0 ** -11, gmax=None):
    """A renderer of the large volume sharing owner's pool: lit 2D TF, its
    own eye distance and screen-space-error level choice."""
    g = owner.MaxGradientMagnitude() if gmax is None else gmax
    r = HIPRaycaster(device)
    r.ShareDataset(owner)
    r.Set2DTrans(tf2d_gaussian_lobes().GetByteArray())
    r.SetRendermode(cfg["mode"])
    r.SetUseLighting(cfg["lighting"])
    r.SetGradientScale(1.0 / g if g > 0 else 1.0)
    r.SetLOD(-1)
    r.SetEarlyRayTermination(ert_eps)
    r.SetViewParameters(cfg["fov"], 0.01, 1000.0)
    r.SetViewPosition((0.0, 0.0, dist))
    r.Resize(w, h)
    r.SetRotation(rot)
    return r


def setup_lod_batch(requests, cfg=LOD_64G, device=0, ert_eps=2.0 ** -11, dims=None, owner=None):
    """The large volume (all kept levels in the HBM pool) shared by one renderer
    per request, each with its own camera distance and screen-space-error level
    choice, in one RenderBatch.  owner: an existing lod_owner() to share."""
    from .batch import RenderBatch
    owner = owner or lod_owner(cfg, device, dims)
    g = owner.MaxGradientMagnitude()
    batch = RenderBatch(device)
    rens = []
    for (dist, rot, w, h) in requests:
        r = lod_renderer(owner, dist, rot, w, h, cfg, device, ert_eps, gmax=g)
        batch.Add(r, rot, w, h)
        rens.append(r)
    return batch, rens, owner


def lod_learned(owner, requests, cal_requests, train_requests, bound=0.5 / 255, seed=0, cap=1_000_000,
                epochs=20):
    """Learned step proposal on the large-volume batch.  Chunk samples of the
    fixed march over train_requests (eye distance, rotation, width, height)
    train a network; the setting of it that renders the cal_requests batch
    fastest within `bound` of the fixed batch is chosen (the fixed march is a
    candidate); then the `requests` batch is timed with and without it and
    compared frame by frame.  Returns a dict of the figures."""
    from .stepnet import StepNet, labels, split_samples
    g = owner.MaxGradientMagnitude()
    rec = []
    for k, (dist, rot, w, h) in enumerate(train_requests):
        r = lod_renderer(owner, dist, rot, w, h, gmax=g)
        rec.append(r.GenerateSegments([rot], cap // max(1, len(train_requests)), keep=0.05, seed=seed + k, fit=True))
        r.close()
    rec = np.concatenate(rec)
    tau = 0.25 / 255
    tgt = labels(rec, tau, tau / 2)
    net = StepNet(0, seed=seed)
    net.set_data(split_samples(rec)[0], tgt)
    losses = net.train(epochs=epochs)

    def frames(b, n):
        return [b.Frame(i) for i in range(n)]

    def timed(b, reps=3):
        b.Render()
        return float(np.median([b.Render().ms for _ in range(reps)]))

    cal, cal_rens, _ = setup_lod_batch(cal_requests, owner=owner)
    cal.SetStepProposal(None)
    cal_ms = timed(cal)
    cal_ref = frames(cal, len(cal_requests))
    best, best_ms, tried = None, cal_ms, []
    for fb in (0.0, 0.02, 0.005):
        for m in (0.0, 1.0, 2.0):
            prop = dict(margin=m, skip_threshold=1.0, fallback_bound=fb)
            cal.SetStepProposal(net, **prop)
            ms = timed(cal, reps=1)
            err = max(float(np.abs(a - b).max()) for a, b in zip(frames(cal, len(cal_requests)), cal_ref))
            tried.append((prop, round(err * 255, 3), round(ms, 3)))
            if err <= bound and ms < best_ms:
                best, best_ms = prop, ms
    cal.close()
    for x in cal_rens:
        x.close()

    batch, rens, _ = setup_lod_batch(requests, owner=owner)
    batch.SetStepProposal(None)
    fixed_ms = timed(batch)
    ref = frames(batch, len(requests))
    if best is not None:
        batch.SetStepProposal(net, **best)
    learned_ms = timed(batch)
    err = max(float(np.abs(a - b).max()) for a, b in zip(frames(batch, len(requests)), ref))
    del ref
    batch.close()
    for x in rens:
        x.close()
    net.close()
    return dict(fixed_ms=round(fixed_ms, 3), learned_ms=round(learned_ms, 3), speedup=round(fixed_ms / learned_ms, 3),
                max_err=err, max_err_x255=round(err * 255, 3),
                proposal=best if best is not None else "fixed march (no setting faster within the bound)",
                calibration_bound_x255=round(bound * 255, 3), chunks=int(len(rec)),
                loss_first=float(losses[0]), loss_last=float(losses[-1]), calibration=tried)


# the three volumes of the batch: (generator, seed, edge)
BATCH_VOLUMES = [(_abi.GEN_BLOBS, 10, 256), (_abi.GEN_ELLIPSOID_SDF, 11, 1024), (_abi.GEN_ELLIPSOID_SDF, 12, 2048)]


def batch48(n=48, seed=7, min_px=256, max_px=2048):
    """Requests of the batched workload: (volume index, TF seed, rotation, width,
    height).  Image edges are log-uniform in [min_px, max_px] (independently for
    width and height), volumes are picked uniformly, every request has its own
    random 1D TF and view rotation."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        w = int(round(np.exp(rng.uniform(np.log(min_px), np.log(max_px)))))
        h = int(round(np.exp(rng.uniform(np.log(min_px), np.log(max_px)))))
        out.append((int(rng.integers(0, len(BATCH_VOLUMES))), 1000 + k, rotation_y(rng.uniform(0, 2 * np.pi)), w, h))
    return out


def orbit(views):
    """Rotations of the model about y, evenly spaced on a full orbit."""
    return [rotation_y(2.0 * np.pi * k / views) for k in range(views)]


def setup(cfg, device=0, keep_linear=False, max_lods=1, ert_eps=0.0, dims=None, width=None, height=None):
    """Creates a HIPRaycaster with the workload's dataset, TF, camera and mode."""
    ren = HIPRaycaster(device)
    d = dims or cfg["dims"]
    if cfg["mode"] == _abi.RM_2DTRANS:
        keep_linear = True   # MaxGradientMagnitude is computed from the linear volume
    ren.GenerateVolume(cfg["gen"], cfg["seed"], cfg["dtype"], d, cfg["brick"], cfg["overlap"],
                       max_lods=max_lods, keep_linear=keep_linear)
    if cfg["mode"] == _abi.RM_1DTRANS:
        ren.Set1DTrans(cfg.get("tf", tf1d_blue_white_red)().GetByteArray())
    elif cfg["mode"] == _abi.RM_2DTRANS:
        ren.Set2DTrans(tf2d_gaussian_lobes().GetByteArray())
        if keep_linear:
            g = ren.MaxGradientMagnitude()
            ren.SetGradientScale(1.0 / g if g > 0 else 1.0)
    ren.SetRendermode(cfg["mode"])
    ren.SetUseLighting(cfg["lighting"])
    if cfg["mode"] == _abi.RM_ISOSURFACE:
        ren.SetIsoValue(cfg["isovalue"])
        ren.SetRefineSteps(cfg["refine_steps"])
    ren.SetEarlyRayTermination(ert_eps)
    ren.SetViewParameters(cfg["fov"], 0.01, 1000.0)
    ren.SetViewPosition(cfg["eye"])
    ren.Resize(width or cfg["width"], height or cfg["height"])
    return ren


def setup_batch(requests, device=0, scale=1, ert_eps=0.0, lighting=False, fov=45.0, eye=(0.0, 0.0, 2.0)):
    """Renderers and the RenderBatch of the batched workload: one HIPRaycaster
    per volume holds its dataset, one per request shares it and holds the
    request's 1D TF.  scale divides every volume edge (reduced test cases)."""
    from .batch import RenderBatch
    owners = {}
    rens = []
    batch = RenderBatch(device)
    for (vi, tf_seed, rot, w, h) in requests:
        if vi not in owners:
            gen, seed, edge = BATCH_VOLUMES[vi]
            e = max(16, edge // scale)
            r0 = HIPRaycaster(device)
            r0.GenerateVolume(gen, seed, np.uint8, (e, e, e), (130, 130, 130) if e >= 512 else (66, 66, 66), 1,
                              max_lods=1)
            owners[vi] = r0
        r = HIPRaycaster(device)
        r.ShareDataset(owners[vi])
        r.Set1DTrans(tf1d_random(tf_seed).GetByteArray())
        r.SetRendermode(_abi.RM_1DTRANS)
        r.SetUseLighting(lighting)
        r.SetEarlyRayTermination(ert_eps)
        r.SetViewParameters(fov, 0.01, 1000.0)
        r.SetViewPosition(eye)
        r.Resize(w, h)
        r.SetRotation(rot)
        batch.Add(r, rot, w, h)
        rens.append(r)
    return batch, rens, owners
