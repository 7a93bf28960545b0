"""Benchmark of the HIP raycasting path on one MI355X.

Workload: 1024^3 uint16 fBm + 64 hollow shells, bricked into 8^3 bricks of
128^3 core + 1 overlap voxel (130^3 stored), lit 2D-TF raycast (4 Gaussian
lobes over value x gradient magnitude), 1920x1080, 16 views on an orbit,
0.5-voxel step (the reference's fRayStepsize), FP32 compositing.
One step = the 16 views rendered in one launch with inputs resident in HBM;
the steps run back to back (each render returns once queued, so the host plans
the next step while the GPU renders this one).

Metric: Gsamples/s = N_ref / t, where N_ref is the number of composited
samples of the fixed-step reference march (per-brick restart, local
a >= 0.99 exit, no cross-brick termination) counted on the device in an
untimed pass; frames/s = 16 / t.  The timed pass uses early ray termination
at a total transmittance of 2^-11 (error <= 2^-11 per channel).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ERT_EPS = 2.0 ** -11


def cpu_baseline(ren, vol, views, threads, budget_s=15.0):
    """The oracle (CPU restatement of the reference raycaster) on bands of rows of
    the first view, from the middle outwards, until about budget_s of host time.
    The oracle's rows are kept and compared with the HIP frame of the same view
    (reference semantics, no early termination): max and mean per-channel error,
    and, when every row was rendered, the composited-sample counts."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    from helpers import oracle_camera_from, oracle_params_from
    from tuvok_amd.transfer import tf2d_gaussian_lobes
    cfg_brick = ren.GetDatasetDesc().brick_size
    ds = O.Dataset(vol, tuple(cfg_brick), 1, lod0_only=True)
    ren.SetRotation(views[0])
    cam = oracle_camera_from(ren, views[0])
    p = oracle_params_from(ren, tf2d=tf2d_gaussian_lobes().GetByteArray())
    p["ert_eps"] = 0.0   # the reference's own semantics: every sample of N_ref is taken
    W, H = cam["width"], cam["height"]
    band = 24
    # bands of rows ordered from the middle of the frame outwards
    starts = sorted(range(0, H, band), key=lambda r: abs(r + band // 2 - H // 2))
    ns_total, dt_total, rows = 0, 0.0, []
    ref = np.zeros((H, W, 4), np.float32)
    for r0 in starts:
        if dt_total >= budget_s:
            break
        roi = (0, r0, W, min(H, r0 + band))
        t0 = time.perf_counter()
        img, ns = O.render(ds, cam, p, roi=roi, threads=threads)
        dt_total += time.perf_counter() - t0
        ns_total += ns
        ref[roi[1]:roi[3]] = img
        rows.append(roi)
    lo, hi = min(r[1] for r in rows), max(r[3] for r in rows)
    # the HIP frame of view 0 with the reference semantics, sample-counted
    eps = ren.GetParams().ert_eps
    ren.SetEarlyRayTermination(0.0)
    ren.SetCountSamples(True)
    ren.SetRotation(views[0])
    hip = ren.Paint()
    hip_ns = int(ren.last_stats.samples)
    ren.SetCountSamples(False)
    ren.SetEarlyRayTermination(eps)
    d = np.abs(hip[lo:hi].astype(np.float64) - ref[lo:hi].astype(np.float64))
    parity = dict(rows=[lo, hi - 1], max_err=float(d.max()), mean_err=float(d.mean()),
                  max_err_x255=round(float(d.max()) * 255, 5),
                  samples_oracle=ns_total if (lo, hi) == (0, H) else None,
                  samples_hip=hip_ns if (lo, hi) == (0, H) else None,
                  samples_equal=(ns_total == hip_ns) if (lo, hi) == (0, H) else None)
    print(f"[bench] view 0, HIP vs oracle: {parity}", file=sys.stderr, flush=True)
    return dict(value=ns_total / dt_total / 1e9, unit="Gsamples/s", cores=threads, kind="port",
                sample=f"oracle (GL-faithful FP32 restatement, no early termination) on rows {lo}..{hi - 1} of view 0 "
                       f"({W}x{hi - lo} rays, {ns_total} samples, {dt_total:.1f} s, {threads} threads)",
                view0_parity=parity)


def hbm_roofline(alg_bytes, ms):
    """Algorithmic bytes of one frame (or batch) over its kernel time against
    MI355X's HBM peak."""
    gbs = alg_bytes / (ms / 1e3) / 1e9
    return {"bound": "hbm", "algorithmic_bytes": int(alg_bytes), "achieved": round(gbs, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5)}


def batch_bytes(requests):
    """V_eff of a batch, deduplicated per dataset: the union of the bricks each
    request's counted render samples (stored bytes), plus 16 B of output and
    4 B of descriptors per pixel.  requests: (renderer, dataset owner key,
    pixels); every renderer has just run a counted render."""
    union, sizes, px = {}, {}, 0
    for r, key, pixels in requests:
        m = r.TouchedMask().astype(bool)
        union[key] = union.get(key, np.zeros_like(m)) | m
        if key not in sizes:
            sizes[key] = r.BrickStoredBytes()
        px += pixels
    return int(sum(int(sizes[k][union[k]].sum()) for k in union)) + 20 * px


def counted_batch_bytes(rens, reqs):
    """batch_bytes of a batch's requests, each counted by one untimed render of
    its renderer (which holds the request's camera and size).  reqs: (dataset
    key, rotation, width, height) per renderer."""
    items = []
    for r, (key, _rot, w, h) in zip(rens, reqs):
        r.SetCountSamples(True)
        r.Paint(download=False)
        items.append((r, key, w * h))
    b = batch_bytes(items)
    for r in rens:
        r.SetCountSamples(False)
    return b


def other_workloads(learned=True):
    """Frame times of the other full-size workloads (reported beside the main
    line, not part of its value): device bricking of the 1024^3 uint16 volume
    and one all-resident GridLeaper pass over it, the 2048^3 uint8 SDF isosurface at 1080p,
    the 48-request batch over three uint8 volumes in one persistent launch, and
    the 32-request batch over the 64 GiB uint16 volume with its 5-level pyramid
    in the HBM pool, also with the learned step proposal (learned=True).  Each
    is built, timed over a few frames and freed."""
    from tuvok_amd import workloads
    res = {}
    try:
        # device bricking of the 1024^3 uint16 volume from its kept linear copy
        # (gather with overlap, extended rows, per-brick min/max), wall clock
        c = dict(workloads.FBM_SHELLS_1024)
        r = workloads.setup(c, keep_linear=True)
        r.Rebrick(c["brick"], c["overlap"], max_lods=1)
        ts, ks = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            r.Rebrick(c["brick"], c["overlap"], max_lods=1)
            ts.append(time.perf_counter() - t0)
            ks.append(r.GetDatasetDesc().build_ms)
        d = r.GetDatasetDesc()
        vox = int(np.prod(c["dims"]))
        stored = d.brick_count * int(np.prod(c["brick"])) * 2
        t, k = float(np.median(ts)), float(np.median(ks))
        # algorithmic bytes: the linear volume read once, every stored brick voxel written once
        res["bricking_1024"] = dict(workload="fbm_shells1024_u16 -> 512 bricks of 130^3", device_ms=round(k, 3),
                                    wall_ms=round(t * 1e3, 2), algorithmic_bytes=vox * 2 + stored,
                                    gb_per_s=round((vox * 2 + stored) / (k / 1e3) / 1e9, 1),
                                    hbm_frac=round((vox * 2 + stored) / (k / 1e3) / 1e9 / HBM_PEAK_GBS, 3))
        # GridLeaper: one all-resident pass over every level, 1080p, lit 2D TF
        r.Rebrick(c["brick"], c["overlap"], max_lods=0)
        r.SetRendermode(c["mode"])
        r.LeapSetResident(None)
        r.SetCountSamples(True)
        _, st = r.LeapRender(first=True)
        samples = int(st.samples)
        r.SetCountSamples(False)
        lt = []
        for _ in range(3):
            _, st = r.LeapRender(first=True)
            lt.append(st.ms)
        m = float(np.median(lt))
        res["leaper_1024"] = dict(workload="fbm_shells1024_u16, all levels resident, lit 2D TF, 1920x1080, one pass",
                                  ms_per_pass=round(m, 3), samples=samples,
                                  gsamples_per_s=round(samples / m / 1e6, 1))
        r.close()
    except Exception as e:
        res["bricking_1024"] = dict(error=f"{type(e).__name__}: {e}")
    print(f"[bench] bricking: {res.get('bricking_1024')} leaper: {res.get('leaper_1024')}", file=sys.stderr,
          flush=True)
    try:
        from tuvok_amd.autotune import dispatch_name, tune_dispatch_order
        c = workloads.SDF_ISO_2048
        # early termination on: for the isosurface it only skips the bricks behind
        # a ray's hit (the same frame as the exact mode, fewer samples)
        r = workloads.setup(c, ert_eps=ERT_EPS)
        iso_order, _ = tune_dispatch_order(r, [np.eye(4, dtype=np.float32)])
        r.SetCountSamples(True)
        r.Paint(download=False)
        alg = batch_bytes([(r, 0, c["width"] * c["height"])])
        r.SetCountSamples(False)
        ms = []
        for _ in range(5):
            r.Paint(download=False)
            ms.append(r.last_stats.ms)
        r.close()
        m = float(np.median(ms))
        res["iso_sdf2048"] = dict(workload=c["name"], ms_per_frame=round(m, 3), frames_per_s=round(1e3 / m, 1),
                                  dispatch=dispatch_name(iso_order), roofline=hbm_roofline(alg, m))
    except Exception as e:
        res["iso_sdf2048"] = dict(error=f"{type(e).__name__}: {e}")
    print(f"[bench] iso: {res['iso_sdf2048']}", file=sys.stderr, flush=True)
    try:
        reqs = workloads.batch48()
        batch, rens, owners = workloads.setup_batch(reqs, ert_eps=ERT_EPS)
        best, _ = batch.Autotune(reps=1)
        batch.Render()
        ms = [batch.Render().ms for _ in range(5)]
        px = sum(w * h for (_, _, _, w, h) in reqs)
        m = float(np.median(ms))
        alg = counted_batch_bytes(rens, [(q[0], q[2], q[3], q[4]) for q in reqs])
        res["batch48"] = dict(requests=len(reqs), mpixels=round(px / 1e6, 2), ms_per_frame=round(m, 3),
                              mpixels_per_s=round(px / m / 1e3, 1), tuning=best, roofline=hbm_roofline(alg, m))
        batch.close()
        for x in rens + list(owners.values()):
            x.close()
    except Exception as e:
        res["batch48"] = dict(error=f"{type(e).__name__}: {e}")
    print(f"[bench] batch48: {res['batch48']}", file=sys.stderr, flush=True)
    owner = None
    try:
        reqs = workloads.lod_requests()
        owner = workloads.lod_owner()
        batch, rens, _ = workloads.setup_lod_batch(reqs, owner=owner)
        d = owner.GetDatasetDesc()
        batch.Render()
        ms = [batch.Render().ms for _ in range(3)]
        px = sum(w * h for (_, _, w, h) in reqs)
        m = float(np.median(ms))
        alg = counted_batch_bytes(rens, [(0, rot, w, h) for (_, rot, w, h) in reqs])
        res["lod64g"] = dict(requests=len(reqs), mpixels=round(px / 1e6, 2), ms_per_frame=round(m, 3),
                             mpixels_per_s=round(px / m / 1e3, 1), pool_gib=round(d.pool_bytes / 2 ** 30, 1),
                             levels=sorted(set(int(x.CurrentLOD()) for x in rens)), roofline=hbm_roofline(alg, m))
        batch.close()
        for x in rens:
            x.close()
    except Exception as e:
        res["lod64g"] = dict(error=f"{type(e).__name__}: {e}")
    print(f"[bench] lod64g: {res['lod64g']}", file=sys.stderr, flush=True)
    if owner is not None and learned:
        # the same batch with the learned step proposal (trained on 12 other
        # views at 960x540, calibrated on 4 other requests within 0.5/255)
        try:
            cal = workloads.lod_requests(n=4, seed=19)
            train = [(d_, rot, 960, 540) for (d_, rot, _, _) in workloads.lod_requests(n=12, seed=29)]
            lr = workloads.lod_learned(owner, reqs, cal, train)
            lr["calibration"] = sorted(lr["calibration"], key=lambda t: t[2])[:4]
            res["lod64g_learned"] = lr
        except Exception as e:
            res["lod64g_learned"] = dict(error=f"{type(e).__name__}: {e}")
        print(f"[bench] lod64g learned: {res['lod64g_learned']}", file=sys.stderr, flush=True)
    if owner is not None:
        owner.close()
    return res


def learned_leg(ren, views, cfg, train_views=48, seed=123):
    """Learned step vs fixed step on the same volume, TF and views, at the
    0.25-voxel reference step (srm 2): chunk samples from `train_views` random
    views at 480x270, training on the device, calibration on 10 other views
    (bound 0.75/255, the fixed march a candidate), then `views` timed with and
    without the proposal and compared pixel by pixel."""
    from tuvok_amd import DISPATCH_QUEUES, DISPATCH_RASTER
    from tuvok_amd.autotune import dispatch_name, tune_dispatch_order
    from tuvok_amd.stepnet import calibrate, train_proposal
    from tuvok_amd.renderer import rotation_y
    rng = np.random.default_rng(seed)
    W, H = cfg["width"], cfg["height"]
    ren.SetSampleRateModifier(2.0)
    t0 = time.perf_counter()
    ren.Resize(480, 270)
    net, info = train_proposal(ren, [rotation_y(a) for a in rng.uniform(0, 2 * np.pi, train_views)], keep=0.02,
                               epochs=20)
    ren.Resize(W, H)
    train_s = time.perf_counter() - t0
    print(f"[bench] learned: {info} in {train_s:.1f} s", file=sys.stderr, flush=True)
    prop, tried = calibrate(ren, net, [rotation_y(a) for a in rng.uniform(0, 2 * np.pi, 10)], bound=0.75 / 255)
    for (p, e, ms) in sorted(tried, key=lambda t: t[2])[:6]:
        print(f"[bench] calibration: {p} max err {e * 255:.3f}/255 {ms:.2f} ms", file=sys.stderr, flush=True)
    ren.SetStepProposal(None)
    orders = {"fixed": tune_dispatch_order(ren, views)[0]}
    fixed = ren.PaintViews(views)
    t_fixed = []
    for _ in range(5):
        ren.PaintViews(views, download=False)
        t_fixed.append(ren.last_stats.ms)
    def timed(prop, use_net=True, label=""):
        ren.SetStepProposal(net if (prop is not None and use_net) else None, **(prop or {}))
        orders[label] = tune_dispatch_order(ren, views)[0]
        img = ren.PaintViews(views)
        ts = []
        for _ in range(5):
            ren.PaintViews(views, download=False)
            ts.append(ren.last_stats.ms)
        return img, float(np.median(ts))

    # the chunked march with the exact skip only (multiplier 1, no network
    # evaluated): what the method gives without the network
    exact, ems = timed(dict(fixed_log2=0, skip_bound=0.0), use_net=False, label="exact_skip")
    learned, lms = timed(prop, label="learned")   # prop None: the calibration kept the fixed march
    ren.SetStepProposal(None)
    ren.SetSampleRateModifier(1.0)
    ren.SetDispatchOrder(DISPATCH_RASTER | DISPATCH_QUEUES)   # the default again
    d = np.abs(learned - fixed)
    fms = float(np.median(t_fixed))
    return dict(reference_step_voxels=0.25, views=len(views), fixed_fps=round(len(views) / (fms / 1e3), 2),
                learned_fps=round(len(views) / (lms / 1e3), 2), speedup=round(fms / lms, 3),
                exact_skip_fps=round(len(views) / (ems / 1e3), 2),
                exact_skip_max_err=float(np.abs(exact - fixed).max()),
                max_err=float(d.max()), mean_err=float(d.mean()), max_err_x255=round(float(d.max()) * 255, 3),
                proposal=prop if prop is not None else "fixed march (no proposal setting faster within the bound)",
                train_s=round(train_s, 1), train=info,
                dispatch={k: dispatch_name(v) for k, v in orders.items()})


PROFILE = os.path.join(ROOT, "profiles", "march_fbm1024_lit2dtf_1080p_x16.md")


# VALU peak of MI355X: 256 CUs x 64 lanes (one wave64 instruction per CU per
# clock: 4 SIMD16 units, 4 clocks each) x 2.4 GHz, lane-instructions per second
VALU_PEAK_TLANE = 256 * 64 * 2.4e9 / 1e12


def valu_roofline(prof, kernel_ms):
    """Useful (active-lane) VALU instructions of the timed launch per second
    against the VALU peak, from the committed counter profile of the same
    kernel (None when its kernel time does not match this run's)."""
    if not prof or "sq_insts_valu" not in prof:
        return None
    useful = prof["sq_insts_valu"] * 64 * prof["lane_util"]
    achieved = useful / (kernel_ms / 1e3) / 1e12
    return {"bound": "valu", "achieved": round(achieved, 2), "peak": round(VALU_PEAK_TLANE, 2),
            "unit": "T lane-instr/s", "frac": round(achieved / VALU_PEAK_TLANE, 4),
            "issue_frac": prof.get("valu_busy"), "lane_util": prof.get("lane_util"),
            "instr_per_sample": prof.get("valu_lane_instr_per_sample"), "binding": True,
            "source": "profiles/march_fbm1024_lit2dtf_1080p_x16.md"}


def profile_figures(kernel_ms):
    """HBM traffic and VALU figures of the timed march kernel from the committed
    rocprofv3 summary, when its kernel time agrees with this run's (within 3%)."""
    import re
    try:
        txt = open(PROFILE).read()
    except OSError:
        return None
    f = dict(re.findall(r"^- ([a-z_]+): ([0-9.eE+-]+)\s*$", txt, re.M))
    try:
        prof_ms = float(f["kernel_avg_ms"])
    except KeyError:
        return None
    if abs(prof_ms - kernel_ms) > 0.03 * kernel_ms:
        return None
    return {k: float(v) for k, v in f.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-step legs")
    ap.add_argument("--no-phantom", action="store_true", help="skip the learned-step leg on the 2048^3 phantom")
    ap.add_argument("--size", type=int, default=1024, help="volume edge (default: the 1024^3 workload)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default dispatch (raster order, work queues) instead of timing the settings")
    ap.add_argument("--no-other", action="store_true",
                    help="skip the isosurface, 48-request batch and 64 GiB LOD batch measurements")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path runs on a single "
              f"MI355X: several GPUs would each render their own views or bricks)", file=sys.stderr)
        sys.exit(2)

    from tuvok_amd import workloads
    cfg = dict(workloads.FBM_SHELLS_1024)
    n = args.size
    dims = (n, n, n)
    t0 = time.perf_counter()
    ren = workloads.setup(cfg, dims=dims, keep_linear=True)
    print(f"[bench] dataset ready in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    views = workloads.orbit(cfg["views"])
    nviews = len(views)

    # untimed counting pass: N_ref with the reference semantics (no cross-brick termination)
    ren.SetEarlyRayTermination(0.0)
    ren.SetCountSamples(True)
    ren.PaintViews(views, download=False)
    n_ref = int(ren.last_stats.samples)
    b_bricks, n_touched = ren.TouchedBytes()
    u_bytes, u_bricks = ren.TouchedBytes(unique=True)
    exact_ms = ren.last_stats.ms
    # executed samples with early ray termination
    ren.SetEarlyRayTermination(ERT_EPS)
    ren.PaintViews(views, download=False)
    n_exec = int(ren.last_stats.samples)
    ren.SetCountSamples(False)
    print(f"[bench] N_ref={n_ref} executed={n_exec} touched bricks={n_touched} ({b_bricks / 1e9:.2f} GB over "
          f"views, {u_bricks} bricks / {u_bytes / 1e9:.2f} GB distinct)", file=sys.stderr, flush=True)

    # the group order of the launch (frames are the same in every order)
    from tuvok_amd.autotune import dispatch_name, tune_dispatch_order
    if args.no_tune:
        from tuvok_amd import DISPATCH_QUEUES, DISPATCH_RASTER
        order, order_ms = DISPATCH_RASTER | DISPATCH_QUEUES, {}
    else:
        order, order_ms = tune_dispatch_order(ren, views)
    print(f"[bench] dispatch: {dispatch_name(order)} "
          f"({', '.join(f'{dispatch_name(k)} {v:.2f} ms' for k, v in order_ms.items())})", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        ren.PaintViews(views, download=False)
    ren.Synchronize()
    # the timed steps: each call returns once its frame is queued, so the host
    # plans the next frame while the GPU renders this one
    t_wall = time.perf_counter()
    ren.EventRecord(0)
    for _ in range(args.steps):
        ren.PaintViews(views, download=False, stats=False)
    ren.EventRecord(1)
    ren.Synchronize()
    wall = time.perf_counter() - t_wall
    total_ms = ren.EventElapsed(0, 1)
    ms_step = total_ms / args.steps
    # the march kernel's own duration (HIP events around the launch), one frame at a time
    kernel_ms = []
    for _ in range(args.steps):
        ren.PaintViews(views, download=False)
        kernel_ms.append(ren.EventElapsed(6, 7))
    k_ms = float(np.mean(kernel_ms))
    pixels = cfg["width"] * cfg["height"] * nviews
    # distinct bricks sampled by the batch + RGBA32F out + 4 B of ray/tile state per pixel
    alg_bytes = u_bytes + pixels * 16 + pixels * 4
    achieved = alg_bytes / (k_ms / 1e3) / 1e9
    value = n_ref / (ms_step / 1e3) / 1e9

    prof = profile_figures(k_ms)
    learned = None
    if not args.no_learned:
        try:
            learned = learned_leg(ren, views, cfg)
        except Exception as e:   # the fixed-step line stays valid; report what failed
            learned = dict(error=f"{type(e).__name__}: {e}")
        print(f"[bench] learned: {learned}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        vol = ren.DownloadVolume()
        thr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        cpu = cpu_baseline(ren, vol, views, thr)
        del vol
    ren.close()
    phantom = None
    if not args.no_learned and not args.no_phantom:
        try:
            pc = workloads.PHANTOM_2048
            t0 = time.perf_counter()
            pren = workloads.setup(pc, ert_eps=ERT_EPS)
            print(f"[bench] phantom ready in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
            phantom = learned_leg(pren, workloads.orbit(4), pc, train_views=32, seed=321)
            phantom["workload"] = pc["name"]
            pren.close()
        except Exception as e:
            phantom = dict(error=f"{type(e).__name__}: {e}")
        print(f"[bench] learned (phantom): {phantom}", file=sys.stderr, flush=True)

    other = None
    if not args.no_other:
        other = other_workloads(learned=not args.no_learned)

    out = {
        "metric": "Gsamples/s (fixed-step-equivalent) + frames/s @1080p",
        "value": round(value, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16 voxels, f32 compositing",
        "data": "synthetic (seeded fBm + shells generated on the device)",
        "config": {"workload": f"fbm_shells{n}_u16_2dtf_lit_1920x1080_x{nviews}views",
                   "volume": list(dims), "brick_stored": list(cfg["brick"]), "overlap": cfg["overlap"],
                   "mode": "2D TF + lighting", "step_voxels": "0.5 (fRayStepsize)", "views": nviews,
                   "ert_eps": ERT_EPS, "dispatch": dispatch_name(order)},
        "frames_per_s": round(nviews / (ms_step / 1e3), 2),
        "samples_ref": n_ref,
        "samples_executed": n_exec,
        "exact_mode_ms_per_step": round(exact_ms, 3),
        "wall_s": round(wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": int(prof["traffic_bytes"]) if prof else None,
                     "algorithmic_bytes_per_launch": int(alg_bytes), "kernel_ms": round(k_ms, 3),
                     "brick_bytes_summed_over_views": int(b_bricks),
                     "valu_busy": prof.get("valu_busy") if prof else None,
                     "lane_util": prof.get("lane_util") if prof else None},
        # the bound that binds: the march issues VALU instructions (value, gradient
        # cross, shading) on nearly every cycle; HBM traffic is a few percent of peak
        "valu": valu_roofline(prof, k_ms),
        "cpu_baseline": cpu,
        "learned_step": learned,
        "learned_step_phantom": phantom,
        "other_workloads": other,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
