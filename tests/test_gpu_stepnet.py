"""The step-proposal network on the GPU (BF16 MFMA kernels, FP32 accumulation)
against the FP64 oracle with the same bfloat16 rounding points, and the
learned-step march against the fixed-step march."""
import numpy as np
import pytest

import oracle_ffi as O
from test_oracle_stepnet import rand_problem
from tuvok_amd import GEN_FBM_SHELLS, RM_2DTRANS, HIPRaycaster, rotation_y, tf2d_gaussian_lobes
from tuvok_amd.stepnet import StepNet, decide, labels, segment_errors, split_samples

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    n = StepNet(0, seed=3)
    yield n
    n.close()


def test_inference_matches_oracle(net):
    th, _, _ = rand_problem(seed=4)
    net.set_params(th)
    x = np.random.default_rng(5).uniform(-1, 1, (1000, 24)).astype(np.float32)
    y = net.infer(x)
    ref = O.sn_forward(th, x, bf16=True)
    err = np.abs(y - ref)
    print(f"infer: max abs err {err.max():.3g}", flush=True)
    assert (err <= 2e-3 * (1 + np.abs(ref))).all()


@pytest.mark.parametrize("n", [64, 200, 1000])
def test_gradient_matches_oracle(net, n):
    th, _, _ = rand_problem(seed=6)
    net.set_params(th)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (n, 24)).astype(np.float32)
    t = np.stack([rng.uniform(0, 1, n), rng.integers(0, 2, n), rng.uniform(0.5, 2, n)], 1).astype(np.float32)
    net.set_data(x, t)
    g, loss = net.grad()
    ref, rloss = O.sn_grad(th, x, t, bf16=True)
    scale = np.abs(ref).max()
    err = np.abs(g - ref)
    print(f"grad n={n}: max err {err.max():.3g} of max |g| {scale:.3g}; loss {loss:.6g} vs {rloss:.6g}", flush=True)
    assert err.max() <= 2e-2 * scale
    assert np.linalg.norm(g - ref) <= 1e-2 * np.linalg.norm(ref)
    assert abs(loss - rloss) <= 1e-4 * abs(rloss)


def test_training_fits_a_teacher(net):
    """Targets from a fixed random teacher network: Adam on the device kernels
    must reduce the loss by a large factor."""
    teacher, _, _ = rand_problem(seed=7)
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, (16384, 24)).astype(np.float32)
    yt = O.sn_forward(teacher, x, bf16=False)
    t = np.stack([yt[:, 0], (yt[:, 1] > 0).astype(np.float64), np.ones(len(x))], 1).astype(np.float32)
    net.set_params(StepNet(0, seed=9).get_params())
    net.set_data(x, t)
    _, l0 = net.grad()
    losses = net.train(epochs=30, batch=1024, lr=3e-3)
    _, l1 = net.grad()
    print(f"teacher fit: loss {l0 / len(x):.4f} -> {l1 / len(x):.4f}; epochs {losses[0]:.4f} .. {losses[-1]:.4f}",
          flush=True)
    assert l1 < 0.3 * l0
    assert losses[-1] < losses[0]


def fbm_renderer(n=96, w=160, h=120, srm=2.0):
    r = HIPRaycaster(0)
    r.GenerateVolume(GEN_FBM_SHELLS, 1, np.uint16, (n, n, n), (34, 34, 34), 1, max_lods=1, keep_linear=True)
    r.Set2DTrans(tf2d_gaussian_lobes().GetByteArray())
    r.SetRendermode(RM_2DTRANS)
    r.SetUseLighting(True)
    r.SetGradientScale(1.0 / r.MaxGradientMagnitude())
    r.SetSampleRateModifier(srm)
    r.SetEarlyRayTermination(0.0)
    r.SetViewParameters(45.0, 0.01, 1000.0)
    r.SetViewPosition((0.0, 0.0, 2.0))
    r.Resize(w, h)
    return r


def test_proposal_of_multiplier_one_equals_fixed_step(net):
    """A proposal that always answers multiplier 1 and never skips gives the
    fixed-step frame bit for bit (same march, same float operations)."""
    r = fbm_renderer()
    for rot in (rotation_y(0.3), rotation_y(2.0)):
        r.SetRotation(rot)
        r.SetStepProposal(None)
        fixed = r.Paint()
        r.SetStepProposal(net, margin=1e9, skip_threshold=1.0)
        learned = r.Paint()
        assert np.array_equal(fixed, learned)
    r.SetStepProposal(None)
    r.close()


def test_segment_samples_are_consistent():
    r = fbm_renderer()
    rec = r.GenerateSegments([rotation_y(a) for a in (0.0, 1.0, 2.0)], cap=200000, keep=0.5, seed=1, fit=True)
    f, tb, steps, c = split_samples(rec)
    print(f"segments: {len(rec)}", flush=True)
    assert len(rec) > 1000
    assert (tb >= 0).all() and (tb <= 1).all() and (steps >= 1).all()
    assert (c[..., 3] >= 0).all() and (c[..., 3] <= 1 + 1e-6).all()
    assert (c[..., :3] <= c[..., 3:4] + 1e-5).all()       # premultiplied colour (lit colour clamped to 1)
    assert (f[:, 23] == 1).all() and np.allclose(f[:, 15], tb)
    assert (steps <= 32).all()
    # the contribution bound (transmittance x the most opacity the chunk's value
    # range allows) holds for the chunk's fixed-step alpha
    assert (f[:, 19] * (1 + 1e-5) + 1e-7 >= tb * c[:, 0, 3]).all()
    e = segment_errors(rec)
    assert (e[:, 0] == 0).all()
    r.close()


def test_cell_grid_matches_brute_force():
    """Per brick, cells of 8^3 stored voxels; each word holds the min and max of
    the cell grown by 4 voxels on every side, clipped to the brick."""
    r = HIPRaycaster(0)
    r.GenerateVolume(GEN_FBM_SHELLS, 5, np.uint16, (70, 61, 50), (34, 34, 34), 1, max_lods=2)
    cg = r.CellGrid()
    off = 0
    for t in range(r.GetDatasetDesc().brick_count):
        b = r.GetBrick(t)   # z, y, x
        nz, ny, nx = b.shape
        cx, cy, cz = [(n + 7) // 8 for n in (nx, ny, nz)]
        for k in range(cz):
            for j in range(cy):
                for i in range(cx):
                    sub = b[max(8 * k - 4, 0):8 * k + 12, max(8 * j - 4, 0):8 * j + 12, max(8 * i - 4, 0):8 * i + 12]
                    w = int(cg[off + (k * cy + j) * cx + i])
                    assert (w & 0xFFFF, w >> 16) == (int(sub.min()), int(sub.max())), (t, i, j, k)
        off += cx * cy * cz
    assert off == len(cg)
    r.close()


def test_exact_skip_leaves_frame_unchanged_with_fewer_samples():
    """The chunked march with multiplier 1 skips chunks whose value range is
    transparent under the TF: same frame bit for bit, fewer samples (phantom
    with a transparent background, lit 1D TF)."""
    from tuvok_amd import GEN_PHANTOM, RM_1DTRANS
    from tuvok_amd.transfer import tf1d_phantom
    r = HIPRaycaster(0)
    r.GenerateVolume(GEN_PHANTOM, 4, np.uint16, (128, 128, 128), (34, 34, 34), 1, max_lods=1)
    r.Set1DTrans(tf1d_phantom().GetByteArray())
    r.SetRendermode(RM_1DTRANS)
    r.SetUseLighting(True)
    r.SetSampleRateModifier(2.0)
    r.SetViewParameters(45.0, 0.01, 1000.0)
    r.SetViewPosition((0.0, 0.0, 2.0))
    r.Resize(160, 120)
    r.SetCountSamples(True)
    for rot in (rotation_y(0.4), rotation_y(2.2)):
        r.SetRotation(rot)
        r.SetStepProposal(None)
        fixed = r.Paint()
        nf = r.last_stats.samples
        r.SetStepProposal(None, fixed_log2=0, skip_bound=0.0)
        skipped = r.Paint()
        ns = r.last_stats.samples
        print(f"exact skip: samples {ns}/{nf} = {ns / nf:.3f}", flush=True)
        assert np.array_equal(fixed, skipped)
        assert ns < 0.7 * nf
    r.SetStepProposal(None)
    r.close()


def test_variation_fallback_orders_samples_and_error():
    """The variation-risk fallback on the chunked march at multiplier 4
    (phantom, lit 1D TF): a threshold no chunk reaches changes nothing; lower
    thresholds force more chunks (and their wavefronts) back to the fixed step:
    the smallest one ends near the multiplier-1 sample count, well above the
    multiplier-4 one, and closer to the fixed-step frame."""
    from tuvok_amd import GEN_PHANTOM, RM_1DTRANS
    from tuvok_amd.transfer import tf1d_phantom
    r = HIPRaycaster(0)
    r.GenerateVolume(GEN_PHANTOM, 4, np.uint16, (128, 128, 128), (34, 34, 34), 1, max_lods=1)
    r.Set1DTrans(tf1d_phantom().GetByteArray())
    r.SetRendermode(RM_1DTRANS)
    r.SetUseLighting(True)
    r.SetSampleRateModifier(2.0)
    r.SetViewParameters(45.0, 0.01, 1000.0)
    r.SetViewPosition((0.0, 0.0, 2.0))
    r.Resize(160, 120)
    r.SetRotation(rotation_y(0.4))
    r.SetCountSamples(True)
    try:
        r.SetStepProposal(None)
        fixed = r.Paint()
        r.SetStepProposal(None, fixed_log2=0, skip_bound=0.0)
        r.Paint()
        n1 = r.last_stats.samples
        got = []
        for tv in (0.0, 1e9, 0.05, 0.005, 0.0005, 1e-7):
            r.SetStepProposal(None, fixed_log2=2, skip_bound=0.0, fallback_tv=tv)
            img = r.Paint()
            got.append((tv, img, r.last_stats.samples, float(np.abs(img - fixed).max())))
            print(f"fallback_tv {tv}: samples {got[-1][2]} (x1: {n1}) max err {got[-1][3] * 255:.3f}/255", flush=True)
    finally:
        r.SetStepProposal(None)
        r.close()
    assert np.array_equal(got[0][1], got[1][1]) and got[0][2] == got[1][2]
    assert got[0][2] < 0.5 * n1 and got[0][2] < got[-1][2] <= 1.05 * n1
    assert got[-1][3] < got[0][3]


def test_learned_step_gain_at_error_bound(net):
    """End to end on a 128^3 uint16 fBm+shells volume, lit 2D TF, 0.25-voxel
    reference step: chunk samples from 48 views, labels, training on the
    device, margin calibrated on 4 other views, evaluated on 3 held-out views.
    The learned march must stay within 1/255 of the fixed-step frame and take
    fewer samples than the same chunked march without the network (multiplier
    1, only the exact skip of transparent chunks): the gain of the network."""
    rng = np.random.default_rng(11)
    r = fbm_renderer(n=128, w=192, h=144)
    train_rots = [rotation_y(a) for a in rng.uniform(0, 2 * np.pi, 48)]
    rec = r.GenerateSegments(train_rots, cap=1_000_000, keep=1.0, seed=2, fit=True)
    tau = 0.25 / 255.0
    tgt = labels(rec, tau=tau, tau_skip=tau / 2)
    f = split_samples(rec)[0]
    k_mean = (tgt[:, 0] * 3).mean()
    print(f"learned: {len(rec)} chunks, mean safe log2 multiplier {k_mean:.2f}, skip {tgt[:, 1].mean():.2f}",
          flush=True)
    net.set_params(StepNet(0, seed=12).get_params())
    net.set_data(f, tgt)
    losses = net.train(epochs=30, batch=4096, lr=2e-3)
    print(f"learned: loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.8 * losses[0]

    exact = dict(margin=1e9, skip_threshold=1.0)   # the chunked march without the network's decisions

    def frame(rot, proposal):
        r.SetRotation(rot)
        if proposal is None:
            r.SetStepProposal(None)
        else:
            r.SetStepProposal(net, **proposal)
        r.SetCountSamples(True)
        img = r.Paint()
        return img, r.last_stats.samples

    val = [rotation_y(a) for a in rng.uniform(0, 2 * np.pi, 4)]
    ref_val = [frame(v, None) for v in val]
    best = None
    for margin in (0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0):
        for skip in (0.999, 1.0):
            prop = dict(margin=margin, skip_threshold=skip, max_log2=3)
            errs, ns = [], 0
            for v, (ref, nref) in zip(val, ref_val):
                img, n = frame(v, prop)
                errs.append(np.abs(img - ref).max())
                ns += n
            if max(errs) <= 0.9 / 255 and (best is None or ns < best[1]):
                best = (prop, ns)
    assert best is not None, "no calibration setting met the error bound"
    prop = best[0]
    test_views = [rotation_y(a) for a in rng.uniform(0, 2 * np.pi, 3)]
    n_fixed = n_exact = n_learned = 0
    worst = 0.0
    for v in test_views:
        ref, nr = frame(v, None)
        img_e, ne = frame(v, exact)
        img, nl = frame(v, prop)
        assert np.array_equal(img_e, ref)   # the exact skip leaves the frame unchanged
        worst = max(worst, float(np.abs(img - ref).max()))
        n_fixed += nr
        n_exact += ne
        n_learned += nl
    r.SetStepProposal(None)
    print(f"learned: proposal {prop}; held-out max err {worst * 255:.3f}/255; samples learned/exact/fixed "
          f"{n_learned}/{n_exact}/{n_fixed} = {n_learned / n_exact:.3f} of the exact-skip march", flush=True)
    assert worst <= 1.0 / 255
    assert n_learned < 0.95 * n_exact
    r.close()


def test_batch_learned_with_multiplier_one_equals_fixed(net):
    """The learned batch (persistent kernel + step proposal) with a proposal of
    multiplier 1 gives the fixed-step batch frames bit for bit."""
    from tuvok_amd import RenderBatch
    r = fbm_renderer()
    b = RenderBatch(0)
    for a, (w, h) in zip((0.2, 1.7, 3.1), ((160, 120), (96, 64), (200, 150))):
        b.Add(r, rotation_y(a), w, h)
    b.Render()
    fixed = [b.Frame(i) for i in range(3)]
    b.SetStepProposal(net, margin=1e9, skip_threshold=1.0)
    b.Render()
    for i in range(3):
        assert np.array_equal(b.Frame(i), fixed[i]), i
    b.SetStepProposal(None)
    b.close()
    r.close()
