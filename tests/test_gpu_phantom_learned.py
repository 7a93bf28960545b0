"""The learned step proposal at full size: the 2048^3 uint16 phantom (40 nested
ellipsoids with noise, 16 GiB) at 1920x1080, lit 1D TF, against the fixed march
at the 0.25-voxel reference step.  The network is trained on the device from
chunk samples of 16 random views (480x270), calibrated on 3 other 1080p views
(the fastest setting within 0.75/255 of the fixed frames), and checked on 4
held-out 1080p views: every frame within 1/255 of the fixed-step frame, with
fewer samples executed."""
import numpy as np
import pytest

from tuvok_amd import workloads
from tuvok_amd.renderer import rotation_y
from tuvok_amd.stepnet import calibrate, train_proposal

pytestmark = pytest.mark.gpu


def test_learned_step_on_the_full_size_phantom():
    cfg = dict(workloads.PHANTOM_2048)
    print(f"phantom {cfg['dims']} at {cfg['width']}x{cfg['height']}", flush=True)
    ren = workloads.setup(cfg, ert_eps=2.0 ** -11)
    try:
        rng = np.random.default_rng(2024)
        ren.SetSampleRateModifier(2.0)            # the 0.25-voxel reference step
        ren.Resize(480, 270)
        net, info = train_proposal(ren, [rotation_y(a) for a in rng.uniform(0, 2 * np.pi, 16)], keep=0.05,
                                   epochs=20)
        print(f"phantom: trained {info}", flush=True)
        ren.Resize(cfg["width"], cfg["height"])
        cal = [rotation_y(a) for a in rng.uniform(0, 2 * np.pi, 3)]
        prop, tried = calibrate(ren, net, cal, bound=0.75 / 255, margins=(0.0, 1.0, 2.0, 3.0),
                                fallbacks=(0.0, 0.01, 0.002))
        print(f"phantom: calibration picked {prop} of {len(tried)} settings", flush=True)
        assert prop is not None, "no proposal setting was faster within the calibration bound"
        held = [rotation_y(a) for a in rng.uniform(0, 2 * np.pi, 4)]
        worst, n_fixed, n_learned = 0.0, 0, 0
        ren.SetCountSamples(True)
        for k, rot in enumerate(held):
            ren.SetStepProposal(None)
            fixed = ren.PaintViews([rot])
            nf = ren.last_stats.samples
            ren.SetStepProposal(net, **prop)
            learned = ren.PaintViews([rot])
            nl = ren.last_stats.samples
            err = float(np.abs(learned - fixed).max())
            print(f"phantom: view {k} max err {err * 255:.3f}/255 samples {nl}/{nf}", flush=True)
            assert fixed[..., 3].max() > 0.5
            worst = max(worst, err)
            n_fixed += nf
            n_learned += nl
        ren.SetStepProposal(None)
        print(f"phantom: held-out max err {worst * 255:.3f}/255, samples {n_learned / n_fixed:.3f} of the "
              f"fixed march", flush=True)
        assert worst <= 1.0 / 255
        assert n_learned < n_fixed
    finally:
        ren.close()
