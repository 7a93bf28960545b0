"""Synthetic workloads of the raycasting path (volume, bricking, TF, camera, views).

BLOBS_256 is the CPU-runnable 1D-TF case (256^3 uint8 blobs, 64^3 cores + 1
voxel overlap, 512x512, one view).  FBM_SHELLS_1024 is the lit 2D-TF case at
scale (1024^3 uint16 fBm + shells, 128^3 cores + 1 voxel overlap, 1920x1080,
16 views on an orbit).
"""
import numpy as np

from . import _abi
from .renderer import HIPRaycaster, rotation_y
from .transfer import tf1d_blue_white_red, tf2d_gaussian_lobes

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
        ren.Set1DTrans(tf1d_blue_white_red().GetByteArray())
    else:
        ren.Set2DTrans(tf2d_gaussian_lobes().GetByteArray())
        if keep_linear:
            g = ren.MaxGradientMagnitude()
            ren.SetGradientScale(1.0 / g if g > 0 else 1.0)
    ren.SetRendermode(cfg["mode"])
    ren.SetUseLighting(cfg["lighting"])
    ren.SetEarlyRayTermination(ert_eps)
    ren.SetViewParameters(cfg["fov"], 0.01, 1000.0)
    ren.SetViewPosition(cfg["eye"])
    ren.Resize(width or cfg["width"], height or cfg["height"])
    return ren
