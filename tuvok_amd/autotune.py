"""Brick-size autotuner: re-bricks a renderer's dataset on the device with each
candidate stored brick size and keeps the fastest for the given views; the
dispatch-order tuner times the group orders of a multi-view render.

The tile shape and grid size of batches are tuned by RenderBatch.Autotune
(tvk_batch_autotune); the step schedule of the learned march by
stepnet.calibrate.  Brick size is a property of the dataset, so it is tuned
here, on the renderer that owns the dataset (its linear volume must be kept).
"""
import numpy as np

from ._abi import DISPATCH_GROUPS, DISPATCH_QUEUES, DISPATCH_RASTER, DISPATCH_ROWS

DISPATCH_SETTINGS = tuple(o | q for q in (DISPATCH_QUEUES, 0) for o in (DISPATCH_RASTER, DISPATCH_GROUPS, DISPATCH_ROWS))


def dispatch_name(d):
    """'raster', 'groups' or 'rows', with '+queues' for the persistent launch."""
    return ("raster", "groups", "rows")[d & 3] + ("+queues" if d & DISPATCH_QUEUES else "")

DEFAULT_BRICKS = (34, 66, 130, 258)   # stored sizes: cores 32, 64, 128, 256 + 1-voxel overlap


def tune_brick_size(ren, rotations, candidates=DEFAULT_BRICKS, overlap=1, max_lods=1, reps=3):
    """Renders `rotations` with every candidate brick size (cubic, stored size
    incl. overlap) and re-bricks with the fastest.  Returns (best size,
    {size: median ms})."""
    times = {}
    for b in candidates:
        ren.Rebrick((b, b, b), overlap, max_lods)
        ren.PaintViews(rotations, download=False)   # warm-up (plans, uploads)
        ms = []
        for _ in range(reps):
            ren.PaintViews(rotations, download=False)
            ms.append(ren.last_stats.ms)
        times[b] = float(np.median(ms))
    best = min(times, key=times.get)
    ren.Rebrick((best, best, best), overlap, max_lods)
    return best, times


def tune_dispatch_order(ren, rotations, orders=DISPATCH_SETTINGS, reps=3):
    """Renders `rotations` with every dispatch setting (order x work queues;
    the renderer's current mode, TF and step settings) and keeps the fastest
    launch.  Frames are the same in every setting; long rays behind many
    non-empty bricks favour the cost orders, dense early-terminating volumes
    the raster one.  Returns (best setting, {setting: median ms})."""
    times = {}
    for o in orders:
        ren.SetDispatchOrder(o)
        ren.PaintViews(rotations, download=False)   # warm-up
        ms = []
        for _ in range(reps):
            ren.PaintViews(rotations, download=False)
            ms.append(ren.last_stats.ms)
        times[o] = float(np.median(ms))
    best = min(times, key=times.get)
    ren.SetDispatchOrder(best)
    return best, times
