"""The 64 GiB volume (4096x4096x2048 uint16 fBm + shells, its whole pyramid,
bricks of 258^3) held within an HBM budget below the pyramid: GridLeaper
passes request the missing bricks through the 509-entry hash table, the bricks
are paged from host memory into the budget's slots, and the passes converge to
the frame of the all-resident pool (lit 2D TF, 1920x1080)."""
import time

import numpy as np
import pytest

from helpers import err_stats
from tuvok_amd import workloads
from tuvok_amd.renderer import rotation_y
from tuvok_amd.transfer import tf2d_gaussian_lobes

pytestmark = pytest.mark.gpu


def test_budget_below_the_pyramid_converges_to_the_resident_frame():
    cfg = dict(workloads.LOD_64G)
    cfg["max_lods"] = 0          # the traversal reads the whole pyramid (down to one voxel)
    t0 = time.perf_counter()
    print(f"lod64g: generating {cfg['dims']} with {cfg['max_lods']} levels", flush=True)
    r = workloads.lod_owner(cfg)
    try:
        d = r.GetDatasetDesc()
        print(f"lod64g: {d.brick_count} bricks in {time.perf_counter() - t0:.1f} s", flush=True)
        r.Set2DTrans(tf2d_gaussian_lobes().GetByteArray())
        r.SetRendermode(cfg["mode"])
        r.SetUseLighting(cfg["lighting"])
        g = r.MaxGradientMagnitude()
        r.SetGradientScale(1.0 / g if g > 0 else 1.0)
        r.SetViewParameters(cfg["fov"], 0.01, 1000.0)
        r.SetViewPosition((0.0, 0.0, 1.15))
        r.Resize(1920, 1080)
        r.SetRotation(rotation_y(0.7))
        full, fst = r.LeapConverge(max_passes=4)
        assert len(fst) == 1 and fst[0].missing == 0          # every brick resident: one pass
        assert full[..., 3].max() > 0.5
        pyramid = d.brick_bytes                                 # reference-layout bytes of all levels
        budget = 48 << 30
        assert budget < pyramid
        t1 = time.perf_counter()
        r.SetPoolBudget(budget)
        st = r.GetPoolStats()
        print(f"lod64g: budget {budget / 2**30:.0f} GiB of a {pyramid / 2**30:.1f} GiB pyramid: {st.slots} slots of "
              f"{st.slot_bytes / 2**20:.1f} MiB ({time.perf_counter() - t1:.1f} s to the host store)", flush=True)
        assert st.slots * st.slot_bytes <= budget and st.slots < d.brick_count
        r.LeapSetResident(None)                                # only the coarsest brick
        img, stats = r.LeapConverge(max_passes=64)
        st = r.GetPoolStats()
        for k, s in enumerate(stats):
            print(f"lod64g: pass {k}: missing {s.missing} unfinished {s.unfinished} {s.ms:.1f} ms", flush=True)
        mx, mean = err_stats(img, full)
        print(f"lod64g: converged after {len(stats)} passes, {st.uploads} uploads "
              f"({st.upload_bytes / 2**30:.1f} GiB, {st.upload_ms:.0f} ms), {st.evictions} evictions: "
              f"max {mx:.3g} mean {mean:.3g}", flush=True)
        assert len(stats) > 1 and stats[-1].missing == 0
        assert st.uploads > 0
        assert mx <= 2.0 / 255 and mean < 1e-3
    finally:
        r.close()
