"""tuvok_amd: MI355X-native volume raycasting path of Tuvok.

The product is libtvkhip.so (HIP kernels for gfx950 + a C ABI, include/tvk_hip.h);
this package holds its build script, a ctypes view of the ABI, the
AbstrRenderer-shaped front end (HIPRaycaster), transfer functions with the
reference's byte semantics, and the synthetic workloads.
"""
from ._abi import (DISPATCH_GROUPS, DISPATCH_QUEUES, DISPATCH_RASTER, DISPATCH_ROWS, GEN_BLOBS,  # noqa: F401
                   GEN_ELLIPSOID_SDF, GEN_FBM_SHELLS, GEN_PHANTOM, RM_1DTRANS, RM_2DTRANS, RM_ISOSURFACE, TvkError)
from .batch import RenderBatch  # noqa: F401
from .renderer import HIPRaycaster, rotation_y  # noqa: F401
from .transfer import (TFPolygon, TransferFunction1D, TransferFunction2D, tf1d_blue_white_red,  # noqa: F401
                       tf2d_gaussian_lobes)
