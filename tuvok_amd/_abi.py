"""ctypes view of the C ABI in include/tvk_hip.h (libtvkhip.so)."""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libtvkhip.so")

TVK_OK = 0
TVK_U8, TVK_U16 = 1, 2
RM_1DTRANS, RM_2DTRANS, RM_ISOSURFACE = 0, 1, 2
GEN_BLOBS, GEN_FBM_SHELLS, GEN_ELLIPSOID_SDF, GEN_PHANTOM = 1, 2, 3, 4
DISPATCH_RASTER, DISPATCH_GROUPS, DISPATCH_ROWS, DISPATCH_QUEUES = 0, 1, 2, 4

# every symbol include/tvk_hip.h declares
EXPORTS = [
    "tvk_create", "tvk_destroy", "tvk_last_error", "tvk_default_params", "tvk_load_volume",
    "tvk_generate_volume", "tvk_share_dataset", "tvk_rebrick", "tvk_download_volume", "tvk_dataset_info", "tvk_lod_info", "tvk_brick_info",
    "tvk_brick_md", "tvk_brick_minmax", "tvk_get_brick", "tvk_get_pool", "tvk_max_gradient",
    "tvk_set_tf1d", "tvk_set_tf2d", "tvk_set_view_matrices", "tvk_set_camera", "tvk_set_rotation",
    "tvk_get_view_matrices", "tvk_set_params", "tvk_get_params", "tvk_current_lod", "tvk_plan",
    "tvk_render", "tvk_render_views", "tvk_render_views_noalloc", "tvk_set_count_samples", "tvk_touched_bytes", "tvk_cell_grid", "tvk_frame_device_ptr", "tvk_display",
    "tvk_event_record", "tvk_event_elapsed", "tvk_synchronize",
    "tvk_default_tuning", "tvk_batch_create", "tvk_batch_destroy", "tvk_batch_last_error", "tvk_batch_clear",
    "tvk_batch_add", "tvk_batch_set_tuning", "tvk_batch_get_tuning", "tvk_batch_set_count_samples",
    "tvk_batch_render", "tvk_batch_frame", "tvk_batch_frame_device_ptr", "tvk_batch_autotune",
    "tvk_stepnet_create", "tvk_stepnet_destroy", "tvk_stepnet_last_error", "tvk_stepnet_param_count",
    "tvk_stepnet_set_params", "tvk_stepnet_get_params", "tvk_stepnet_infer", "tvk_stepnet_set_data",
    "tvk_stepnet_grad", "tvk_stepnet_train", "tvk_set_step_proposal", "tvk_generate_segments",
    "tvk_batch_set_step_proposal", "tvk_leap_set_resident", "tvk_leap_render", "tvk_leap_missing",
    "tvk_leap_page_in", "tvk_leap_resume_state",
    "tvk_set_bricks", "tvk_upload_brick", "tvk_set_pool_budget", "tvk_get_pool_stats",
    "tvk_leap_set_request_table", "tvk_leap_requests", "tvk_pool_resident",
    "tvk_rasterize_tf2d", "tvk_set_tf2d_swatches", "tvk_set_clip_plane", "tvk_crop_dataset", "tvk_pick", "tvk_set_compat", "tvk_set_lod_filter", "tvk_set_dispatch_order", "tvk_open_uvf", "tvk_touched_mask",
]


class Params(C.Structure):
    _fields_ = [("mode", C.c_int), ("lighting", C.c_int), ("srm", C.c_float), ("isovalue", C.c_float),
                ("lod", C.c_int), ("tf_scale", C.c_float), ("grad_scale", C.c_float),
                ("domain_scale", C.c_float * 3), ("light_ambient", C.c_float * 3),
                ("light_diffuse", C.c_float * 3), ("light_specular", C.c_float * 3),
                ("light_dir", C.c_float * 3), ("iso_color", C.c_float * 3), ("refine_steps", C.c_int),
                ("clamp_to_edge", C.c_int), ("ert_eps", C.c_float), ("screen_space_error", C.c_float)]


class BrickRec(C.Structure):
    _fields_ = [("key_index", C.c_uint64), ("toc_index", C.c_uint64), ("center", C.c_float * 3),
                ("extension", C.c_float * 3), ("voxel_count", C.c_uint32 * 3), ("tex_min", C.c_float * 3),
                ("tex_max", C.c_float * 3), ("distance", C.c_float), ("is_empty", C.c_int)]


class DatasetDesc(C.Structure):
    _fields_ = [("lod_count", C.c_int), ("brick_count", C.c_uint64), ("pool_bytes", C.c_uint64),
                ("brick_bytes", C.c_uint64),
                ("range_min", C.c_double), ("range_max", C.c_double), ("bit_width", C.c_int),
                ("dims", C.c_uint64 * 3), ("brick_size", C.c_uint32 * 3), ("overlap", C.c_uint32),
                ("build_ms", C.c_float)]


class Tuning(C.Structure):
    _fields_ = [("tile_w", C.c_int), ("tile_h", C.c_int), ("groups_per_cu", C.c_int), ("use_graph", C.c_int)]


class ProposalParams(C.Structure):
    _fields_ = [("margin", C.c_float), ("skip_threshold", C.c_float), ("max_log2", C.c_int),
                ("fallback_bound", C.c_float), ("skip_bound", C.c_float), ("fixed_log2", C.c_int),
                ("fallback_tv", C.c_float)]


class LeapStats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("missing", C.c_uint64), ("unfinished", C.c_uint64), ("ms", C.c_float)]


class BrickKey(C.Structure):
    _fields_ = [("lod", C.c_uint32), ("index", C.c_uint64), ("minmax", C.c_double * 2)]


class PoolStats(C.Structure):
    _fields_ = [("pool_bytes", C.c_uint64), ("slots", C.c_uint64), ("slot_bytes", C.c_uint64),
                ("resident", C.c_uint64), ("uploads", C.c_uint64), ("evictions", C.c_uint64),
                ("upload_bytes", C.c_uint64), ("upload_ms", C.c_float), ("parts", C.c_uint32)]


class UvfInfo(C.Structure):
    _fields_ = [("file_version", C.c_uint64), ("octree_version", C.c_uint32), ("octree_offset", C.c_uint64),
                ("dims", C.c_uint64 * 3), ("brick_size", C.c_uint32 * 3), ("aspect", C.c_double * 3),
                ("overlap", C.c_uint32), ("bit_width", C.c_int), ("lod_count", C.c_int),
                ("brick_count", C.c_uint64), ("zlib_bricks", C.c_uint64)]


class GradientStop(C.Structure):
    _fields_ = [("pos", C.c_float), ("rgba", C.c_float * 4)]


class Swatch(C.Structure):
    _fields_ = [("radial", C.c_int), ("first_point", C.c_uint32), ("n_points", C.c_uint32),
                ("gradient", C.c_float * 4), ("first_stop", C.c_uint32), ("n_stops", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("bricks_in_list", C.c_uint32), ("bricks_nonempty", C.c_uint32),
                ("ms", C.c_float)]


_lib = None


def load(path=LIB_PATH):
    """Loads libtvkhip.so.  Raises if it has not been built: there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing; build it with `python -m tuvok_amd.build`")
    L = C.CDLL(path)
    vp, u64p, u32p = C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)
    fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
    sig = {
        "tvk_create": (C.c_int, [C.POINTER(vp), C.c_int]),
        "tvk_destroy": (None, [vp]),
        "tvk_last_error": (C.c_char_p, [vp]),
        "tvk_default_params": (C.c_int, [C.POINTER(Params)]),
        "tvk_load_volume": (C.c_int, [vp, vp, C.c_int, u64p, u32p, C.c_uint32, C.c_int, C.c_int, C.c_int]),
        "tvk_generate_volume": (C.c_int, [vp, C.c_int, C.c_uint64, C.c_int, u64p, u32p, C.c_uint32, C.c_int,
                                          C.c_int, C.c_int]),
        "tvk_download_volume": (C.c_int, [vp, vp, C.c_size_t]),
        "tvk_share_dataset": (C.c_int, [vp, vp]),
        "tvk_dataset_info": (C.c_int, [vp, C.POINTER(DatasetDesc)]),
        "tvk_lod_info": (C.c_int, [vp, C.c_int, u64p, u64p, u64p, dp]),
        "tvk_brick_info": (C.c_int, [vp, C.c_uint64, u64p, u32p, u64p]),
        "tvk_brick_md": (C.c_int, [vp, C.c_int, C.c_uint64, fp, fp, u32p]),
        "tvk_brick_minmax": (C.c_int, [vp, C.c_uint64, dp]),
        "tvk_get_brick": (C.c_int, [vp, C.c_uint64, vp, C.c_size_t]),
        "tvk_get_pool": (C.c_int, [vp, vp, C.c_size_t]),
        "tvk_max_gradient": (C.c_int, [vp, fp]),
        "tvk_set_tf1d": (C.c_int, [vp, C.POINTER(C.c_uint8), C.c_int]),
        "tvk_set_tf2d": (C.c_int, [vp, C.POINTER(C.c_uint8), C.c_int, C.c_int]),
        "tvk_set_view_matrices": (C.c_int, [vp, fp, fp, fp, fp, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                            C.c_float]),
        "tvk_set_camera": (C.c_int, [vp, fp, fp, fp, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_uint32]),
        "tvk_set_rotation": (C.c_int, [vp, fp]),
        "tvk_set_clip_plane": (C.c_int, [vp, fp, C.c_int]),
        "tvk_set_compat": (C.c_int, [vp, C.c_int, C.c_int]),
        "tvk_set_lod_filter": (C.c_int, [vp, C.c_int]),
        "tvk_set_dispatch_order": (C.c_int, [vp, C.c_int]),
        "tvk_open_uvf": (C.c_int, [vp, C.c_char_p, C.c_uint64, C.c_void_p]),
        "tvk_touched_mask": (C.c_int, [vp, C.POINTER(C.c_uint8), C.c_size_t]),
        "tvk_crop_dataset": (C.c_int, [vp, C.c_int, fp]),
        "tvk_pick": (C.c_int, [vp, C.c_uint32, C.c_uint32, fp]),
        "tvk_get_view_matrices": (C.c_int, [vp, fp, fp]),
        "tvk_set_params": (C.c_int, [vp, C.POINTER(Params)]),
        "tvk_get_params": (C.c_int, [vp, C.POINTER(Params)]),
        "tvk_current_lod": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "tvk_plan": (C.c_int, [vp, C.POINTER(BrickRec), C.c_int, C.POINTER(C.c_int)]),
        "tvk_render": (C.c_int, [vp, fp, C.POINTER(Stats)]),
        "tvk_render_views": (C.c_int, [vp, C.c_int, fp, fp, C.POINTER(Stats)]),
        "tvk_render_views_noalloc": (C.c_int, [vp, C.c_int, fp, vp, C.c_size_t, C.POINTER(Stats)]),
        "tvk_set_count_samples": (C.c_int, [vp, C.c_int]),
        "tvk_touched_bytes": (C.c_int, [vp, u64p, u64p, u64p, u64p]),
        "tvk_cell_grid": (C.c_int, [vp, C.POINTER(C.c_uint32), C.c_uint64, u64p]),
        "tvk_rebrick": (C.c_int, [vp, C.POINTER(C.c_uint32), C.c_uint32, C.c_int]),
        "tvk_display": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint8), C.c_size_t]),
        "tvk_frame_device_ptr": (C.c_int, [vp, C.POINTER(vp)]),
        "tvk_event_record": (C.c_int, [vp, C.c_int]),
        "tvk_event_elapsed": (C.c_int, [vp, C.c_int, C.c_int, fp]),
        "tvk_synchronize": (C.c_int, [vp]),
        "tvk_default_tuning": (C.c_int, [C.POINTER(Tuning)]),
        "tvk_batch_create": (C.c_int, [C.POINTER(vp), C.c_int]),
        "tvk_batch_destroy": (None, [vp]),
        "tvk_batch_last_error": (C.c_char_p, [vp]),
        "tvk_batch_clear": (C.c_int, [vp]),
        "tvk_batch_add": (C.c_int, [vp, vp, fp, C.c_uint32, C.c_uint32, C.POINTER(C.c_int)]),
        "tvk_batch_set_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "tvk_batch_get_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "tvk_batch_set_count_samples": (C.c_int, [vp, C.c_int]),
        "tvk_batch_render": (C.c_int, [vp, C.POINTER(Stats)]),
        "tvk_batch_frame": (C.c_int, [vp, C.c_int, fp]),
        "tvk_batch_frame_device_ptr": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
        "tvk_batch_autotune": (C.c_int, [vp, C.c_int, C.POINTER(Tuning), fp]),
        "tvk_stepnet_create": (C.c_int, [C.POINTER(vp), C.c_int, C.c_uint64]),
        "tvk_stepnet_destroy": (None, [vp]),
        "tvk_stepnet_last_error": (C.c_char_p, [vp]),
        "tvk_stepnet_param_count": (C.c_int, []),
        "tvk_stepnet_set_params": (C.c_int, [vp, fp, C.c_int]),
        "tvk_stepnet_get_params": (C.c_int, [vp, fp, C.c_int]),
        "tvk_stepnet_infer": (C.c_int, [vp, fp, C.c_int, fp]),
        "tvk_stepnet_set_data": (C.c_int, [vp, fp, fp, C.c_int]),
        "tvk_stepnet_grad": (C.c_int, [vp, C.c_int, fp]),
        "tvk_stepnet_train": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint64, fp]),
        "tvk_set_step_proposal": (C.c_int, [vp, vp, C.POINTER(ProposalParams)]),
        "tvk_batch_set_step_proposal": (C.c_int, [vp, vp, C.POINTER(ProposalParams)]),
        "tvk_leap_set_resident": (C.c_int, [vp, C.POINTER(C.c_uint8), C.c_size_t]),
        "tvk_leap_render": (C.c_int, [vp, C.c_int, fp, C.POINTER(LeapStats)]),
        "tvk_leap_missing": (C.c_int, [vp, C.POINTER(C.c_uint8), C.c_size_t, u64p]),
        "tvk_leap_page_in": (C.c_int, [vp, C.c_uint64, u64p]),
        "tvk_leap_resume_state": (C.c_int, [vp, fp, C.c_size_t]),
        "tvk_set_bricks": (C.c_int, [vp, C.c_int, u64p, u32p, C.c_uint32, C.c_int, C.POINTER(BrickKey), C.c_size_t,
                                     C.c_uint64]),
        "tvk_upload_brick": (C.c_int, [vp, C.c_uint32, C.c_uint64, vp, C.c_size_t]),
        "tvk_set_pool_budget": (C.c_int, [vp, C.c_uint64]),
        "tvk_leap_set_request_table": (C.c_int, [vp, C.c_uint32, C.c_uint32]),
        "tvk_leap_requests": (C.c_int, [vp, u32p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "tvk_pool_resident": (C.c_int, [vp, C.POINTER(C.c_uint8), C.c_size_t]),
        "tvk_rasterize_tf2d": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_uint8), C.c_int, C.POINTER(Swatch), C.c_int,
                                         fp, C.c_size_t, C.POINTER(GradientStop), C.c_size_t, C.POINTER(C.c_uint8)]),
        "tvk_set_tf2d_swatches": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_uint8), C.c_int, C.POINTER(Swatch),
                                            C.c_int, fp, C.c_size_t, C.POINTER(GradientStop), C.c_size_t]),
        "tvk_get_pool_stats": (C.c_int, [vp, C.POINTER(PoolStats)]),
        "tvk_generate_segments": (C.c_int, [vp, C.c_int, fp, C.c_uint32, C.c_float, C.c_uint32, fp,
                                            C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class TvkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tvk error {code}: {msg}")
        self.code = code
