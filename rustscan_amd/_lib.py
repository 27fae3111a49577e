"""ctypes binding of librgs_hip.so (the C ABI declared in include/rgs.h).

The library is the only implementation of the render/train path: there is no
CPU fallback.  Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_float, c_int, c_int64, c_uint32, c_void_p

from .build import LIB_PATH

RGS_OK = 0
RGS_ERR_INVALID = 1
RGS_ERR_CAPACITY = 2
RGS_ERR_HIP = 3
RGS_ERR_STATE = 4
RGS_ERR_UNSUPPORTED = 5

RGS_MODE_PER_VIEW = 0
RGS_MODE_BATCHED = 1
RGS_SORT_NOALLOC_MAX_PAIRS = 2097152

RGS_DENSIFY_ACCUMULATE = 1
RGS_DENSIFY_ACTUAL_VISIBILITY = 2
RGS_DENSIFY_COLLECT_VISIBILITY = 4
DENSIFY_FIELDS = ["grad_2d", "screen_grad_2d", "abs_grad_2d", "abs_pixel_grad_2d", "pixel_coverage", "camera_depth",
                  "grad_color", "num_observations", "visible_observations", "actual_visible_observations"]

RGS_GRAPH_OFF = 0
RGS_GRAPH_ON = 1

RGS_DEBUG_DEPTH_ORDER = 0
RGS_DEBUG_ISECT_TILES = 1
RGS_DEBUG_ISECT_GIDS = 2
RGS_DEBUG_TILE_RANGES = 3
RGS_DEBUG_PROJECTED = 4
RGS_DEBUG_TILE_COUNTS = 5
RGS_DEBUG_IMAGE = 6
RGS_DEBUG_RASTER_GRADS = 7

STAGES = ["project_fwd", "depth_sort", "count_scan", "emit", "tile_sort", "tile_ranges", "raster_fwd", "raster_bwd",
          "project_bwd", "adam"]


class RgsCamera(ctypes.Structure):
    _fields_ = [
        ("viewmat", c_float * 16),
        ("fx", c_float),
        ("fy", c_float),
        ("cx", c_float),
        ("cy", c_float),
        ("width", c_uint32),
        ("height", c_uint32),
        ("cam_pos", c_float * 3),
    ]


class RgsAdamCfg(ctypes.Structure):
    _fields_ = [
        ("beta1", c_float),
        ("beta2", c_float),
        ("eps", c_float),
        ("lr_transforms", c_float * 10),
        ("lr_sh", c_float * 25),
        ("lr_opacity", c_float),
        ("weight_decay", c_float),
    ]


class RgsLossCfg(ctypes.Structure):
    _fields_ = [
        ("combined", c_int),
        ("l1_weight", c_float),
        ("ssim_weight", c_float),
        ("gradient_weight", c_float),
        ("robust_delta", c_float),
        ("outlier_threshold", c_float),
        ("outlier_weight", c_float),
    ]


class RgsStepCfg(ctypes.Structure):
    _fields_ = [
        ("mode", c_int),
        ("graph", c_int),
        ("cov_blur", c_float),
        ("background", c_float * 3),
        ("loss_weight", c_float),
        ("check_loss", c_int),
        ("adam", RgsAdamCfg),
        ("densify", c_int),
        ("loss", RgsLossCfg),
    ]


class RgsTopologyCfg(ctypes.Structure):
    _fields_ = [
        ("iterations", c_uint32),
        ("max_initial_gaussians", c_uint32),
        ("frame_shuffle_seed", ctypes.c_uint64),
        ("refine_every", c_uint32),
        ("opacity_reset_interval", c_uint32),
        ("opacity_reset_mode", c_uint32),
        ("topology_freeze_after_epoch", ctypes.c_int32),
        ("growth_freeze_after_epoch", ctypes.c_int32),
        ("prune_until_epoch", ctypes.c_int32),
        ("split_score_mode", c_uint32),
        ("growth_stop_iter", c_uint32),
        ("growth_grad_threshold", c_float),
        ("split_grad_threshold", c_float),
        ("depth_scale_gamma", c_float),
        ("growth_select_fraction", c_float),
        ("opacity_decay", c_float),
        ("scale_decay", c_float),
        ("prune_mode", c_uint32),
        ("prune_min_age", c_uint32),
        ("prune_invisible_epochs", c_uint32),
        ("prune_visibility_dry_run", c_uint32),
        ("prune_opacity_threshold", c_float),
        ("prune_visibility_threshold", c_float),
        ("prune_high_opacity_threshold", c_float),
    ]


class RgsTopologySchedule(ctypes.Structure):
    _fields_ = [("completed_epoch", ctypes.c_int32), ("densify", ctypes.c_int32), ("prune", ctypes.c_int32),
                ("reset_opacity", ctypes.c_int32), ("allow_extra_growth", ctypes.c_int32)]


class RgsTopologyResult(ctypes.Structure):
    _fields_ = [
        ("schedule", RgsTopologySchedule),
        ("before", c_int64), ("after", c_int64), ("added", c_int64), ("pruned", c_int64),
        ("clone_candidates", c_int64), ("split_candidates", c_int64), ("growth_candidates", c_int64),
        ("prune_candidates", c_int64), ("replacement_count", c_int64), ("extra_growth_count", c_int64),
        ("selected_count", c_int64), ("mutated", ctypes.c_int32), ("adam_reset", ctypes.c_int32),
    ]


class RgsRasterCfg(ctypes.Structure):
    _fields_ = [("tile_width", c_uint32), ("batch_depth", c_uint32), ("persistent", c_uint32)]


class RgsCounts(ctypes.Structure):
    _fields_ = [("num_visible", c_int64), ("num_intersections", c_int64), ("num_tiles", c_int64)]


class RgsFrameMetrics(ctypes.Structure):
    _fields_ = [("psnr_db", c_float), ("render_grad", c_float), ("target_grad", c_float), ("render_lap", c_float),
                ("target_lap", c_float), ("sharpness_grad_ratio", c_float), ("sharpness_lap_ratio", c_float)]


class RgsSplatMetadata(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("final_loss", c_float), ("gaussian_count", ctypes.c_uint64),
                ("sh_degree", c_uint32)]


class RgsError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"rgs error {code}: {message}")
        self.code = code


_LIB = None

_SIGNATURES = {
    "rgs_ctx_create": (c_int, [c_int, POINTER(c_void_p)]),
    "rgs_ctx_destroy": (None, [c_void_p]),
    "rgs_last_error": (ctypes.c_char_p, [c_void_p]),
    "rgs_sync": (c_int, [c_void_p]),
    "rgs_get_stream": (c_void_p, [c_void_p]),
    "rgs_set_raster_config": (c_int, [c_void_p, POINTER(RgsRasterCfg)]),
    "rgs_get_raster_config": (c_int, [c_void_p, POINTER(RgsRasterCfg)]),
    "rgs_splats_create": (c_int, [c_void_p, c_int64, c_int, POINTER(c_void_p)]),
    "rgs_splats_destroy": (None, [c_void_p, c_void_p]),
    "rgs_splats_count": (c_int64, [c_void_p]),
    "rgs_splats_sh_degree": (c_int, [c_void_p]),
    "rgs_splats_import_aos": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "rgs_splats_export_aos": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "rgs_splats_export_grads": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "rgs_splats_export_visible": (c_int, [c_void_p, c_void_p, c_void_p]),
    "rgs_adam_reset": (c_int, [c_void_p, c_void_p]),
    "rgs_densify_export": (c_int, [c_void_p, c_void_p, c_void_p]),
    "rgs_densify_reset": (c_int, [c_void_p, c_void_p]),
    "rgs_densify_import": (c_int, [c_void_p, c_void_p, c_void_p]),
    "rgs_topology_should_apply": (c_int, [POINTER(RgsTopologyCfg), c_uint32, c_uint32]),
    "rgs_topology_get_schedule": (c_int, [POINTER(RgsTopologyCfg), c_uint32, c_uint32, POINTER(RgsTopologySchedule)]),
    "rgs_topology_step": (c_int, [c_void_p, c_void_p, POINTER(RgsTopologyCfg), c_uint32, c_uint32,
                                  POINTER(RgsTopologyResult)]),
    "rgs_topology_export_state": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "rgs_topology_import_state": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "rgs_adam_get_steps": (c_int, [c_void_p, c_void_p, c_void_p]),
    "rgs_adam_set_steps": (c_int, [c_void_p, c_void_p, c_void_p]),
    "rgs_adam_export_moments": (c_int, [c_void_p, c_void_p] + [c_void_p] * 6),
    "rgs_adam_import_moments": (c_int, [c_void_p, c_void_p] + [c_void_p] * 6),
    "rgs_tiled_render": (c_int, [c_void_p, c_void_p, POINTER(RgsCamera), POINTER(c_float), c_float, c_void_p,
                                 c_void_p]),
    "rgs_render_splats": (c_int, [c_void_p, c_void_p, POINTER(RgsCamera), POINTER(c_float), c_float, c_void_p,
                                  c_void_p]),
    "rgs_image_metrics": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, POINTER(RgsFrameMetrics)]),
    "rgs_evaluate_frame": (c_int, [c_void_p, c_void_p, POINTER(RgsCamera), c_float, c_void_p, c_void_p,
                                   POINTER(RgsFrameMetrics)]),
    "rgs_save_splats": (c_int, [c_void_p, ctypes.c_char_p, c_int64, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, POINTER(RgsSplatMetadata)]),
    "rgs_scene_load": (c_int, [c_void_p, ctypes.c_char_p, POINTER(c_void_p)]),
    "rgs_scene_info": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_int), POINTER(RgsSplatMetadata)]),
    "rgs_scene_get": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "rgs_scene_free": (None, [c_void_p]),
    "rgs_scene_to_splats": (c_int, [c_void_p, c_void_p, POINTER(c_void_p)]),
    "rgs_splats_save": (c_int, [c_void_p, c_void_p, ctypes.c_char_p, POINTER(RgsSplatMetadata)]),
    "rgs_io_last_error": (ctypes.c_char_p, []),
    "rgs_render_forward": (c_int, [c_void_p, c_void_p, POINTER(RgsCamera), c_int, POINTER(c_float), c_float,
                                   POINTER(c_void_p)]),
    "rgs_render_backward": (c_int, [c_void_p, c_void_p, POINTER(c_void_p)]),
    "rgs_checkpoint_create": (c_int, [c_void_p, POINTER(c_void_p)]),
    "rgs_checkpoint_destroy": (None, [c_void_p, c_void_p]),
    "rgs_render_forward_checkpoint": (c_int, [c_void_p, c_void_p, POINTER(RgsCamera), c_int, POINTER(c_float), c_float,
                                              POINTER(c_void_p), c_void_p]),
    "rgs_render_backward_checkpoint": (c_int, [c_void_p, c_void_p, c_void_p, POINTER(c_void_p)]),
    "rgs_l1_loss_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_float, c_void_p,
                                 POINTER(c_float)]),
    "rgs_adam_step": (c_int, [c_void_p, c_void_p, POINTER(RgsAdamCfg)]),
    "rgs_image_loss_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_float, POINTER(RgsLossCfg),
                                    c_void_p, POINTER(c_float)]),
    "rgs_train_step": (c_int, [c_void_p, c_void_p, POINTER(RgsCamera), POINTER(c_void_p), c_int,
                               POINTER(RgsStepCfg), POINTER(c_float)]),
    "rgs_autotune": (c_int, [c_void_p, c_void_p, POINTER(RgsCamera), POINTER(c_void_p), c_int, POINTER(RgsStepCfg),
                             POINTER(RgsRasterCfg)]),
    "rgs_get_counts": (c_int, [c_void_p, POINTER(RgsCounts), c_int]),
    "rgs_skipped_steps": (c_int64, [c_void_p]),
    "rgs_sort_pairs_u32": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64]),
    "rgs_inclusive_scan_u32": (c_int, [c_void_p, c_void_p, c_void_p, c_int64]),
    "rgs_sort_pairs_u32_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64]),
    "rgs_debug_get": (c_int, [c_void_p, c_int, c_int, c_void_p, POINTER(c_int64)]),
    "rgs_set_debug": (c_int, [c_void_p, c_int]),
    "rgs_get_fragment_counts": (c_int, [c_void_p, POINTER(c_int64), c_int]),
    "rgs_set_timing": (c_int, [c_void_p, c_int]),
    "rgs_get_timing": (c_int, [c_void_p, POINTER(ctypes.c_double), POINTER(c_int64), c_int]),
    "rgs_set_launch_limit": (c_int, [c_void_p, c_int64]),
    "rgs_version": (ctypes.c_char_p, []),
}


def lib_path() -> str:
    return LIB_PATH


def lib():
    """Load librgs_hip.so from the package directory (raises if missing)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(ctx, code: int) -> None:
    if code != RGS_OK:
        msg = lib().rgs_last_error(ctx).decode() if ctx else ""
        raise RgsError(code, msg)
