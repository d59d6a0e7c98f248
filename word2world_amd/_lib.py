"""ctypes binding of ``libw2w_raster.so`` (C ABI: include/w2w_raster.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libw2w_raster.so")

W2W_OK, W2W_ERR_INVALID, W2W_ERR_HIP, W2W_ERR_CAPACITY = 0, 1, 2, 3
_STATUS_TEXT = {W2W_ERR_INVALID: "invalid argument", W2W_ERR_HIP: "HIP launch failure",
                W2W_ERR_CAPACITY: "workspace too small"}

_vp = ctypes.c_void_p


class W2WFrame(ctypes.Structure):
    _fields_ = [("P", ctypes.c_int32), ("M", ctypes.c_int32), ("D", ctypes.c_int32),
                ("W", ctypes.c_int32), ("H", ctypes.c_int32), ("tile", ctypes.c_int32),
                ("chunk", ctypes.c_int32), ("reserved", ctypes.c_int32), ("focal_x", ctypes.c_float), ("focal_y", ctypes.c_float),
                ("means", _vp), ("scales", _vp), ("rotations", _vp), ("opacities", _vp),
                ("sh", _vp), ("viewmatrix", _vp), ("projmatrix", _vp), ("campos", _vp),
                ("background", _vp)]


class W2WWorkspace(ctypes.Structure):
    _fields_ = [("geom", _vp), ("geom_bytes", ctypes.c_size_t), ("binning", _vp),
                ("binning_bytes", ctypes.c_size_t), ("image", _vp), ("image_bytes", ctypes.c_size_t),
                ("capacity", ctypes.c_int64)]


class W2WLayout(ctypes.Structure):
    _fields_ = [(n, ctypes.c_size_t) for n in (
        "geom_bytes", "binning_bytes", "image_bytes", "splats", "radii", "rect", "counts", "perm",
        "offsets", "item_off", "sorted_ids", "sorted_tiles", "ranges", "n_contrib", "final_T",
        "item_frame", "table", "grad_rec", "grad_flag", "item_grad", "tile_key_bytes")]


class W2WCaps(ctypes.Structure):
    _fields_ = [("scenes", ctypes.c_int32), ("frames", ctypes.c_int32), ("tile", ctypes.c_int32),
                ("chunk", ctypes.c_int32), ("reserved", ctypes.c_int32), ("depth_split", ctypes.c_int32), ("items", ctypes.c_int64), ("gaussians", ctypes.c_int64),
                ("tiles", ctypes.c_int64), ("pixels", ctypes.c_int64), ("capacity", ctypes.c_int64)]


class W2WStoreGroups(ctypes.Structure):
    _fields_ = [("n_groups", ctypes.c_int32), ("width", ctypes.c_int32 * 8), ("means", ctypes.c_int32),
                ("opacity", ctypes.c_int32), ("svec", ctypes.c_int32), ("quat", ctypes.c_int32),
                ("param", ctypes.c_void_p * 8), ("exp_avg", ctypes.c_void_p * 8),
                ("exp_avg_sq", ctypes.c_void_p * 8)]


class W2WDensifyConfig(ctypes.Structure):
    _fields_ = [("grad_thresh", ctypes.c_float), ("min_opacity", ctypes.c_float),
                ("clone_scale_max", ctypes.c_float), ("split_scale_min", ctypes.c_float),
                ("prune_scale_max", ctypes.c_float), ("split_shrink", ctypes.c_float),
                ("split_factor", ctypes.c_int32), ("pad", ctypes.c_int32)]


EXPORTS = ["w2w_version", "w2w_splat_extents", "w2w_block_touches", "w2w_workspace_layout", "w2w_batch_workspace_layout",
           "w2w_batch_table_bytes", "w2w_batch_encode", "w2w_forward_batch", "w2w_backward_batch", "w2w_backward_batch_sh_adam", "w2w_forward", "w2w_backward", "w2w_sorted_keys",
           "w2w_sort_temp_bytes", "w2w_sort_pairs_u64", "w2w_scan_temp_bytes",
           "w2w_exclusive_scan_u32", "w2w_exclusive_scan_u32_noalloc", "w2w_prof_enable", "w2w_prof_reset", "w2w_prof_read", "w2w_stats_enable", "w2w_adam_step", "w2w_activate", "w2w_activate_backward",
           "w2w_l2_loss", "w2w_adam_step_hyper", "w2w_activate_rows", "w2w_activate_backward_rows",
           "w2w_adam_ranges", "w2w_densify_temp_bytes", "w2w_densify_plan", "w2w_densify_apply",
           "w2w_densify_stats", "w2w_voxel_temp_bytes", "w2w_voxel_field"]

KERNEL_IDS = {0: "fill", 1: "preprocess", 2: "depth_sort_hist", 3: "depth_sort_pass", 4: "scan",
              5: "emit", 6: "tile_sort_hist", 7: "tile_sort_pass", 8: "ranges", 9: "blend",
              10: "blend_bwd", 11: "preprocess_bwd", 12: "adam", 13: "sort_hist", 14: "sort_pass",
              15: "activate", 16: "loss", 17: "activate_bwd", 18: "sh_bwd", 19: "densify",
              20: "voxel", 21: "count_open", 22: "emit_far", 23: "blend_far",
              24: "tile_order"}

_lib = None
ABI_VERSION = 9  # W2W_ABI_VERSION of include/w2w_raster.h


def load(path: str = LIB_PATH):
    """Load the library (does not touch the GPU)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with word2world_amd.build.build() "
                           "(the rasterizer has no fallback path)")
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    L.w2w_version.restype = ctypes.c_int
    L.w2w_version.argtypes = []
    L.w2w_block_touches.restype = ctypes.c_int
    L.w2w_block_touches.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    L.w2w_splat_extents.restype = ctypes.c_int
    L.w2w_splat_extents.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    L.w2w_workspace_layout.restype = ctypes.c_int
    L.w2w_workspace_layout.argtypes = [ctypes.c_int32] * 4 + [ctypes.c_int64, P(W2WLayout)]
    L.w2w_batch_workspace_layout.restype = ctypes.c_int
    L.w2w_batch_workspace_layout.argtypes = [P(W2WCaps), P(W2WLayout)]
    L.w2w_batch_table_bytes.restype = ctypes.c_size_t
    L.w2w_batch_table_bytes.argtypes = [ctypes.c_int32, ctypes.c_int32]
    L.w2w_batch_encode.restype = ctypes.c_int
    L.w2w_batch_encode.argtypes = [_vp, ctypes.c_size_t, P(W2WCaps), ctypes.c_int32] + [_vp] * 14 + [_vp]
    L.w2w_forward_batch.restype = ctypes.c_int
    L.w2w_forward_batch.argtypes = [_vp, P(W2WCaps), P(W2WWorkspace)] + [_vp] * 10
    L.w2w_backward_batch.restype = ctypes.c_int
    L.w2w_backward_batch.argtypes = [_vp, P(W2WCaps), P(W2WWorkspace)] + [_vp] * 16
    L.w2w_backward_batch_sh_adam.restype = ctypes.c_int
    L.w2w_backward_batch_sh_adam.argtypes = ([_vp, P(W2WCaps), P(W2WWorkspace)] + [_vp] * 17
                                             + [ctypes.c_int32, ctypes.c_int32, _vp, _vp])
    L.w2w_forward.restype = ctypes.c_int
    L.w2w_forward.argtypes = [P(W2WFrame), P(W2WWorkspace), _vp, _vp, _vp, _vp, _vp]
    L.w2w_backward.restype = ctypes.c_int
    L.w2w_backward.argtypes = [P(W2WFrame), P(W2WWorkspace)] + [_vp] * 12
    L.w2w_sorted_keys.restype = ctypes.c_int
    L.w2w_sorted_keys.argtypes = [P(W2WFrame), P(W2WWorkspace), _vp, _vp]
    L.w2w_sort_temp_bytes.restype = ctypes.c_size_t
    L.w2w_sort_temp_bytes.argtypes = [ctypes.c_int64, ctypes.c_int32]
    L.w2w_sort_pairs_u64.restype = ctypes.c_int
    L.w2w_sort_pairs_u64.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_int32, _vp,
                                     ctypes.c_size_t, _vp]
    L.w2w_scan_temp_bytes.restype = ctypes.c_size_t
    L.w2w_scan_temp_bytes.argtypes = [ctypes.c_int64]
    L.w2w_exclusive_scan_u32.restype = ctypes.c_int
    L.w2w_exclusive_scan_u32.argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp, ctypes.c_size_t, _vp]
    L.w2w_exclusive_scan_u32_noalloc.restype = ctypes.c_int
    L.w2w_exclusive_scan_u32_noalloc.argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp]
    L.w2w_prof_enable.restype = ctypes.c_int
    L.w2w_prof_enable.argtypes = [ctypes.c_int32, ctypes.c_uint32]
    L.w2w_prof_reset.restype = ctypes.c_int
    L.w2w_prof_reset.argtypes = []
    L.w2w_prof_read.restype = ctypes.c_int
    L.w2w_prof_read.argtypes = [_vp, _vp, ctypes.c_int32, P(ctypes.c_int32)]
    L.w2w_stats_enable.restype = ctypes.c_int
    L.w2w_stats_enable.argtypes = [_vp]
    L.w2w_adam_step.restype = ctypes.c_int
    L.w2w_adam_step.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_int32, _vp, _vp,
                                ctypes.c_int64, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp]
    L.w2w_activate.restype = ctypes.c_int
    L.w2w_activate.argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp, _vp, _vp, _vp]
    L.w2w_activate_backward.restype = ctypes.c_int
    L.w2w_activate_backward.argtypes = [_vp] * 6 + [ctypes.c_int64] + [_vp] * 4
    L.w2w_l2_loss.restype = ctypes.c_int
    L.w2w_l2_loss.argtypes = [_vp, P(W2WCaps), _vp, _vp, _vp, _vp, _vp, _vp]
    L.w2w_adam_step_hyper.restype = ctypes.c_int
    L.w2w_adam_step_hyper.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_int32, _vp, _vp,
                                      _vp, _vp, _vp, _vp, _vp]
    L.w2w_activate_rows.restype = ctypes.c_int
    L.w2w_activate_rows.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp, _vp]
    L.w2w_activate_backward_rows.restype = ctypes.c_int
    L.w2w_activate_backward_rows.argtypes = [_vp] * 7 + [ctypes.c_int32, ctypes.c_int64] + [_vp] * 4
    L.w2w_adam_ranges.restype = ctypes.c_int
    L.w2w_adam_ranges.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int64, _vp,
                                  ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp]
    L.w2w_densify_temp_bytes.restype = ctypes.c_size_t
    L.w2w_densify_temp_bytes.argtypes = [ctypes.c_int64]
    L.w2w_densify_plan.restype = ctypes.c_int
    L.w2w_densify_plan.argtypes = [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp, _vp]
    L.w2w_densify_apply.restype = ctypes.c_int
    L.w2w_densify_apply.argtypes = [_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int64, _vp,
                                    ctypes.c_int64, _vp]
    L.w2w_voxel_temp_bytes.restype = ctypes.c_size_t
    L.w2w_voxel_temp_bytes.argtypes = [ctypes.c_int64, ctypes.c_int32]
    L.w2w_voxel_field.restype = ctypes.c_int
    L.w2w_voxel_field.argtypes = [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                  _vp, _vp, ctypes.c_size_t, _vp]
    L.w2w_densify_stats.restype = ctypes.c_int
    L.w2w_densify_stats.argtypes = [_vp] * 6 + [_vp]
    if L.w2w_version() != ABI_VERSION:
        raise RuntimeError(f"{path} has C ABI version {L.w2w_version()}, this binding expects "
                           f"{ABI_VERSION}: rebuild it with word2world_amd.build.build()")
    _lib = L
    return L


def check(status: int, what: str) -> None:
    if status != W2W_OK:
        raise RuntimeError(f"{what} failed: {_STATUS_TEXT.get(status, status)} (status {status})")


def layout(P: int, W: int, H: int, tile: int, capacity: int) -> W2WLayout:
    out = W2WLayout()
    check(load().w2w_workspace_layout(P, W, H, tile, capacity, ctypes.byref(out)),
          "w2w_workspace_layout")
    return out


def batch_layout(caps: W2WCaps) -> W2WLayout:
    out = W2WLayout()
    check(load().w2w_batch_workspace_layout(ctypes.byref(caps), ctypes.byref(out)),
          "w2w_batch_workspace_layout")
    return out


def prof_read(max_records: int = 1 << 16):
    """[(kernel name, ms)] of the records since the last reset (synchronises)."""
    import numpy as np
    ids = np.zeros(max_records, np.int32)
    ms = np.zeros(max_records, np.float32)
    n = ctypes.c_int32(0)
    check(load().w2w_prof_read(ids.ctypes.data, ms.ctypes.data, max_records, ctypes.byref(n)),
          "w2w_prof_read")
    return [(KERNEL_IDS.get(int(i), str(int(i))), float(t)) for i, t in zip(ids[:n.value], ms[:n.value])]
