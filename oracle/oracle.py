"""Python driver for the CPU oracle (``liboracle.so`` built from w2w_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, never by the product package.

``forward`` runs the reference stage order of ``gaussianRasterizer.cu:118-280``:
preprocess (projection.cu, ADRculling.cu, gaussianForwardPass.cu) -> inclusive
scan (gaussianRasterizer.cu:77-117) -> key emission in Gaussian order
(prepareSort.cu:239-301) -> stable LSD sort of the 64-bit keys
(prepareSort.cu:303-360) -> per-tile ranges (gaussianTiles.cu:6-33) -> blend
(pixelGaussiansFull.cu:3-98).  ``backward`` runs the reverse blend
(backward.cu:521-665 structure) and the per-Gaussian chain rule.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_c_int, _c_float, _c_i64 = ctypes.c_int, ctypes.c_float, ctypes.c_int64


def build(force: bool = False) -> str:
    """Compile liboracle.so with oracle/Makefile."""
    if force or not os.path.exists(_LIB_PATH) or (
            os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "w2w_oracle.c"))):
        subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_expf.restype = ctypes.c_float
        L.oracle_expf.argtypes = [ctypes.c_float]
        L.oracle_logf.restype = ctypes.c_float
        L.oracle_logf.argtypes = [ctypes.c_float]
        L.oracle_preprocess.restype = _c_int
        L.oracle_preprocess.argtypes = [
            _c_int, _c_int, _c_int, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
            _c_int, _c_int, _c_int, _c_float, _c_float,
            _f32p, _f32p, _f32p, _f32p, _i32p, _u32p, _u32p, _f32p]
        L.oracle_inclusive_scan.restype = ctypes.c_uint64
        L.oracle_inclusive_scan.argtypes = [_c_int, _u32p, _u32p]
        L.oracle_emit.restype = _c_int
        L.oracle_emit.argtypes = [_c_int, _c_int, _c_int, _u32p, _u32p, _u32p, _f32p, _u64p, _u32p]
        L.oracle_sort_pairs.restype = _c_int
        L.oracle_sort_pairs.argtypes = [_c_i64, _u64p, _u32p, _c_int]
        L.oracle_ranges.restype = _c_int
        L.oracle_ranges.argtypes = [_c_i64, _u64p, _c_int, _u32p]
        L.oracle_blend.restype = _c_int
        L.oracle_blend.argtypes = [
            _c_int, _c_int, _c_int, _u32p, _u32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
            _f32p, _f32p, _f32p, _u32p, _f32p, _c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_blend_backward.restype = _c_int
        L.oracle_blend_backward.argtypes = [
            _c_int, _c_int, _c_int, _u32p, _u32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
            _u32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
        L.oracle_preprocess_backward.restype = _c_int
        L.oracle_preprocess_backward.argtypes = [
            _c_int, _c_int, _c_int, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _c_int,
            _c_int, _c_float, _c_float, _i32p, _f32p, _f32p, _f32p, _f32p,
            _f32p, _f32p, _f32p, _f32p, _f32p]
        L.oracle_voxel_field.restype = None
        L.oracle_voxel_field.argtypes = [_f32p, _f32p, _f32p, _f32p, _c_int, _c_int, _c_int, _c_float]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def tile_grid(W: int, H: int, ts: int):
    return (W + ts - 1) // ts, (H + ts - 1) // ts


def preprocess(means, scales, rots, opac, sh, D, view, proj, campos, bg, W, H, ts, fx, fy):
    L = lib()
    means, scales, rots, opac, sh = map(_f32, (means, scales, rots, opac, sh))
    N = means.shape[0]
    M = sh.shape[1] if N else 1
    out = dict(m2d=np.zeros((N, 2), np.float32), depth=np.zeros(N, np.float32),
               conic=np.zeros((N, 3), np.float32), rgb=np.zeros((N, 3), np.float32),
               radii=np.zeros(N, np.int32), rect=np.zeros((N, 4), np.uint32),
               counts=np.zeros(N, np.uint32), cov3d=np.zeros((N, 6), np.float32))
    if N:
        L.oracle_preprocess(N, M, D, means, scales, rots, opac.reshape(-1), sh.reshape(-1), _f32(view),
                            _f32(proj), _f32(campos), _f32(bg), W, H, ts, fx, fy, out["m2d"],
                            out["depth"], out["conic"], out["rgb"], out["radii"], out["rect"],
                            out["counts"], out["cov3d"])
    return out


def forward(means, scales, rots, opac, sh, D, view, proj, campos, bg, W, H, ts, fx, fy,
            record: int = 0) -> dict:
    """Full forward.  ``record`` > 0 also returns each pixel's contributor ids
    (at most ``record`` per pixel) as ``rec_count`` [H*W] and ``rec_ids``."""
    L = lib()
    opac = _f32(opac).reshape(-1)
    pre = preprocess(means, scales, rots, opac, sh, D, view, proj, campos, bg, W, H, ts, fx, fy)
    N = pre["counts"].shape[0]
    scan = np.zeros(N, np.uint32)
    K = int(L.oracle_inclusive_scan(N, pre["counts"], scan)) if N else 0
    keys = np.zeros(max(K, 1), np.uint64)
    vals = np.zeros(max(K, 1), np.uint32)
    if K:
        L.oracle_emit(N, W, ts, pre["rect"], pre["counts"], scan, pre["depth"], keys, vals)
    keys, vals = keys[:K].copy(), vals[:K].copy()
    unsorted_keys, unsorted_vals = keys.copy(), vals.copy()
    tx, ty = tile_grid(W, H, ts)
    T = tx * ty
    tbits = max(1, int(T - 1).bit_length())
    if K:
        L.oracle_sort_pairs(K, keys, vals, 32 + tbits)
    ranges = np.zeros((T, 2), np.uint32)
    if K:
        assert L.oracle_ranges(K, keys, T, ranges) == 0
    HW = W * H
    color = np.zeros((3, H, W), np.float32)
    alpha = np.zeros((H, W), np.float32)
    depth_img = np.zeros((H, W), np.float32)
    n_contrib = np.zeros(HW, np.uint32)
    final_T = np.zeros(HW, np.float32)
    rec_count = np.zeros(HW, np.uint32) if record else None
    rec_ids = np.zeros(HW * record, np.uint32) if record else None
    ids_arg = vals if K else np.zeros(1, np.uint32)
    rc = L.oracle_blend(W, H, ts, ranges, ids_arg, pre["m2d"], pre["conic"], opac, pre["rgb"],
                        pre["depth"], _f32(bg), color, alpha, depth_img, n_contrib, final_T, record,
                        rec_count.ctypes.data if record else None,
                        rec_ids.ctypes.data if record else None)
    if rc != 0:
        raise RuntimeError(f"oracle_blend failed ({rc}); raise record")
    out = dict(pre)
    out.update(scan=scan, K=K, unsorted_keys=unsorted_keys, unsorted_vals=unsorted_vals, keys=keys,
               vals=vals, ranges=ranges, color=color, alpha=alpha, depth_img=depth_img,
               n_contrib=n_contrib, final_T=final_T, opac=opac, W=W, H=H, ts=ts, D=D)
    if record:
        out["rec_count"] = rec_count
        out["rec_ids"] = rec_ids.reshape(HW, record)
    return out


def backward(fwd: dict, means, scales, rots, sh, view, proj, campos, bg, fx, fy,
             dL_dcolor, dL_dalpha, dL_ddepth) -> dict:
    """Hand-written backward of ``forward``; returns the reference's gradient
    tuple fields (rasterize_points.cu:139-148,188) plus the per-Gaussian
    intermediates."""
    L = lib()
    W, H, ts, D = fwd["W"], fwd["H"], fwd["ts"], fwd["D"]
    means, scales, rots, sh = map(_f32, (means, scales, rots, sh))
    N = means.shape[0]
    M = sh.shape[1] if N else 1
    g_m2d = np.zeros((N, 2), np.float32)
    g_conic = np.zeros((N, 3), np.float32)
    g_opac = np.zeros(N, np.float32)
    g_rgb = np.zeros((N, 3), np.float32)
    g_depth = np.zeros(N, np.float32)
    ids_arg = fwd["vals"] if fwd["K"] else np.zeros(1, np.uint32)
    if N:
        L.oracle_blend_backward(W, H, ts, fwd["ranges"], ids_arg, fwd["m2d"], fwd["conic"],
                                fwd["opac"], fwd["rgb"], fwd["depth"], _f32(bg), fwd["n_contrib"],
                                fwd["final_T"], _f32(dL_dcolor), _f32(dL_dalpha), _f32(dL_ddepth),
                                g_m2d, g_conic, g_opac, g_rgb, g_depth)
    dm = np.zeros((N, 3), np.float32)
    dcov = np.zeros((N, 6), np.float32)
    dsh = np.zeros((N, M, 3), np.float32)
    dsc = np.zeros((N, 3), np.float32)
    drot = np.zeros((N, 4), np.float32)
    if N:
        L.oracle_preprocess_backward(N, M, D, means, scales, rots, sh.reshape(-1), _f32(view),
                                     _f32(proj), _f32(campos), W, H, fx, fy, fwd["radii"], g_m2d,
                                     g_conic, g_rgb, g_depth, dm, dcov, dsh, dsc, drot)
    m2d3 = np.zeros((N, 3), np.float32)
    m2d3[:, :2] = g_m2d
    return dict(dL_dmeans2D=m2d3, dL_dcolors=g_rgb, dL_dopacity=g_opac.reshape(N, 1),
                dL_dmeans3D=dm, dL_dcov3D=dcov, dL_dsh=dsh, dL_dscales=dsc, dL_drotations=drot,
                g_conic=g_conic, g_depth=g_depth)


def expf(x: float) -> float:
    return float(lib().oracle_expf(float(x)))


def logf(x: float) -> float:
    return float(lib().oracle_logf(float(x)))


def voxel_field(means, covs, opacities, resolution: int, num_blocks: int, relax_ratio: float) -> np.ndarray:
    """compute_voxel_field.cu:25-79 on the CPU: occ [res^3] indexed
    ix + iy res + iz res^2 (reshape (res, res, res) gives [iz, iy, ix])."""
    means, covs, opac = _f32(means).reshape(-1, 3), _f32(covs).reshape(-1, 6), _f32(opacities).reshape(-1)
    occ = np.zeros(resolution ** 3, np.float32)
    n = means.shape[0]
    lib().oracle_voxel_field(occ, means if n else np.zeros((1, 3), np.float32),
                             covs if n else np.zeros((1, 6), np.float32),
                             opac if n else np.zeros(1, np.float32), n, resolution, num_blocks, relax_ratio)
    return occ

