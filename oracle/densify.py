"""CPU restatement of one adaptive-density-control cycle.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Follows the reference phase by phase, as gaussians_handler.adc_cycle
(gaussians_handler.py:185-194) runs them, on numpy float32 arrays:

1. clone_cycle (gaussians_handler.py:148-165, Densification.py:33-52):
   mask = |accum/(count+1e-5)| >= thr  &  any(exp(svec) <= 0.04); the masked
   rows are appended (torch.cat, OptimizerManager.densify_on_optimizer,
   OptimizerManager.py:33-45: zero moments); statistics get zeros appended.
2. split_cycle (gaussians_handler.py:131-146, Densification.py:56-107), on the
   enlarged set: mask = |g| >= thr  &  any(exp(svec) > 4); children
   = rows[mask] repeated `factor` times block-major (tensor.repeat), means
   R(q/|q|) @ (std * z) + mean with std = exp(svec) and z the caller's
   standard normals (torch.normal(0, std) = std * randn), svec
   log(std / (factor * 0.8)), the rest copied; children appended with zero
   moments, then the masked rows removed (OptimizerManager.prune_optimizer,
   OptimizerManager.py:23-31).
3. statistics reset (gaussians_handler.py:190-191).
4. prune_cycle (gaussians_handler.py:167-183, Densification.py:167-187):
   sigmoid(opacity) < min_opacity  |  max(exp(svec)) > 0.1 * 4.

Arrays: params / exp_avg / exp_avg_sq are dicts of the six groups
(means [n,3], sh_dc [n,1,3], sh_ac [n,K,3], opacity [n,1], svec [n,3],
quaternion [n,4]).
"""
from __future__ import annotations

import numpy as np

GROUPS = ("means", "sh_dc", "sh_ac", "opacity", "svec", "quaternion")
f32 = np.float32


def _cat(d: dict, ext: dict) -> dict:
    return {k: np.concatenate([d[k], ext[k].astype(f32)], 0) for k in GROUPS}


def _take(d: dict, idx) -> dict:
    return {k: d[k][idx] for k in GROUPS}


def _zeros_like(d: dict) -> dict:
    return {k: np.zeros_like(v) for k, v in d.items()}


def _grad_mask(accum, count, thr):
    g = accum / (count + f32(1e-5))
    return np.abs(g) >= f32(thr)


def _rotmats(q):
    """Densification.py:75-90 (quaternion normalised by its explicit norm)."""
    nrm = np.sqrt(q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2] + q[:, 3] * q[:, 3])
    qn = q / nrm[:, None]
    w, x, y, z = qn[:, 0], qn[:, 1], qn[:, 2], qn[:, 3]
    two, one = f32(2), f32(1)
    R = np.empty((q.shape[0], 3, 3), f32)
    R[:, 0, 0] = one - two * (y * y + z * z)
    R[:, 1, 0] = two * (x * y + w * z)
    R[:, 2, 0] = two * (x * z - w * y)
    R[:, 0, 1] = two * (x * y - w * z)
    R[:, 1, 1] = one - two * (x * x + z * z)
    R[:, 2, 1] = two * (y * z + w * x)
    R[:, 0, 2] = two * (x * z + w * y)
    R[:, 1, 2] = two * (y * z - w * x)
    R[:, 2, 2] = one - two * (x * x + y * y)
    return R


def adc_cycle(params: dict, exp_avg: dict, exp_avg_sq: dict, accum, count, normals,
              grad_thresh: float, min_opacity: float, clone_scale_max: float = 0.04,
              split_scale_min: float = 4.0, prune_scale_max: float = 0.4, split_factor: int = 2,
              split_shrink: float = 0.8):
    """Returns (params, exp_avg, exp_avg_sq, (n_clone, n_split, n_prune))."""
    P = {k: np.asarray(params[k], f32).copy() for k in GROUPS}
    M = {k: np.asarray(exp_avg[k], f32).copy() for k in GROUPS}
    V = {k: np.asarray(exp_avg_sq[k], f32).copy() for k in GROUPS}
    acc = np.asarray(accum, f32).reshape(-1)
    cnt = np.asarray(count, f32).reshape(-1)
    z = np.asarray(normals, f32).reshape(-1, 3)

    # 1. clone
    scales = np.exp(P["svec"])
    clone = _grad_mask(acc, cnt, grad_thresh) & (scales <= f32(clone_scale_max)).any(-1)
    n_clone = int(clone.sum())
    ext = _take(P, clone)
    P, M, V = _cat(P, ext), _cat(M, _zeros_like(ext)), _cat(V, _zeros_like(ext))
    acc = np.concatenate([acc, np.zeros(n_clone, f32)])
    cnt = np.concatenate([cnt, np.zeros(n_clone, f32)])

    # 2. split
    scales = np.exp(P["svec"])
    split = _grad_mask(acc, cnt, grad_thresh) & (scales > f32(split_scale_min)).any(-1)
    S = int(split.sum())
    if z.shape[0] < split_factor * S:
        raise ValueError("too few normal samples")
    sel = _take(P, split)
    rep = lambda a: np.concatenate([a] * split_factor, 0)
    std = rep(scales[split])
    sample = std * z[:split_factor * S]
    R = _rotmats(rep(sel["quaternion"]))
    # (R @ s) row by row, in the order ((r0 s0 + r1 s1) + r2 s2)
    rs = (R[:, :, 0] * sample[:, None, 0] + R[:, :, 1] * sample[:, None, 1]) + R[:, :, 2] * sample[:, None, 2]
    div = f32(float(split_factor) * float(f32(split_shrink)))
    children = {"means": rs + rep(sel["means"]), "sh_dc": rep(sel["sh_dc"]), "sh_ac": rep(sel["sh_ac"]),
                "opacity": rep(sel["opacity"]), "svec": np.log(std / div), "quaternion": rep(sel["quaternion"])}
    P, M, V = _cat(P, children), _cat(M, _zeros_like(children)), _cat(V, _zeros_like(children))
    keep = np.concatenate([~split, np.ones(split_factor * S, bool)])
    P, M, V = _take(P, keep), _take(M, keep), _take(V, keep)

    # 4. prune (statistics restart at zero: the caller sizes them from the result)
    opac = f32(1) / (f32(1) + np.exp(-P["opacity"][:, 0]))
    scales = np.exp(P["svec"])
    prune = (opac < f32(min_opacity)) | (scales.max(-1) > f32(prune_scale_max))
    n_prune = int(prune.sum())
    P, M, V = _take(P, ~prune), _take(M, ~prune), _take(V, ~prune)
    return P, M, V, (n_clone, S, n_prune)
