"""Host side of on-device densification (w2w_densify_plan / w2w_densify_apply).

A store here is a flat float32 tensor, group-major, with the rows of one or
more segments (scenes) back to back inside each group, plus its two Adam
moment tensors of the same layout.  `densify_segments` runs one adc_cycle
(gaussians_handler.py:185-194: clone, split, statistics reset, prune; the
Adam remap of OptimizerManager.py:23-45) on every segment and returns the new
store: each segment's surviving rows stay in place of the segment, followed
by its clones and then its split children, as the reference orders them.
One host synchronisation reads the per-segment counts (the new sizes).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

# clone / split / prune thresholds of the reference trainer
# (GaussianTrainer.py:125, gaussians_handler.py:131-183)
DEFAULTS = dict(clone_scale_max=0.04, split_scale_min=4.0, prune_scale_max=0.1 * 4, split_factor=2,
                split_shrink=0.8)


def config(grad_thresh: float, min_opacity: float, **kw) -> _lib.W2WDensifyConfig:
    o = dict(DEFAULTS, **kw)
    return _lib.W2WDensifyConfig(grad_thresh, min_opacity, o["clone_scale_max"], o["split_scale_min"],
                                 o["prune_scale_max"], o["split_shrink"], int(o["split_factor"]), 0)


def _groups(widths, roles, bufs, rows_total: int, first_row: int) -> _lib.W2WStoreGroups:
    g = _lib.W2WStoreGroups()
    g.n_groups = len(widths)
    for k, w in enumerate(widths):
        g.width[k] = w
    g.means, g.opacity, g.svec, g.quat = roles
    start = 0
    for k, w in enumerate(widths):
        off = 4 * (start + first_row * w)
        g.param[k] = bufs[0].data_ptr() + off
        g.exp_avg[k] = bufs[1].data_ptr() + off
        g.exp_avg_sq[k] = bufs[2].data_ptr() + off
        start += w * rows_total
    return g


def densify_segments(widths: list, roles: tuple, params: torch.Tensor, exp_avg: torch.Tensor,
                     exp_avg_sq: torch.Tensor, ns: list, accum: torch.Tensor, count: torch.Tensor,
                     cfg: _lib.W2WDensifyConfig, generator: torch.Generator | None = None,
                     normals: torch.Tensor | None = None):
    """widths: floats per row of each group; roles: group indices of (means,
    opacity logit, svec, quaternion); ns: rows per segment; accum/count: one
    float per row.  normals: [sum over segments of factor * splits, 3]
    standard normal samples, segment by segment (drawn with `generator` when
    None).  Returns (params, exp_avg, exp_avg_sq, new ns, counts) with
    counts[s] = (clones, splits, pruned) of segment s."""
    dev = params.device
    if dev.type != "cuda":
        raise RuntimeError("densification runs on the device (HIP); the store is on " + str(dev))
    L = _lib.load()
    s = torch.cuda.current_stream().cuda_stream
    G = int(sum(ns))
    per_row = int(sum(widths))
    for t in (params, exp_avg, exp_avg_sq):
        if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() < G * per_row:
            raise ValueError("store tensors must be contiguous float32 of at least rows * row width")
    accum = accum.float().contiguous()
    count = count.float().contiguous()
    if accum.numel() < G or count.numel() < G:
        raise ValueError("statistics need one value per row")
    offs = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    tb = [(int(L.w2w_densify_temp_bytes(int(n))) + 255) // 256 * 256 for n in ns]
    toff = np.concatenate([[0], np.cumsum(tb)]).astype(np.int64)
    temp = torch.empty(max(int(toff[-1]), 1), dtype=torch.uint8, device=dev)
    counts = torch.zeros(len(ns), 8, dtype=torch.int64, device=dev)
    bufs = (params, exp_avg, exp_avg_sq)
    for k, n in enumerate(ns):
        gin = _groups(widths, roles, bufs, G, int(offs[k]))
        _lib.check(L.w2w_densify_plan(ctypes.byref(gin), int(n), accum.data_ptr() + 4 * int(offs[k]),
                                      count.data_ptr() + 4 * int(offs[k]), ctypes.byref(cfg),
                                      temp.data_ptr() + int(toff[k]), tb[k], counts[k].data_ptr(), s),
                   "w2w_densify_plan")
    c = counts.cpu().numpy()
    ns2 = [int(x) for x in c[:, 3]]
    need = [int(cfg.split_factor) * int(x) for x in c[:, 1]]
    noff = np.concatenate([[0], np.cumsum(need)]).astype(np.int64)
    if normals is None:
        gdev = generator.device if generator is not None else dev
        normals = torch.randn(int(noff[-1]), 3, generator=generator, device=gdev)
    normals = normals.to(dev, torch.float32).contiguous()
    if normals.dim() != 2 or normals.shape[1] != 3 or normals.shape[0] < int(noff[-1]):
        raise ValueError(f"densify needs [{int(noff[-1])}, 3] normal samples, got {tuple(normals.shape)}")
    G2 = int(sum(ns2))
    out = tuple(torch.empty(max(G2 * per_row, 1), dtype=torch.float32, device=dev) for _ in range(3))
    offs2 = np.concatenate([[0], np.cumsum(ns2)]).astype(np.int64)
    for k, n in enumerate(ns):
        gin = _groups(widths, roles, bufs, G, int(offs[k]))
        gout = _groups(widths, roles, out, G2, int(offs2[k]))
        _lib.check(L.w2w_densify_apply(ctypes.byref(gin), int(n), ctypes.byref(cfg),
                                       temp.data_ptr() + int(toff[k]),
                                       normals.data_ptr() + 12 * int(noff[k]), need[k], ctypes.byref(gout),
                                       ns2[k], s), "w2w_densify_apply")
    k = G2 * per_row
    return out[0][:k], out[1][:k], out[2][:k], ns2, [tuple(int(v) for v in r[:3]) for r in c]
