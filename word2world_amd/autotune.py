"""Tile-size autotuner.

The tile edge (8, 16 or 32 pixels) trades sort and emission work (more,
smaller tiles: more instances) against blend work (fewer, larger tiles: longer
per-tile lists and coarser culling).  `sweep` times one fwd+bwd of a view at
each edge on the device; `Autotuner` keeps the winner per (Gaussian-count
bucket, width, height) so a trainer or server tunes once per workload shape.

The reference's second knob, the LDS batch of its blend (64/128/256), has no
counterpart here: both blend kernels gather in wave-width chunks of 64 with a
one-chunk prefetch and the forward uses no LDS.
"""
from __future__ import annotations

import math

import torch

from . import raster

TILES = (8, 16, 32)


def _fwd_bwd(m, s, r, o, sh, deg, cam, ts, ws, g, z):
    color, depth, alpha, radii, ws = raster.forward(m, s, r, o, sh, deg, cam, ts, ws)
    raster.backward(m, s, r, o, sh, deg, cam, ts, ws, g, z, z)
    return ws


def sweep(means, scales, rotations, opacities, sh, sh_degree: int, cam: raster.Camera,
          tiles=TILES, reps: int = 3) -> dict:
    """{tile: ms per fwd+bwd} for one view (inputs activated, on the device)."""
    out = {}
    H, W = int(cam.height), int(cam.width)
    g = torch.full((3, H, W), 1e-3, dtype=torch.float32, device=means.device)
    z = torch.zeros((H, W), dtype=torch.float32, device=means.device)
    for ts in tiles:
        # size the workspace (and let the capacity policy learn K)
        _, _, _, _, _, ws = raster.forward_checked(means, scales, rotations, opacities, sh, sh_degree,
                                                   cam, ts)
        ws = _fwd_bwd(means, scales, rotations, opacities, sh, sh_degree, cam, ts, ws, g, z)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(reps):
            ws = _fwd_bwd(means, scales, rotations, opacities, sh, sh_degree, cam, ts, ws, g, z)
        b.record()
        torch.cuda.synchronize()
        if int(ws.status()[1]):
            raise RuntimeError(f"instance capacity overflowed while timing tile {ts}")
        out[ts] = a.elapsed_time(b) / reps
        del ws
    return out


class Autotuner:
    """Best tile edge per workload shape, measured on first use."""

    def __init__(self, tiles=TILES, reps: int = 3):
        self.tiles, self.reps = tuple(tiles), reps
        self.table: dict = {}
        self.timings: dict = {}

    @staticmethod
    def key(P: int, W: int, H: int):
        return (int(math.log2(max(P, 1))), int(W), int(H))

    def best_tile(self, means, scales, rotations, opacities, sh, sh_degree, cam) -> int:
        k = self.key(int(means.shape[0]), int(cam.width), int(cam.height))
        if k not in self.table:
            t = sweep(means, scales, rotations, opacities, sh, sh_degree, cam, self.tiles, self.reps)
            self.timings[k] = t
            self.table[k] = min(t, key=t.get)
        return self.table[k]
