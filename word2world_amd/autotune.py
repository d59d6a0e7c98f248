"""Autotuner for the rasterizer's two speed knobs (same outputs for every
choice):

* the tile edge (8, 16 or 32 pixels), which trades sort and emission work
  (more, smaller tiles: more instances) against blend work (fewer, larger
  tiles: longer per-tile lists and coarser culling);
* the forward blend's chunk (64, 128 or 256 instances gathered per wave and
  step of its walk, one to four per lane), the counterpart of the reference
  blend's shared-memory batch of 64/128/256 Gaussians
  (pixelGaussiansFull.cu:41-76): it sets how many splat-record gathers each
  lane keeps in flight against how much a wave loads past the point where
  its pixels saturate.

`sweep` times one fwd+bwd of a view for every (tile, chunk) on the device;
`Autotuner` keeps the winner per (Gaussian-count bucket, width, height) so a
trainer or server tunes once per workload shape; `sweep_trainer` times a few
steps of a caller-built trainer per candidate (the batched form, where one
choice serves every scene of the batch).
"""
from __future__ import annotations

import math
import time

import torch

from . import raster

TILES = (8, 16, 32)
CHUNKS = (64, 128, 256)


def candidates(tiles=TILES, chunks=CHUNKS) -> list:
    return [(t, c) for t in tiles for c in chunks]


def label(key) -> str:
    return f"tile{key[0]}/chunk{key[1]}"


def _fwd_bwd(m, s, r, o, sh, deg, cam, ts, chunk, ws, g, z):
    color, depth, alpha, radii, ws = raster.forward(m, s, r, o, sh, deg, cam, ts, ws, chunk=chunk)
    raster.backward(m, s, r, o, sh, deg, cam, ts, ws, g, z, z)
    return ws


def sweep(means, scales, rotations, opacities, sh, sh_degree: int, cam: raster.Camera,
          tiles=TILES, chunks=CHUNKS, reps: int = 3) -> dict:
    """{(tile, chunk): ms per fwd+bwd} for one view (inputs activated, on the
    device)."""
    out = {}
    H, W = int(cam.height), int(cam.width)
    g = torch.full((3, H, W), 1e-3, dtype=torch.float32, device=means.device)
    z = torch.zeros((H, W), dtype=torch.float32, device=means.device)
    for ts in tiles:
        # size the workspace (and let the capacity policy learn K)
        _, _, _, _, _, ws = raster.forward_checked(means, scales, rotations, opacities, sh, sh_degree,
                                                   cam, ts)
        for ch in chunks:
            ws = _fwd_bwd(means, scales, rotations, opacities, sh, sh_degree, cam, ts, ch, ws, g, z)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            a.record()
            for _ in range(reps):
                ws = _fwd_bwd(means, scales, rotations, opacities, sh, sh_degree, cam, ts, ch, ws, g, z)
            b.record()
            torch.cuda.synchronize()
            if int(ws.status()[1]):
                raise RuntimeError(f"instance capacity overflowed while timing tile {ts}, chunk {ch}")
            out[(ts, ch)] = a.elapsed_time(b) / reps
        del ws
    return out


def sweep_trainer(make, cands=None, warmup: int = 3, steps: int = 5) -> dict:
    """{(tile, chunk): ms per step} of trainers built by make(tile, chunk)
    (each built, stepped `warmup` times untimed, timed over `steps` steps and
    dropped before the next).  A candidate whose batch the trainer refuses
    (more tile instances than one workspace holds) is left out."""
    out = {}
    for ts, ch in (cands or candidates()):
        try:
            t = make(ts, ch)
        except RuntimeError as e:
            if "tile instances" not in str(e):
                raise
            torch.cuda.empty_cache()
            continue
        for _ in range(warmup):
            t.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            t.step()
        torch.cuda.synchronize()
        out[(ts, ch)] = (time.perf_counter() - t0) / steps * 1e3
        if t.overflow() or t.skipped_overflow():
            raise RuntimeError(f"instance capacity overflowed while timing tile {ts}, chunk {ch}")
        del t
        torch.cuda.empty_cache()
    return out


class Autotuner:
    """Best (tile edge, chunk) per workload shape, measured on first use."""

    def __init__(self, tiles=TILES, chunks=CHUNKS, reps: int = 3):
        self.tiles, self.chunks, self.reps = tuple(tiles), tuple(chunks), reps
        self.table: dict = {}
        self.timings: dict = {}

    @staticmethod
    def key(P: int, W: int, H: int):
        return (int(math.log2(max(P, 1))), int(W), int(H))

    def best(self, means, scales, rotations, opacities, sh, sh_degree, cam) -> tuple:
        k = self.key(int(means.shape[0]), int(cam.width), int(cam.height))
        if k not in self.table:
            t = sweep(means, scales, rotations, opacities, sh, sh_degree, cam, self.tiles, self.chunks,
                      self.reps)
            self.timings[k] = t
            self.table[k] = min(t, key=t.get)
        return self.table[k]

    def best_tile(self, means, scales, rotations, opacities, sh, sh_degree, cam) -> int:
        return self.best(means, scales, rotations, opacities, sh, sh_degree, cam)[0]
