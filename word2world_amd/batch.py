"""Batched form: many scenes and views per launch.

A batch packs the Gaussians of several scenes back to back and lists the
frames (views) to render; every stage of the rasterizer then runs as one
launch over all (frame, Gaussian) items, all tiles of all frames and all
pixels.  Sizes are capacities: a workspace sized once serves every batch that
fits, the kernels read the actual counts from the device-side table, and the
launch sequence can be captured in a HIP graph (`torch.cuda.graph`).

The single-view entry points of `raster` are the one-scene, one-frame case of
the same kernels.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib


@dataclass
class FrameSpec:
    viewmatrix: np.ndarray   # 16 floats, column-major W2C (Renderer.py:135)
    projmatrix: np.ndarray   # 16 floats
    campos: np.ndarray       # 3
    width: int
    height: int
    focal_x: float
    focal_y: float
    background: np.ndarray = field(default_factory=lambda: np.ones(3, np.float32))
    backward: bool = True
    # emit into the tiles the Gaussian's cull extent meets only (fewer
    # instances; identical images and gradients, but K, keys, ranges and
    # n_contrib are no longer the reference's)
    tight: bool = False


@dataclass
class SceneSpec:
    n: int                   # Gaussians
    sh_degree: int
    n_coeffs: int            # SH coefficients per Gaussian (row stride)
    frames: list             # FrameSpec per view


def frame_from_camera(cam, background=None, backward=True, tight=False) -> FrameSpec:
    """From a word2world_amd.camera.Camera."""
    bg = np.ones(3, np.float32) if background is None else np.asarray(background, np.float32)
    return FrameSpec(cam.view_flat(), cam.proj_flat(), cam.campos_np(), cam.width, cam.height,
                     cam.focal_x, cam.focal_y, bg, backward, tight)


def caps_for(scenes: list, tile: int, capacity: int, scenes_cap: int | None = None,
             frames_cap: int | None = None, depth_split: float = 0.0, chunk: int = 0) -> _lib.W2WCaps:
    """Capacities of a batch.  depth_split in (0, 1): depth-sliced binning
    with that fraction of the items (nearest first) in the near slice; chunk:
    instances per forward-blend chunk (64, 128 or 256; 0 = 64)."""
    g = sum(s.n for s in scenes)
    items = sum(s.n * len(s.frames) for s in scenes)
    frames = [f for s in scenes for f in s.frames]
    tiles = sum(((f.width + tile - 1) // tile) * ((f.height + tile - 1) // tile) for f in frames)
    pix = sum(f.width * f.height for f in frames)
    c = _lib.W2WCaps()
    c.scenes = scenes_cap or max(len(scenes), 1)
    c.frames = frames_cap or max(len(frames), 1)
    c.tile = tile
    c.chunk = int(chunk)
    c.depth_split = int(round(1000 * depth_split)) if 0.0 < depth_split < 1.0 else 0
    if 0.0 < depth_split < 1.0:
        c.depth_split = min(max(c.depth_split, 1), 999)
    c.items, c.gaussians, c.tiles, c.pixels, c.capacity = items, g, max(tiles, 1), max(pix, 1), capacity
    return c


def encode(scenes: list, caps: _lib.W2WCaps, bases=None, sh_offs=None,
           target_offs=None) -> tuple[np.ndarray, np.ndarray]:
    """Host bytes of the batch table and the totals (G, items, frames, tiles,
    pixels, SH floats).  `bases` / `sh_offs`: per-scene row and SH float offset
    in a resident store (None = packed in batch order)."""
    L = _lib.load()
    nbytes = L.w2w_batch_table_bytes(caps.scenes, caps.frames)
    buf = np.zeros(nbytes, np.uint8)
    frames = [f for s in scenes for f in s.frames]
    nf = max(len(frames), 1)
    view = np.zeros((nf, 16), np.float32)
    proj = np.zeros((nf, 16), np.float32)
    cp = np.zeros((nf, 3), np.float32)
    bg = np.zeros((nf, 3), np.float32)
    foc = np.zeros((nf, 2), np.float32)
    wh = np.zeros((nf, 2), np.int32)
    flags = np.zeros(nf, np.uint32)
    for k, f in enumerate(frames):
        view[k], proj[k], cp[k], bg[k] = f.viewmatrix, f.projmatrix, f.campos, f.background
        foc[k] = (f.focal_x, f.focal_y)
        wh[k] = (f.width, f.height)
        flags[k] = (1 if f.backward else 0) | (2 if f.tight else 0)
    sn = np.array([s.n for s in scenes] or [0], np.int64)
    sv = np.array([len(s.frames) for s in scenes] or [1], np.int32)
    sm = np.array([s.n_coeffs for s in scenes] or [1], np.int32)
    sd = np.array([s.sh_degree for s in scenes] or [0], np.int32)
    totals = np.zeros(6, np.int64)
    b = None if bases is None else np.ascontiguousarray(bases, np.int64)
    so = None if sh_offs is None else np.ascontiguousarray(sh_offs, np.int64)
    to = None if target_offs is None else np.ascontiguousarray(target_offs, np.int64)
    st = L.w2w_batch_encode(buf.ctypes.data, nbytes, ctypes.byref(caps), len(scenes), sn.ctypes.data, sv.ctypes.data, sm.ctypes.data,
                            sd.ctypes.data, view.ctypes.data, proj.ctypes.data, cp.ctypes.data,
                            bg.ctypes.data, foc.ctypes.data, wh.ctypes.data, flags.ctypes.data,
                            None if b is None else b.ctypes.data,
                            None if so is None else so.ctypes.data,
                            None if to is None else to.ctypes.data, totals.ctypes.data)
    _lib.check(st, "w2w_batch_encode")
    return buf, totals


class BatchWorkspace:
    """Device state for batches up to `caps`."""

    def __init__(self, caps: _lib.W2WCaps, device):
        self.caps = caps
        self.device = torch.device(device)
        self.layout = _lib.batch_layout(caps)
        L = self.layout
        # zeroed once: depth-sliced binning keeps a tile hint here between
        # forwards (any content is correct, an empty one is the neutral start)
        self.geom = torch.zeros(max(L.geom_bytes, 1), dtype=torch.uint8, device=device)
        self.binning = torch.empty(max(L.binning_bytes, 1), dtype=torch.uint8, device=device)
        self.image = torch.empty(max(L.image_bytes, 1), dtype=torch.uint8, device=device)
        nbytes = _lib.load().w2w_batch_table_bytes(caps.scenes, caps.frames)
        self.table = torch.zeros(nbytes, dtype=torch.uint8, device=device)
        self.table_host = torch.zeros(nbytes, dtype=torch.uint8).pin_memory()
        self._copied = None  # event after the last host->device table copy
        self.struct = _lib.W2WWorkspace(self.geom.data_ptr(), L.geom_bytes, self.binning.data_ptr(),
                                        L.binning_bytes, self.image.data_ptr(), L.image_bytes,
                                        caps.capacity)
        P, pix = int(caps.items), int(caps.pixels)
        f32 = dict(dtype=torch.float32, device=device)
        self.color = torch.empty(3 * pix, **f32)
        self.depth = torch.empty(pix, **f32)
        self.alpha = torch.empty(pix, **f32)
        self.radii = torch.empty(max(P, 1), dtype=torch.int32, device=device)
        self.totals = np.zeros(6, np.int64)

    def load_table(self, scenes: list, bases=None, sh_offs=None, target_offs=None) -> np.ndarray:
        """Encode on the host and copy into the device table (stream-ordered)."""
        buf, totals = encode(scenes, self.caps, bases, sh_offs, target_offs)
        if self._copied is not None:
            self._copied.synchronize()  # the pinned staging buffer is free again
        self.table_host.numpy()[:] = buf
        self.table.copy_(self.table_host, non_blocking=True)
        self._copied = torch.cuda.Event()
        self._copied.record()
        self.totals = totals
        return totals

    def status(self) -> torch.Tensor:
        """[K (all instances of the batch), overflow flag]."""
        return self.geom[:8].view(torch.int32)

    def sliced_counts(self) -> torch.Tensor:
        """Depth-sliced binning, last forward: [near depth positions, their
        instances, far instances on open unhinted tiles, unfinished blocks of
        unhinted tiles, far instances on hinted tiles, near-list instances]."""
        return self.geom[8:32].view(torch.int32)

    def final_T(self) -> torch.Tensor:
        """Final transmittance of every pixel of the encoded batch."""
        n, o = int(self.totals[4]), self.layout.final_T
        return self.image[o:o + 4 * n].view(torch.float32)

    def frame_views(self, scenes: list):
        """(color [3,H,W], depth [H,W], alpha [H,W]) views per frame."""
        out, pix = [], 0
        for s in scenes:
            for f in s.frames:
                hw = f.width * f.height
                out.append((self.color[3 * pix:3 * (pix + hw)].view(3, f.height, f.width),
                            self.depth[pix:pix + hw].view(f.height, f.width),
                            self.alpha[pix:pix + hw].view(f.height, f.width)))
                pix += hw
        return out


def _stream():
    return torch.cuda.current_stream().cuda_stream


def forward(ws: BatchWorkspace, means, scales, rots, opac, sh) -> None:
    """Renders the table's frames into ws.color / ws.depth / ws.alpha (packed
    per frame) and ws.radii (per item).  No synchronisation."""
    st = _lib.load().w2w_forward_batch(ws.table.data_ptr(), ctypes.byref(ws.caps),
                                       ctypes.byref(ws.struct), means.data_ptr(), scales.data_ptr(),
                                       rots.data_ptr(), opac.data_ptr(), sh.data_ptr(),
                                       ws.color.data_ptr(), ws.depth.data_ptr(), ws.alpha.data_ptr(),
                                       ws.radii.data_ptr(), _stream())
    _lib.check(st, "w2w_forward_batch")


class BatchGrads:
    def __init__(self, caps: _lib.W2WCaps, sh_floats: int, device):
        f32 = dict(dtype=torch.float32, device=device)
        I, G = max(int(caps.items), 1), max(int(caps.gaussians), 1)
        self.means2D = torch.empty(I, 3, **f32)
        self.colors = torch.empty(I, 3, **f32)
        self.opacity = torch.empty(G, **f32)
        self.means3D = torch.empty(G, 3, **f32)
        self.cov3D = torch.empty(G, 6, **f32)
        self.sh = torch.empty(max(sh_floats, 1), **f32)
        self.scales = torch.empty(G, 3, **f32)
        self.rotations = torch.empty(G, 4, **f32)


def backward(ws: BatchWorkspace, grads: BatchGrads, means, scales, rots, sh, dL_dcolor,
             dL_ddepth=None, dL_dalpha=None, sh_grad_out=None) -> None:
    p = lambda t: None if t is None else t.data_ptr()
    shg = grads.sh if sh_grad_out is None else sh_grad_out
    st = _lib.load().w2w_backward_batch(
        ws.table.data_ptr(), ctypes.byref(ws.caps), ctypes.byref(ws.struct), means.data_ptr(),
        scales.data_ptr(), rots.data_ptr(), sh.data_ptr(), dL_dcolor.data_ptr(), p(dL_ddepth),
        p(dL_dalpha), grads.means2D.data_ptr(), grads.colors.data_ptr(), grads.opacity.data_ptr(),
        grads.means3D.data_ptr(), grads.cov3D.data_ptr(), shg.data_ptr(), grads.scales.data_ptr(),
        grads.rotations.data_ptr(), _stream())
    _lib.check(st, "w2w_backward_batch")
