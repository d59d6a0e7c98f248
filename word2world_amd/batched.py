"""Batched trainer: many scenes optimised together, one launch per stage.

All scenes' parameters live in one flat HBM store, group-major:
``means [G,3] | sh [G,M,3] | opacity logits [G] | svec [G,3] | quaternion
[G,4]`` with the scenes back to back inside each group, so the SH group *is*
the packed SH array the rasterizer reads and the means group its packed
means.  One step is a fixed sequence of native launches -- activation,
batched forward, L2 loss, batched backward (writing the means gradient
straight into the flat gradient; the SH coefficients take their Adam step
inside the SH backward, so their gradient is never stored), activation
backward, Adam over the other groups -- with
no autograd and no host synchronisation, so it can be captured once in a HIP
graph and replayed; the per-step learning rates reach the captured Adam
kernel through a small device buffer refreshed before each replay.

Learning rates follow optimizer_config.json (means 0.01 with the
misc_utils.py:11-27 schedule, SH DC 0.01, SH rest 0.0005, opacity 0.05, svec
0.005, quaternion 0.005); Adam eps 1e-15 (gaussians_handler.py:40).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib, batch
from .train import LRS, mean_lr

GROUPS = ("means", "sh", "opacity", "svec", "quaternion")
MAX_CAPACITY = (1 << 30) - 2  # tile instances per batch (30-bit look-back values in the sort)


class BatchedStore:
    def __init__(self, scenes: list, device):
        """scenes: list of word2world_amd.scene.GaussianScene (equal SH width)."""
        self.device = torch.device(device)
        ms = {s.sh_ac.shape[1] + 1 for s in scenes}
        if len(ms) != 1:
            raise ValueError("all scenes of a batched store need the same SH coefficient count")
        self.M = ms.pop()
        self.ns = [s.n for s in scenes]
        self.degrees = [s.sh_degree for s in scenes]
        self.G = G = sum(self.ns)
        cat = lambda f: torch.cat([f(s) for s in scenes]).reshape(-1).float()
        parts = [cat(lambda s: s.means), cat(lambda s: torch.cat((s.sh_dc, s.sh_ac), 1)),
                 cat(lambda s: s.opacity), cat(lambda s: s.svec), cat(lambda s: s.quaternion)]
        self.sizes = [3 * G, 3 * self.M * G, G, 3 * G, 4 * G]
        self.ends = np.cumsum(self.sizes).astype(np.int64)
        self.flat = torch.cat(parts).to(self.device).contiguous()
        self.grad = torch.zeros_like(self.flat)
        self.exp_avg = torch.zeros_like(self.flat)
        self.exp_avg_sq = torch.zeros_like(self.flat)
        self.step_count = 0
        f32 = dict(dtype=torch.float32, device=self.device)
        self.scales = torch.empty(G, 3, **f32)
        self.opac = torch.empty(G, **f32)
        self.rots = torch.empty(G, 4, **f32)
        self.g_scales = torch.empty(G, 3, **f32)
        self.g_opac = torch.empty(G, **f32)
        self.g_rots = torch.empty(G, 4, **f32)
        self.g_cov = torch.empty(G, 6, **f32)
        # Adam hyper-parameters (w2w_adam_step_hyper layout) in [0, 32), and in
        # [32, 64) the same for a launch over the groups after the SH block
        # (opacity, svec, quaternion as groups 0-2)
        self.hyper = torch.zeros(64, **f32)
        # ring of pinned staging buffers: a buffer is rewritten only after the
        # copy that last read it has completed
        self._ring = [torch.zeros(64, dtype=torch.float32).pin_memory() for _ in range(4)]
        self._ring_ev = [None] * 4
        self._ring_i = 0

    def rebuild(self, flat: torch.Tensor, exp_avg: torch.Tensor, exp_avg_sq: torch.Tensor, ns: list):
        """Adopt a re-laid-out store (after densification): new rows per scene."""
        self.ns = [int(n) for n in ns]
        self.G = G = sum(self.ns)
        self.sizes = [3 * G, 3 * self.M * G, G, 3 * G, 4 * G]
        self.ends = np.cumsum(self.sizes).astype(np.int64)
        self.flat, self.exp_avg, self.exp_avg_sq = flat, exp_avg, exp_avg_sq
        self.grad = torch.zeros_like(self.flat)
        f32 = dict(dtype=torch.float32, device=self.device)
        self.scales = torch.empty(G, 3, **f32)
        self.opac = torch.empty(G, **f32)
        self.rots = torch.empty(G, 4, **f32)
        self.g_scales = torch.empty(G, 3, **f32)
        self.g_opac = torch.empty(G, **f32)
        self.g_rots = torch.empty(G, 4, **f32)
        self.g_cov = torch.empty(G, 6, **f32)

    def upload_hyper(self, h: np.ndarray) -> None:
        k = self._ring_i
        self._ring_i = (k + 1) % len(self._ring)
        if self._ring_ev[k] is not None:
            self._ring_ev[k].synchronize()
        self._ring[k].numpy()[:] = h
        self.hyper.copy_(self._ring[k], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self._ring_ev[k] = ev

    def group(self, t: torch.Tensor, name: str) -> torch.Tensor:
        k = GROUPS.index(name)
        a = 0 if k == 0 else int(self.ends[k - 1])
        return t[a:int(self.ends[k])]

    def view(self, name: str, t: torch.Tensor | None = None) -> torch.Tensor:
        t = self.flat if t is None else t
        g = self.group(t, name)
        shapes = {"means": (self.G, 3), "sh": (self.G, self.M, 3), "opacity": (self.G,),
                  "svec": (self.G, 3), "quaternion": (self.G, 4)}
        return g.view(shapes[name])

    def hyper_values(self, step: int, lrs: dict, b1=0.9, b2=0.999, eps=1e-15) -> np.ndarray:
        bc1 = 1.0 - b1 ** step
        bc2 = 1.0 - b2 ** step
        h = np.zeros(64, np.float32)
        h[0] = -lrs["means"] / bc1
        h[1] = -lrs["sh_dc"] / bc1
        h[8 + 1] = -lrs["sh_ac"] / bc1
        h[2] = -lrs["opacity"] / bc1
        h[3] = -lrs["svec"] / bc1
        h[4] = -lrs["quaternion"] / bc1
        h[16] = 1.0 / np.sqrt(bc2)
        h[17], h[18], h[19], h[20] = 1.0 - b1, b2, 1.0 - b2, eps
        h[32:35] = h[2:5]
        h[48:53] = h[16:21]
        return h


class BatchedTrainer:
    """fwd + bwd + Adam for every scene of the batch in one launch sequence."""

    def __init__(self, scenes: list, cameras: list, targets: list | None, device, tile: int = 16,
                 capacity: int | None = None, tight: bool = True, densify_every: int = 0,
                 densify_grad: float = 2e-4, densify_min_opacity: float = 0.005,
                 generator: torch.Generator | None = None, depth_split: float = 0.0,
                 densify_options: dict | None = None, chunk: int = 0, fused_sh_adam: bool = True):
        """cameras[k]: word2world_amd.camera.Camera views of scene k;
        targets[k][v]: [3,H,W] target image (or None for render-only);
        tight: emit instances only where a Gaussian's cull extent reaches
        (same images and gradients, fewer instances);
        densify_every: run one densification cycle (clone / split / prune
        with the Adam remap, gaussians_handler.py:185-194) every that many
        steps, with the statistics of gaussians_handler.py:116-118 collected
        on the device each step (0 = never); `generator` draws the split
        samples; depth_split in (0, 1): depth-sliced binning (the nearest
        fraction of the items binned and blended first, the rest only on
        tiles with unsaturated pixels; same images and gradients);
        densify_options: the cycle's other thresholds (clone_scale_max,
        split_scale_min, prune_scale_max, split_factor, split_shrink of
        word2world_amd.densify.config; the reference trainer's values when
        absent); chunk: instances per forward-blend chunk (64, 128 or 256;
        0 = 64), a speed knob with the same outputs; fused_sh_adam: the SH
        backward takes the SH coefficients' Adam step itself (their gradient
        is never written; the same parameters bit for bit)."""
        self.fused_sh_adam = bool(fused_sh_adam)
        self.store = BatchedStore(scenes, device)
        self.chunk = int(chunk)
        self.depth_split = float(depth_split)
        self.device = torch.device(device)
        self.tile = tile
        self.specs = [batch.SceneSpec(s.n, s.sh_degree, self.store.M,
                                      [batch.frame_from_camera(c, tight=tight) for c in cams])
                      for s, cams in zip(scenes, cameras)]
        self.densify_every = int(densify_every)
        self.densify_grad, self.densify_min_opacity = densify_grad, densify_min_opacity
        self.densify_options = dict(densify_options or {})
        self.generator = generator
        self.densify_log: list = []
        self._size_workspace(capacity)
        pix = int(self.caps.pixels)
        self.target = torch.zeros(3 * pix, dtype=torch.float32, device=device)
        if targets is not None:
            o = 0
            for tl in targets:
                for t in tl:
                    n = t.numel()
                    self.target[o:o + n] = t.reshape(-1)
                    o += n
        self.dL_dcolor = torch.empty(3 * pix, dtype=torch.float32, device=device)
        self.loss = torch.zeros(max(int(self.caps.frames), 1), dtype=torch.float32, device=device)
        self.tile_loss = torch.zeros(max(int(self.caps.tiles), 1), dtype=torch.float32, device=device)
        self.graph = None
        self._recapture = False
        self._fresh_ws, self._prev_overflow = False, False
        # sticky overflow: the Adam launch skips an overflowed step (its
        # images and gradients are incomplete) and sets this word; step()
        # reads it back without waiting and grows the workspace when set
        self.sticky = torch.zeros(1, dtype=torch.int32, device=device)
        self._sticky_host = torch.zeros(1, dtype=torch.int32).pin_memory()
        self._sticky_ev = None
        self.grow_log: list = []
        self.it = 0
        st = self.store
        self._reset_stats()
        self._period = np.array([0, 3 * st.M, 0, 0, 0], np.int32)
        self._split = np.array([0, 3, 0, 0, 0], np.int32)

    def _size_workspace(self, capacity: int | None = None):
        """Workspace for the current scenes; without a fixed capacity, size the
        instance buffers from one forward pass (K is counted in full even when
        it overflows the probe capacity)."""
        items = sum(sp.n * len(sp.frames) for sp in self.specs)
        cap = capacity if capacity is not None else min(MAX_CAPACITY, max(1 << 20, 8 * items))
        self._alloc(cap)
        if capacity is None:
            self._forward()
            K, ovf = (int(v) for v in self.ws.status().cpu())
            if K >= MAX_CAPACITY:
                raise RuntimeError(f"batch needs {K} tile instances; split it (limit {MAX_CAPACITY})")
            want = min(MAX_CAPACITY, int(K * 1.15) + 4096)
            if ovf or want > 2 * cap or want < cap // 2:
                self._alloc(want)

    def _reset_stats(self):
        G = max(self.store.G, 1) if self.densify_every else 1
        self.accum = torch.zeros(G, dtype=torch.float32, device=self.device)
        self.count = torch.zeros(G, dtype=torch.float32, device=self.device)

    def _alloc(self, capacity: int):
        self.caps = batch.caps_for(self.specs, self.tile, capacity, depth_split=self.depth_split,
                                   chunk=self.chunk)
        self.ws = None
        self.grads = None
        self.ws = batch.BatchWorkspace(self.caps, self.device)
        self.ws.status().zero_()
        self.ws.load_table(self.specs)
        self.grads = batch.BatchGrads(self.caps, 0, self.device)

    # ---- the launch sequence ----
    def _forward(self):
        L = _lib.load()
        st, ws = self.store, self.ws
        s = torch.cuda.current_stream().cuda_stream
        _lib.check(L.w2w_activate(st.view("svec").data_ptr(), st.view("opacity").data_ptr(),
                                  st.view("quaternion").data_ptr(), st.G, st.scales.data_ptr(),
                                  st.opac.data_ptr(), st.rots.data_ptr(), s), "w2w_activate")
        _lib.check(L.w2w_forward_batch(ws.table.data_ptr(), ctypes.byref(self.caps),
                                       ctypes.byref(ws.struct), st.view("means").data_ptr(),
                                       st.scales.data_ptr(), st.rots.data_ptr(), st.opac.data_ptr(),
                                       st.group(st.flat, "sh").data_ptr(), ws.color.data_ptr(),
                                       ws.depth.data_ptr(), ws.alpha.data_ptr(),
                                       ws.radii.data_ptr() if self.densify_every else None, s),
                   "w2w_forward_batch")

    def _body(self):
        L = _lib.load()
        st, ws, g = self.store, self.ws, self.grads
        s = torch.cuda.current_stream().cuda_stream
        G = st.G
        means = st.view("means")
        sh = st.group(st.flat, "sh")
        self._forward()
        _lib.check(L.w2w_l2_loss(ws.table.data_ptr(), ctypes.byref(self.caps), ws.color.data_ptr(),
                                 self.target.data_ptr(), self.dL_dcolor.data_ptr(),
                                 self.loss.data_ptr(), self.tile_loss.data_ptr(), s), "w2w_l2_loss")
        skip = ws.geom.data_ptr() + 4  # the forward's overflow word: an overflowed step updates nothing
        if self.fused_sh_adam:
            _lib.check(L.w2w_backward_batch_sh_adam(
                ws.table.data_ptr(), ctypes.byref(self.caps), ctypes.byref(ws.struct), means.data_ptr(),
                st.scales.data_ptr(), st.rots.data_ptr(), sh.data_ptr(), self.dL_dcolor.data_ptr(), None,
                None, g.means2D.data_ptr(), g.colors.data_ptr(), st.g_opac.data_ptr(),
                st.view("means", st.grad).data_ptr(), st.g_cov.data_ptr(),
                st.group(st.exp_avg, "sh").data_ptr(), st.group(st.exp_avg_sq, "sh").data_ptr(),
                st.g_scales.data_ptr(), st.g_rots.data_ptr(), st.hyper.data_ptr(), 1, 9, skip, s),
                "w2w_backward_batch_sh_adam")
        else:
            _lib.check(L.w2w_backward_batch(
                ws.table.data_ptr(), ctypes.byref(self.caps), ctypes.byref(ws.struct), means.data_ptr(),
                st.scales.data_ptr(), st.rots.data_ptr(), sh.data_ptr(), self.dL_dcolor.data_ptr(), None,
                None, g.means2D.data_ptr(), g.colors.data_ptr(), st.g_opac.data_ptr(),
                st.view("means", st.grad).data_ptr(), st.g_cov.data_ptr(),
                st.group(st.grad, "sh").data_ptr(), st.g_scales.data_ptr(), st.g_rots.data_ptr(), s),
                "w2w_backward_batch")
        if self.densify_every:
            _lib.check(L.w2w_densify_stats(ws.table.data_ptr(), ctypes.byref(self.caps),
                                           g.means2D.data_ptr(), ws.radii.data_ptr(), self.accum.data_ptr(),
                                           self.count.data_ptr(), s), "w2w_densify_stats")
        _lib.check(L.w2w_activate_backward(
            st.scales.data_ptr(), st.opac.data_ptr(), st.view("quaternion").data_ptr(),
            st.g_scales.data_ptr(), st.g_opac.data_ptr(), st.g_rots.data_ptr(), G,
            st.view("svec", st.grad).data_ptr(), st.view("opacity", st.grad).data_ptr(),
            st.view("quaternion", st.grad).data_ptr(), s), "w2w_activate_backward")
        if not self.fused_sh_adam:
            _lib.check(L.w2w_adam_step_hyper(st.flat.data_ptr(), st.grad.data_ptr(),
                                             st.exp_avg.data_ptr(), st.exp_avg_sq.data_ptr(),
                                             int(st.ends[-1]), len(GROUPS), st.ends.ctypes.data,
                                             self._period.ctypes.data, self._split.ctypes.data,
                                             st.hyper.data_ptr(), skip, self.sticky.data_ptr(), s),
                       "w2w_adam_step_hyper")
            return
        # the SH block took its step in the backward: the means, then the
        # groups after the SH block (their rates at hyper[32 + k])
        _lib.check(L.w2w_adam_step_hyper(st.flat.data_ptr(), st.grad.data_ptr(),
                                         st.exp_avg.data_ptr(), st.exp_avg_sq.data_ptr(),
                                         int(st.ends[0]), 1, st.ends.ctypes.data, None, None,
                                         st.hyper.data_ptr(), skip, self.sticky.data_ptr(), s),
                   "w2w_adam_step_hyper")
        b = int(st.ends[1])
        self._tail_ends = (st.ends[2:5] - b).astype(np.int64)
        _lib.check(L.w2w_adam_step_hyper(st.flat.data_ptr() + 4 * b, st.grad.data_ptr() + 4 * b,
                                         st.exp_avg.data_ptr() + 4 * b, st.exp_avg_sq.data_ptr() + 4 * b,
                                         int(st.ends[-1]) - b, 3, self._tail_ends.ctypes.data, None, None,
                                         st.hyper.data_ptr() + 4 * 32, skip, self.sticky.data_ptr(), s),
                   "w2w_adam_step_hyper")

    def _upload_hyper(self):
        self.it += 1
        st = self.store
        h = st.hyper_values(self.it, dict(LRS, means=mean_lr(self.it), sh_dc=LRS["sh_dc"],
                                          sh_ac=LRS["sh_ac"]))
        st.upload_hyper(h)

    def step(self):
        """One fwd+bwd+Adam step of every scene (graph replay when captured),
        then the densification cycle when one is due.  A step whose instance
        capacity overflowed updates no parameter; the trainer notices it a
        step or two later (without waiting on the device) and grows the
        workspace."""
        if self._sticky_ev is not None and self._sticky_ev.query() and int(self._sticky_host[0]):
            self.grow()
        self._upload_hyper()
        self._fresh_ws = False
        if self.graph is not None:
            self.graph.replay()
        else:
            self._body()
            if self._recapture:
                self._recapture = False
                self.capture()
        self._sticky_host.copy_(self.sticky, non_blocking=True)
        self._sticky_ev = torch.cuda.Event()
        self._sticky_ev.record()
        if self.densify_every and self.it % self.densify_every == 0:
            self.densify()
        return self.loss

    def skipped_overflow(self) -> bool:
        """Whether any step since the workspace last grew overflowed its
        instance capacity (and so updated nothing).  Synchronises."""
        return bool(int(self.sticky.cpu()[0]))

    def grow(self):
        """Re-size the instance capacity from a forward of the current scenes,
        clear the sticky overflow word and re-capture the graph if there was
        one."""
        old = int(self.caps.capacity)
        torch.cuda.synchronize()
        self._size_workspace()
        self.sticky.zero_()
        self._sticky_host.zero_()
        self._sticky_ev = None
        self._fresh_ws = True
        self._prev_overflow = False
        if self.graph is not None:
            self.graph = None
            self._recapture = True
        self.grow_log.append((self.it, old, int(self.caps.capacity)))

    def densify(self, normals: torch.Tensor | None = None):
        """One densification cycle of every scene on the device; the store,
        the workspace (re-sized for the new Gaussian counts) and, when one was
        captured, the graph (re-captured after the next eager step) follow.
        normals: the split children's standard normal samples, scene by scene
        (drawn with the trainer's generator when None)."""
        from . import densify as dens
        st = self.store
        self._prev_overflow = self.overflow()
        cfg = dens.config(self.densify_grad, self.densify_min_opacity, **self.densify_options)
        p, m, v, ns, counts = dens.densify_segments([3, 3 * st.M, 1, 3, 4], (0, 2, 3, 4), st.flat,
                                                    st.exp_avg, st.exp_avg_sq, st.ns, self.accum,
                                                    self.count, cfg, self.generator, normals)
        st.rebuild(p, m, v, ns)
        for sp, n in zip(self.specs, ns):
            sp.n = n
        self._size_workspace()
        self._fresh_ws = True
        self._reset_stats()
        if self.graph is not None:
            self.graph = None
            self._recapture = True
        self.densify_log.append((self.it, counts))
        return counts

    def capture(self):
        """Capture the launch sequence into a HIP graph (after one eager step
        has run, so every buffer exists)."""
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._body()
        torch.cuda.synchronize()
        self.graph = g

    def overflow(self) -> bool:
        """Whether the last step overflowed the instance capacity (its images
        and gradients are then invalid).  After a densification cycle the
        workspace is new: the answer then covers the step before the cycle."""
        if self._fresh_ws:
            return self._prev_overflow
        return bool(int(self.ws.status()[1]))

    def frame_images(self):
        return self.ws.frame_views(self.specs)
