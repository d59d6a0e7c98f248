"""Step-level training harness on the HIP rasterizer: parameters and Adam
state resident in one flat HBM store, a fused Adam kernel, and densification
with Adam-state remap.

Follows the reference trainer:
* parameter groups and learning rates: gaussians_handler.py:27-50 and
  optimizer_config.json (means 0.01, sh_dc 0.01, sh_ac 0.0005, opacity 0.05,
  svec 0.005, quaternion 0.005); Adam eps 1e-15;
* mean learning-rate schedule: exponential from 0.01 to 0.0002 over 300 steps
  (misc_utils.py:11-27, gaussians_handler.py:46,52-56);
* activation before the rasterizer: Renderer.py:164-167;
* densification statistics: gaussians_handler.py:116-118 (per visible
  Gaussian, count += 1 and accum += ||dL/dmeans2D[:, :2]||);
* densification cycle: gaussians_handler.py:185-194 with Densification.py
  (clone: grad >= thr and any scale <= 0.04; split: grad >= thr and any scale
  > 4, two samples N(0, scale) rotated about the mean, svec = log(scale/1.6),
  originals removed; prune: opacity < thr or max scale > 0.4) and the Adam
  state remap of OptimizerManager.py:23-45 (new rows get zero moments, pruned
  rows are dropped), as one on-device stream compaction (w2w_densify_*).

Instead of six separate leaves the six groups are views of one flat float32
buffer, so autograd accumulates every gradient into one flat buffer and one
fused kernel updates all groups.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, raster
from . import densify as dens
from .ply import save_ply
from .scene import GaussianScene, load_scene

GROUPS = ("means", "sh_dc", "sh_ac", "opacity", "svec", "quaternion")
LRS = {"means": 0.01, "sh_dc": 0.01, "sh_ac": 0.0005, "opacity": 0.05, "svec": 0.005,
       "quaternion": 0.005}


def mean_lr(step: int, lr_init: float = 0.01, lr_final: float = 0.0002, max_steps: int = 300) -> float:
    """misc_utils.create_exponential_learning_rate_schedule(0.01, 0.0002, 300)."""
    if lr_init == lr_final:
        return lr_init
    if step < 0:
        return 0.0
    t = float(np.clip(step / max_steps, 0, 1))
    return float(np.exp(np.log(lr_init) * (1 - t) + np.log(lr_final) * t))


class GaussianStore:
    """Flat, HBM-resident parameters (+ gradient, Adam moments) of one scene."""

    def __init__(self, params: dict, sh_degree: int, device):
        self.sh_degree = sh_degree
        self.device = torch.device(device)
        self.n = int(params["means"].shape[0])
        self.m_ac = int(params["sh_ac"].shape[1])
        self.flat = torch.cat([params[k].reshape(-1).float() for k in GROUPS]).to(self.device)
        self.flat.requires_grad_(True)
        self.exp_avg = torch.zeros_like(self.flat, requires_grad=False)
        self.exp_avg_sq = torch.zeros_like(self.flat, requires_grad=False)
        self.step = 0
        self.grad_accum = torch.zeros(self.n, device=self.device)
        self.grad_count = torch.zeros(self.n, device=self.device)

    @classmethod
    def from_ply(cls, path: str, device, sh_degree: int | None = None) -> "GaussianStore":
        """A store initialised from a Gaussian PLY (fresh Adam state)."""
        sc = load_scene(path, sh_degree)
        return cls(sc.params(), sc.sh_degree, device)

    def save_ply(self, path: str) -> None:
        """The current raw parameters in the format of GaussianIO.save_as_ply."""
        v = {k: t.detach() for k, t in self.views().items()}
        save_ply(path, v["means"], v["sh_dc"], v["sh_ac"], v["opacity"], v["svec"], v["quaternion"])

    # ---- layout ----
    def sizes(self) -> list[int]:
        n, a = self.n, self.m_ac
        return [3 * n, 3 * n, 3 * a * n, n, 3 * n, 4 * n]

    def group_ends(self) -> np.ndarray:
        return np.cumsum(self.sizes()).astype(np.int64)

    def views(self, t: torch.Tensor | None = None) -> dict:
        t = self.flat if t is None else t
        n, a = self.n, self.m_ac
        shapes = [(n, 3), (n, 1, 3), (n, a, 3), (n, 1), (n, 3), (n, 4)]
        out, o = {}, 0
        for k, sz, sh in zip(GROUPS, self.sizes(), shapes):
            out[k] = t[o:o + sz].view(sh)
            o += sz
        return out

    def activated(self):
        """Renderer.py:164-167 on views of the flat store (differentiable)."""
        v = self.views()
        opac = torch.sigmoid(v["opacity"])
        scales = torch.exp(v["svec"])
        rots = torch.nn.functional.normalize(v["quaternion"])
        sh = torch.cat((v["sh_dc"], v["sh_ac"]), dim=1)
        return v["means"], scales, rots, opac, sh

    # ---- optimizer ----
    def adam_step(self, lrs: dict | None = None, beta1=0.9, beta2=0.999, eps=1e-15) -> None:
        if self.flat.grad is None:
            return
        self.step += 1
        lrs = dict(LRS, **(lrs or {}))
        ends = self.group_ends()
        lr = np.array([lrs[k] for k in GROUPS], np.float32)
        st = _lib.load().w2w_adam_step(self.flat.data_ptr(), self.flat.grad.data_ptr(),
                                       self.exp_avg.data_ptr(), self.exp_avg_sq.data_ptr(),
                                       int(ends[-1]), len(GROUPS), ends.ctypes.data, lr.ctypes.data,
                                       self.step, beta1, beta2, eps,
                                       torch.cuda.current_stream().cuda_stream)
        _lib.check(st, "w2w_adam_step")

    def zero_grad(self):
        self.flat.grad = None

    # ---- densification (gaussians_handler.py:116-194, Densification.py) ----
    @torch.no_grad()
    def collect(self, means2d_grad: torch.Tensor, visible: torch.Tensor):
        """gaussians_handler.collect_densification_info."""
        self.grad_count += visible.float()
        self.grad_accum += torch.where(visible, means2d_grad[:, :2].norm(dim=-1),
                                       torch.zeros((), device=self.device))

    @torch.no_grad()
    def densify(self, grad_thresh: float, min_opacity: float, generator: torch.Generator | None = None,
                normals: torch.Tensor | None = None, clone_scale_max: float = 0.04,
                split_scale_min: float = 4.0, prune_scale_max: float = 0.1 * 4, split_factor: int = 2,
                split_shrink: float = 0.8):
        """One adc_cycle (gaussians_handler.py:185-194) on the device: clone,
        split, statistics reset, prune, with the Adam-state remap of
        OptimizerManager.py:23-45 (w2w_densify_plan / w2w_densify_apply).
        `normals`: standard normal samples [>= split_factor * n_split, 3] for
        the split children (drawn with `generator` when not given).  One host
        synchronisation reads the counts.  Returns (n_clone, n_split, n_prune)."""
        cfg = dens.config(grad_thresh, min_opacity, clone_scale_max=clone_scale_max,
                          split_scale_min=split_scale_min, prune_scale_max=prune_scale_max,
                          split_factor=split_factor, split_shrink=split_shrink)
        widths = [3, 3, 3 * self.m_ac, 1, 3, 4]
        p, m, v, ns, counts = dens.densify_segments(widths, (0, 3, 4, 5), self.flat.detach(), self.exp_avg,
                                                    self.exp_avg_sq, [self.n], self.grad_accum,
                                                    self.grad_count, cfg, generator, normals)
        n_clone, n_split, n_prune = counts[0]
        self.n = ns[0]
        self.flat = p.requires_grad_(True)
        self.exp_avg, self.exp_avg_sq = m, v
        self.grad_accum = torch.zeros(self.n, device=self.device)
        self.grad_count = torch.zeros(self.n, device=self.device)
        return n_clone, n_split, n_prune


@dataclass
class TrainConfig:
    sh_degree: int = 3
    tile: int = 16
    densify_every: int = 0          # 0 = never
    densify_grad: float = 2e-4
    densify_min_opacity: float = 0.005


class PlainTrainer:
    """One scene, one view at a time, as GaussianTrainer.optimizaiton_iteration
    does (GaussianTrainer.py:104-114), with an L2-to-target loss replacing the
    SDS guidance."""

    def __init__(self, scene: GaussianScene, cams: list, targets: list, cfg: TrainConfig, device):
        self.store = GaussianStore(scene.params(), scene.sh_degree, device)
        self.cams, self.targets, self.cfg = cams, targets, cfg
        self.it = 0
        self.last_workspaces: list = []

    def step(self) -> torch.Tensor:
        self.it += 1
        st = self.store
        st.zero_grad()
        m, s, r, o, sh = st.activated()
        sh = sh.contiguous()
        loss = torch.zeros((), device=st.device)
        m2ds, radii_all = [], []
        self.last_workspaces = []
        for cam, tgt in zip(self.cams, self.targets):
            m2d = torch.zeros(st.n, 3, device=st.device, requires_grad=True)
            img, radii, depth, alpha = raster.rasterize(m, m2d, sh, o, s, r, cam, self.cfg.sh_degree,
                                                        self.cfg.tile,
                                                        workspace_cb=self.last_workspaces.append)
            loss = loss + ((img.clamp(0, 1) - tgt) ** 2).mean()
            m2ds.append(m2d)
            radii_all.append(radii)
        loss.backward()
        st.adam_step({"means": mean_lr(self.it)})
        if self.cfg.densify_every:
            for m2d, radii in zip(m2ds, radii_all):
                if m2d.grad is not None:
                    st.collect(m2d.grad, radii > 0)
            if self.it % self.cfg.densify_every == 0:
                st.densify(self.cfg.densify_grad, self.cfg.densify_min_opacity)
        return loss.detach()

    def overflow(self) -> bool:
        return any(int(w.status()[1]) for w in self.last_workspaces)

    def observe_capacity(self) -> bool:
        """Synchronise, record each view's instance count with the capacity
        policy (so the next step's workspaces fit) and say whether any view
        of the last step overflowed (its images and gradients are then
        invalid and the step should be discarded)."""
        ovf = False
        for cam, w in zip(self.cams, self.last_workspaces):
            K, o = (int(v) for v in w.status().cpu())
            raster.POLICY.observe((self.store.n, int(cam.width), int(cam.height), self.cfg.tile), K)
            ovf = ovf or bool(o)
        return ovf

