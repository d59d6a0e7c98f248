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
  rows are dropped).

Instead of six separate leaves the six groups are views of one flat float32
buffer, so autograd accumulates every gradient into one flat buffer and one
fused kernel updates all groups.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, raster
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
    def _rebuild(self, rows: dict, keep: torch.Tensor | None, new_rows: dict | None):
        """Re-layout the store: keep[i] rows stay (in order), new rows are
        appended with zero Adam moments (OptimizerManager.py:23-45)."""
        old = {k: v.detach() for k, v in self.views().items()}
        om = self.views(self.exp_avg)
        ov = self.views(self.exp_avg_sq)
        parts, pm, pv = {}, {}, {}
        for k in GROUPS:
            p, m1, m2 = old[k], om[k], ov[k]
            if keep is not None:
                p, m1, m2 = p[keep], m1[keep], m2[keep]
            if new_rows is not None:
                p = torch.cat([p, new_rows[k].to(p.dtype)], 0)
                z = torch.zeros_like(new_rows[k], dtype=p.dtype)
                m1, m2 = torch.cat([m1, z], 0), torch.cat([m2, z], 0)
            parts[k], pm[k], pv[k] = p, m1, m2
        self.n = int(parts["means"].shape[0])
        self.flat = torch.cat([parts[k].reshape(-1) for k in GROUPS]).contiguous().requires_grad_(True)
        self.exp_avg = torch.cat([pm[k].reshape(-1) for k in GROUPS]).contiguous()
        self.exp_avg_sq = torch.cat([pv[k].reshape(-1) for k in GROUPS]).contiguous()

    @torch.no_grad()
    def densify(self, grad_thresh: float, min_opacity: float, generator: torch.Generator | None = None):
        """One adc_cycle (gaussians_handler.py:185-194): clone, split, reset
        statistics, prune.  Returns (n_clone, n_split, n_prune)."""
        dev = self.device
        # clone (Densification.py:33-52, clone_cycle threshold 0.04)
        v = self.views()
        scales = torch.exp(v["svec"])
        g = self.grad_accum / (self.grad_count + 1e-5)
        clone = (g.abs() >= grad_thresh) & (scales <= 0.04).any(dim=-1)
        n_clone = int(clone.sum())
        if n_clone:
            self._rebuild(None, None, {k: t.detach()[clone] for k, t in v.items()})
            self.grad_accum = torch.cat([self.grad_accum, torch.zeros(n_clone, device=dev)])
            self.grad_count = torch.cat([self.grad_count, torch.zeros(n_clone, device=dev)])
        # split (Densification.py:56-107, split_cycle threshold 4, factor 2, shrink 0.8)
        v = self.views()
        scales = torch.exp(v["svec"])
        g = self.grad_accum / (self.grad_count + 1e-5)
        split = (g.abs() >= grad_thresh) & (scales > 4.0).any(dim=-1)
        n_split = int(split.sum())
        if n_split:
            sel = {k: t.detach()[split] for k, t in v.items()}
            rep = lambda t: t.repeat((2,) + (1,) * (t.dim() - 1))
            q = rep(sel["quaternion"])
            qn = q / q.norm(dim=1, keepdim=True)
            w, x, y, z = qn.unbind(1)
            R = torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                             2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                             2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
                            1).view(-1, 3, 3)
            std = rep(scales[split])
            sample = torch.normal(torch.zeros_like(std), std, generator=generator)
            new = {"means": (R @ sample.unsqueeze(-1)).squeeze(-1) + rep(sel["means"]),
                   "sh_dc": rep(sel["sh_dc"]), "sh_ac": rep(sel["sh_ac"]),
                   "opacity": rep(sel["opacity"]), "svec": torch.log(std / (2 * 0.8)),
                   "quaternion": q}
            self._rebuild(None, None, new)
            keep = torch.cat([~split, torch.ones(2 * n_split, dtype=torch.bool, device=dev)])
            self._rebuild(None, keep, None)
        self.grad_accum = torch.zeros(self.n, device=dev)
        self.grad_count = torch.zeros(self.n, device=dev)
        # prune (Densification.py:167-187; prune_cycle passes svec_thresh 4 -> 0.4)
        v = self.views()
        opac = torch.sigmoid(v["opacity"]).squeeze(-1)
        scales = torch.exp(v["svec"])
        prune = (opac < min_opacity) | (scales.max(dim=1).values > 0.1 * 4)
        n_prune = int(prune.sum())
        if n_prune:
            self._rebuild(None, ~prune, None)
            self.grad_accum = self.grad_accum[~prune]
            self.grad_count = self.grad_count[~prune]
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

