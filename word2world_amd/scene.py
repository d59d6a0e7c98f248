"""Seeded synthetic Gaussian scenes with the parameter distributions of the
benchmark workloads.

Parameters are the *raw* (pre-activation) values held by the reference's
``GaussiansHandler`` (gaussians_handler.py:27-50): ``means`` [N,3], ``svec``
log-scales [N,3], ``quaternion`` [N,4] (w,x,y,z), ``opacity`` logits [N,1],
``sh_dc`` [N,1,3] and ``sh_ac`` [N,M-1,3].  ``activate`` applies what
``Renderer.py:164-167`` applies before the rasterizer call.

Distributions: means uniform on the sphere of radius 0.5 plus N(0, 0.02)
jitter; log-scales N(-4, 0.5); quaternions normalise(N(0,1)^4); opacity
logits N(0,1); SH DC N(0, 0.3); SH coefficients beyond DC N(0, 0.1).
Tensors are generated on the CPU with a seeded ``torch.Generator``.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch


@dataclass
class GaussianScene:
    means: torch.Tensor
    svec: torch.Tensor
    quaternion: torch.Tensor
    opacity: torch.Tensor
    sh_dc: torch.Tensor
    sh_ac: torch.Tensor
    sh_degree: int

    @property
    def n(self) -> int:
        return int(self.means.shape[0])

    def to(self, device) -> "GaussianScene":
        f = lambda t: t.to(device)
        return GaussianScene(f(self.means), f(self.svec), f(self.quaternion), f(self.opacity),
                             f(self.sh_dc), f(self.sh_ac), self.sh_degree)

    def params(self) -> dict:
        return {"means": self.means, "sh_dc": self.sh_dc, "sh_ac": self.sh_ac,
                "opacity": self.opacity, "svec": self.svec, "quaternion": self.quaternion}


def make_scene(n: int, sh_degree: int, seed: int, n_coeffs: int | None = None) -> GaussianScene:
    """A seeded scene; ``n_coeffs`` defaults to (degree+1)^2."""
    m = (sh_degree + 1) ** 2 if n_coeffs is None else n_coeffs
    g = torch.Generator().manual_seed(seed)
    d = torch.randn(n, 3, generator=g)
    d = d / d.norm(dim=1, keepdim=True).clamp_min(1e-12)
    means = 0.5 * d + 0.02 * torch.randn(n, 3, generator=g)
    svec = -4.0 + 0.5 * torch.randn(n, 3, generator=g)
    q = torch.randn(n, 4, generator=g)
    q = q / q.norm(dim=1, keepdim=True).clamp_min(1e-12)
    opacity = torch.randn(n, 1, generator=g)
    sh_dc = 0.3 * torch.randn(n, 1, 3, generator=g)
    sh_ac = 0.1 * torch.randn(n, m - 1, 3, generator=g)
    if sh_degree == 0:
        sh_ac.zero_()
    return GaussianScene(means.float(), svec.float(), q.float(), opacity.float(), sh_dc.float(),
                         sh_ac.float(), sh_degree)


def activate(scene_params: dict):
    """``Renderer.py:164-167``: sigmoid, exp, normalize, cat(dc, ac)."""
    opac = torch.sigmoid(scene_params["opacity"])
    scales = torch.exp(scene_params["svec"])
    rots = torch.nn.functional.normalize(scene_params["quaternion"])
    sh = torch.cat((scene_params["sh_dc"], scene_params["sh_ac"]), dim=1)
    return scene_params["means"], scales, rots, opac, sh


def save_scene(path: str, scene: GaussianScene) -> None:
    """Writes the raw parameters as a Gaussian PLY (GaussianIO.save_as_ply)."""
    from .ply import save_ply
    save_ply(path, scene.means, scene.sh_dc, scene.sh_ac, scene.opacity, scene.svec, scene.quaternion)


def load_scene(path: str, sh_degree: int | None = None) -> GaussianScene:
    """Reads a Gaussian PLY.  The SH degree follows from the AC coefficient
    count ((degree+1)^2 - 1 per channel) unless given."""
    from .ply import load_ply
    d = load_ply(path)
    K = int(d["sh_ac"].shape[1])
    if sh_degree is None:
        deg = int(round((K + 1) ** 0.5)) - 1
        if (deg + 1) ** 2 != K + 1:
            raise ValueError(f"{K} AC coefficients per channel is not a full SH degree; pass sh_degree")
        sh_degree = deg
    elif (sh_degree + 1) ** 2 > K + 1:
        raise ValueError(f"degree {sh_degree} needs {(sh_degree + 1) ** 2 - 1} AC coefficients, file has {K}")
    t = lambda a: torch.from_numpy(a)
    return GaussianScene(t(d["means"]), t(d["svec"]), t(d["quaternion"]), t(d["opacity"]).reshape(-1, 1),
                         t(d["sh_dc"]), t(d["sh_ac"]), sh_degree)
