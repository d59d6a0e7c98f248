"""Cameras in the exact form the reference's trainer hands to the rasterizer.

Follows ``GaussianTrainer.py:50-83`` (orbital look-at pose, inverse, row
flips, transposed storage), ``misc_utils.py:143-173`` (projection matrix),
``DynamicCamera.py:5-11`` and ``utils_for_camera.py:11-12`` (fovy 49.1 deg,
near 0.01, far 100, fovx from the aspect ratio).

The 4x4 matrices are stored the way ``Renderer.py:135-136`` passes them:
``viewmatrix`` is ``W2C^T`` and ``projmatrix`` is ``W2C^T @ P^T``, both
row-major, i.e. the column-major storage of ``W2C`` and ``P @ W2C``.
``campos`` is the caller's value, ``-c2w[:3, 3]`` (GaussianTrainer.py:83); the
rasterizer uses it as given for the SH view direction.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch


def projection_matrix(z_near: float, z_far: float, fov_x: float, fov_y: float) -> torch.Tensor:
    """``misc_utils.get_projection_matrix`` (misc_utils.py:143-173), float32."""
    tx = math.tan(fov_x / 2.0)
    ty = math.tan(fov_y / 2.0)
    p = torch.zeros(4, 4)
    p[0, 0] = 1.0 / tx
    p[1, 1] = 1.0 / ty
    p[2, 2] = z_far / (z_far - z_near)
    p[2, 3] = -(z_far * z_near) / (z_far - z_near)
    p[3, 2] = 1.0
    return p


def _normalize32(v: np.ndarray) -> np.ndarray:
    """``utils_for_camera.vector_ops(..., normalize=True)`` in float32."""
    t = torch.tensor(v, dtype=torch.float32)
    n = torch.sqrt(torch.clamp(torch.dot(t, t), min=1e-20))
    return (t / n).numpy()


@dataclass
class Camera:
    width: int
    height: int
    fovx: float
    fovy: float
    viewmatrix: torch.Tensor  # [4,4] float32, W2C^T
    projmatrix: torch.Tensor  # [4,4] float32, W2C^T @ P^T
    campos: torch.Tensor      # [3] float32, -c2w[:3,3]
    c2w: np.ndarray           # [4,4] float32 pose

    @property
    def tanfovx(self) -> float:
        return math.tan(self.fovx * 0.5)

    @property
    def tanfovy(self) -> float:
        return math.tan(self.fovy * 0.5)

    @property
    def focal_x(self) -> float:
        """``Renderer.py:39`` with the requested width instead of 800 (PD-1)."""
        return float(np.float32(self.width / (2.0 * self.tanfovx)))

    @property
    def focal_y(self) -> float:
        return float(np.float32(self.height / (2.0 * self.tanfovy)))

    def view_flat(self) -> np.ndarray:
        return np.ascontiguousarray(self.viewmatrix.numpy().reshape(-1), dtype=np.float32)

    def proj_flat(self) -> np.ndarray:
        return np.ascontiguousarray(self.projmatrix.numpy().reshape(-1), dtype=np.float32)

    def campos_np(self) -> np.ndarray:
        return np.ascontiguousarray(self.campos.numpy(), dtype=np.float32)


def orbit_camera(azimuth_deg: float, elevation_deg: float, radius: float, width: int, height: int,
                 fovy_deg: float = 49.1, z_near: float = 0.01, z_far: float = 100.0) -> Camera:
    """One camera of ``GaussianTrainer.render_training_views`` (GaussianTrainer.py:50-83)."""
    el = np.deg2rad(elevation_deg)
    az = np.deg2rad(azimuth_deg)
    x = radius * np.cos(el) * np.sin(az)
    y = -radius * np.sin(el)
    z = radius * np.cos(el) * np.cos(az)
    target = np.zeros(3, dtype=np.float32)
    campos = np.array([x, y, z]) + target
    pose = np.eye(4, dtype=np.float32)
    forward = _normalize32(campos - target)
    up = np.array([0, 1, 0], dtype=np.float32)
    right = _normalize32(np.cross(up, forward))
    up = _normalize32(np.cross(forward, right))
    pose[:3, :3] = np.stack([right, up, forward], axis=1)
    pose[:3, 3] = campos
    c2w = torch.tensor(pose, dtype=torch.float32)
    fovy = float(torch.deg2rad(torch.tensor(fovy_deg, dtype=torch.float32)))
    fovx = float(2 * torch.atan(torch.tan(torch.tensor(fovy) / 2) * width / height))
    w2c = torch.linalg.inv(c2w)
    w2c[1:3, :3] *= -1
    w2c[:3, 3] *= -1
    view = w2c.transpose(0, 1).contiguous()
    proj = projection_matrix(z_near, z_far, fovx, fovy).transpose(0, 1).contiguous()
    full = view @ proj
    cc = -c2w[:3, 3]
    return Camera(width, height, fovx, fovy, view, full.contiguous(), cc.contiguous(), pose)


def orbit_views(width: int, height: int, n_views: int = 4, radius: float = 2.5,
                elevation_deg: float | list = 0.0, azimuth0: float = 0.0) -> list[Camera]:
    """``n_views`` cameras at azimuth0 + k*360/n (the trainer's 0/90/180/270 for n=4)."""
    els = elevation_deg if isinstance(elevation_deg, (list, tuple)) else [elevation_deg] * n_views
    return [orbit_camera(azimuth0 + k * 360.0 / n_views, els[k], radius, width, height)
            for k in range(n_views)]
