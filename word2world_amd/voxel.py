"""Occupancy field for mesh export (compute_voxel_field_wrapper.py, on the
device): `compute_voxel_fields` replaces `computeVoxelFeildsCuda`
(compute_voxel_field_wrapper.py:75-107) and `prepare_gaussians_for_extraction`
mirrors the reference helper of the same name (:111-130).

The field is the one of compute_voxel_field.cu:25-79 (w2w_voxel_field): for
every grid point, the opacity-weighted Gaussian density of the Gaussians
whose means lie in the relaxed box of the point's block.  Inputs stay on the
device; the reference copies them to the host and back.
"""
from __future__ import annotations

import torch

from . import _lib


def voxel_field(means: torch.Tensor, covs: torch.Tensor, opacities: torch.Tensor, resolution: int,
                num_blocks: int, relax_ratio: float) -> torch.Tensor:
    """means [N,3], covs [N,6] (xx, xy, xz, yy, yz, zz), opacities [N] or
    [N,1], device float32.  Returns occ [res, res, res] indexed [iz, iy, ix]
    (the reference's reshape of its flat ix + iy res + iz res^2 array)."""
    dev = means.device
    if dev.type != "cuda":
        raise RuntimeError("voxel_field runs on the device (HIP); inputs are on " + str(dev))
    L = _lib.load()
    m = means.detach().float().contiguous()
    c = covs.detach().float().contiguous()
    o = opacities.detach().float().reshape(-1).contiguous()
    n = int(m.shape[0])
    if m.shape != (n, 3) or c.shape != (n, 6) or o.shape != (n,):
        raise ValueError("voxel_field: means [N,3], covs [N,6], opacities [N]")
    occ = torch.empty(resolution ** 3, dtype=torch.float32, device=dev)
    tb = int(L.w2w_voxel_temp_bytes(n, int(num_blocks)))
    if tb == 0:
        raise ValueError(f"voxel_field: unsupported num_blocks {num_blocks}")
    temp = torch.empty(tb, dtype=torch.uint8, device=dev)
    _lib.check(L.w2w_voxel_field(m.data_ptr() if n else None, c.data_ptr() if n else None,
                                 o.data_ptr() if n else None, n, int(resolution), int(num_blocks),
                                 float(relax_ratio), occ.data_ptr(), temp.data_ptr(), tb,
                                 torch.cuda.current_stream().cuda_stream), "w2w_voxel_field")
    return occ.view(resolution, resolution, resolution)


def compute_voxel_fields(means, covs, opacities, resolution, num_blocks, relax_ratio) -> torch.Tensor:
    """computeVoxelFeildsCuda (compute_voxel_field_wrapper.py:75-107), same
    arguments and result layout."""
    return voxel_field(means, covs, opacities, resolution, num_blocks, relax_ratio)


def _rotation(q: torch.Tensor) -> torch.Tensor:
    """misc_utils.quaternion_to_rotation_matrix (misc_utils.py:82-112)."""
    norm = torch.sqrt(q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2] + q[:, 3] * q[:, 3])
    q = q / norm[:, None]
    w, x, y, z = q.unbind(1)
    return torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                        2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                        2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1).view(-1, 3, 3)


@torch.no_grad()
def prepare_gaussians_for_extraction(means, stds, quaternions, opacities, opacity_threshold):
    """compute_voxel_field_wrapper.prepare_gaussians_for_extraction: keep
    opacities above the threshold, centre and scale the means into
    [-0.9, 0.9], covariance R S S^T R^T as (xx, xy, xz, yy, yz, zz).  Returns
    (means, covs, opacities, scale, center)."""
    mask = (opacities > opacity_threshold).squeeze(-1)
    means, stds, quaternions, opacities = means[mask], stds[mask], quaternions[mask], opacities[mask]
    mn, mx = means.amin(0), means.amax(0)
    center = (mn + mx) / 2
    scale = 1.8 / (mx - mn).amax().item()
    means = (means - center) * scale
    stds = stds * scale
    Lm = _rotation(quaternions) @ torch.diag_embed(stds)
    cov = Lm @ Lm.transpose(1, 2)
    covs = torch.stack([cov[:, 0, 0], cov[:, 0, 1], cov[:, 0, 2], cov[:, 1, 1], cov[:, 1, 2], cov[:, 2, 2]], 1)
    return means, covs, opacities, scale, center
