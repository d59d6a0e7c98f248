"""Shared scene/camera/oracle plumbing for the tests."""
import numpy as np
import torch

from oracle import oracle
from word2world_amd.camera import orbit_views
from word2world_amd.scene import activate, make_scene

BG = np.ones(3, np.float32)


def scene_np(n, deg, seed, scale_mul=1.0):
    sc = make_scene(n, deg, seed)
    m, s, r, o, sh = [t.numpy().copy() for t in activate(sc.params())]
    return m, s * np.float32(scale_mul), r, o.reshape(-1), sh


def cam_args(cam):
    return (cam.view_flat(), cam.proj_flat(), cam.campos_np(), BG)


def oracle_forward(arrs, deg, cam, ts=16, record=0):
    m, s, r, o, sh = arrs
    return oracle.forward(m, s, r, o, sh, deg, *cam_args(cam), cam.width, cam.height, ts,
                          cam.focal_x, cam.focal_y, record=record)


def gpu_camera(cam, device):
    from word2world_amd.raster import camera_from_tensors
    return camera_from_tensors(cam.viewmatrix, cam.projmatrix, cam.campos, cam.width, cam.height,
                               cam.focal_x, cam.focal_y, BG, device=device)


def to_dev(arrs, device):
    return [torch.from_numpy(np.ascontiguousarray(a)).to(device) for a in arrs]


def unpack_rect(rect_i32):
    r = rect_i32.astype(np.int64) & 0xffffffff
    return np.stack([r[:, 0] & 0xffff, r[:, 0] >> 16, r[:, 1] & 0xffff, r[:, 1] >> 16], 1)
