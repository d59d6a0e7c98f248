"""Shared scene/camera/oracle plumbing for the tests.

Parity bars: integer outputs and the final transmittance bit-exact; images
within IMG_TOL max-abs; gradients within GRAD_TOL of the reference, measured
as max|g - g_ref| <= GRAD_TOL * max|g_ref| per tensor (rel_err)."""
import numpy as np
import torch

from oracle import oracle
from word2world_amd import raster
from word2world_amd.camera import orbit_views
from word2world_amd.scene import activate, make_scene

BG = np.ones(3, np.float32)
IMG_TOL = 1e-4
GRAD_TOL = 1e-3


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


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def run_gpu(arrs, deg, cam, dev, ts):
    m, s, r, o, sh = to_dev(arrs, dev)
    gcam = gpu_camera(cam, dev)
    K, color, depth, alpha, radii, ws = raster.forward_checked(m, s, r, o, sh, deg, gcam, ts)
    torch.cuda.synchronize()
    keys = raster.sorted_keys(m, s, r, o, sh, deg, gcam, ts, ws, K).cpu().numpy().view(np.uint64)
    return dict(K=K, color=color.cpu().numpy(), depth=depth.cpu().numpy(),
                alpha=alpha.cpu().numpy(), radii=radii.cpu().numpy(),
                counts=ws.counts().cpu().numpy().view(np.uint32),
                rect=unpack_rect(ws.rect().cpu().numpy()), keys=keys,
                ids=ws.sorted_ids(K).cpu().numpy().view(np.uint32),
                ranges=ws.ranges().cpu().numpy().view(np.uint32),
                n_contrib=ws.n_contrib().cpu().numpy().view(np.uint32),
                final_T=ws.final_T().cpu().numpy(), splats=ws.splats().cpu().numpy(),
                ws=ws, gcam=gcam, dev_arrs=(m, s, r, o, sh))


def compare_forward(f, g, ts):
    print(f"K oracle={f['K']} gpu={g['K']}", flush=True)
    np.testing.assert_array_equal(g["radii"], f["radii"])
    np.testing.assert_array_equal(g["counts"], f["counts"])
    vis = f["counts"] > 0
    np.testing.assert_array_equal(g["rect"][vis], f["rect"][vis])
    assert g["K"] == f["K"]
    np.testing.assert_array_equal(g["keys"], f["keys"])
    np.testing.assert_array_equal(g["ids"], f["vals"])
    np.testing.assert_array_equal(g["ranges"], f["ranges"])
    np.testing.assert_array_equal(g["n_contrib"], f["n_contrib"])
    np.testing.assert_array_equal(g["final_T"], f["final_T"])
    # per-Gaussian preprocess outputs of visible Gaussians: bit-exact too
    sp = g["splats"]
    np.testing.assert_array_equal(sp[vis, 0:2], f["m2d"][vis])
    np.testing.assert_array_equal(sp[vis, 2:4], f["conic"][vis, 0:2])
    np.testing.assert_array_equal(sp[vis, 4], f["conic"][vis, 2])
    np.testing.assert_array_equal(sp[vis, 6], f["depth"][vis])
    assert np.abs(g["color"] - f["color"]).max() <= IMG_TOL
    assert np.abs(g["alpha"] - f["alpha"]).max() <= IMG_TOL
    assert np.abs(g["depth"] - f["depth_img"]).max() <= IMG_TOL
