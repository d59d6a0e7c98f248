"""GPU parity: the HIP rasterizer against the CPU oracle on the same seeded
inputs.  Integer outputs (radii, tile counts and rects, K, sorted 64-bit keys,
per-tile ranges, n_contrib) and the final transmittance must be bit-exact;
images within 1e-4 max-abs; gradients within 1e-3 of the float64 autograd
gradient, measured as max|g - g_ref| <= 1e-3 * max|g_ref| per tensor."""
import numpy as np
import pytest
import torch

from helpers import (GRAD_TOL, IMG_TOL, cam_args, compare_forward as _compare_forward, gpu_camera,
                     oracle_forward, rel_err as _rel, run_gpu as _run_gpu, scene_np, to_dev)
from oracle import oracle, torch_ref
from word2world_amd import raster
from word2world_amd.camera import orbit_camera, orbit_views

pytestmark = pytest.mark.gpu

@pytest.mark.parametrize("n,deg,res,ts,scale_mul,seed", [
    (4096, 0, 256, 16, 1.0, 0),     # the CPU-path workload, one view per case below
    (4096, 3, 256, 16, 1.0, 1),
    (3000, 3, 200, 8, 2.0, 2),      # ragged image (200 % 16 != 0), tile 8
    (3000, 1, 190, 32, 2.0, 3),     # tile 32, ragged
    (20000, 2, 384, 16, 1.0, 4),
])
def test_forward_bit_exact(gpu, n, deg, res, ts, scale_mul, seed):
    arrs = scene_np(n, deg, seed, scale_mul)
    for k, cam in enumerate(orbit_views(res, res)):
        print(f"case n={n} deg={deg} res={res} ts={ts} view={k}", flush=True)
        f = oracle_forward(arrs, deg, cam, ts)
        g = _run_gpu(arrs, deg, cam, gpu, ts)
        _compare_forward(f, g, ts)


def test_forward_large_scene_properties(gpu):
    """100k Gaussians, SH3, 512^2 (the one-scene GPU workload): bit-exact
    against the oracle on one view."""
    arrs = scene_np(100_000, 3, 7)
    cam = orbit_views(512, 512)[2]
    f = oracle_forward(arrs, 3, cam, 16)
    g = _run_gpu(arrs, 3, cam, gpu, 16)
    _compare_forward(f, g, 16)


@pytest.mark.parametrize("n,deg,res,ts,scale_mul,seed", [
    (300, 0, 64, 16, 3.0, 11),
    (300, 3, 64, 16, 3.0, 12),
    (2000, 3, 96, 8, 3.0, 13),
    (2000, 2, 100, 32, 3.0, 14),
])
def test_backward_vs_autograd(gpu, n, deg, res, ts, scale_mul, seed):
    arrs = scene_np(n, deg, seed, scale_mul)
    cam = orbit_views(res, res)[1]
    print(f"case n={n} deg={deg} res={res} ts={ts}", flush=True)
    _check_backward(gpu, arrs, deg, cam, ts, seed)


def _check_backward(gpu, arrs, deg, cam, ts, seed, record=1024):
    """GPU backward against float64 autograd over the oracle's recorded
    contributor lists, and the intermediate gradients against the oracle's
    float32 hand backward; returns the oracle forward."""
    m, s, r, o, sh = arrs
    res = cam.width
    f = oracle_forward(arrs, deg, cam, ts, record=record)
    rng = np.random.default_rng(seed)
    gC = rng.standard_normal((3, res, res)).astype(np.float32)
    gA = rng.standard_normal((res, res)).astype(np.float32)
    gD = rng.standard_normal((res, res)).astype(np.float32)
    ref = torch_ref.grads_fixed(f, m, s, r, o, sh, *cam_args(cam), cam.focal_x, cam.focal_y,
                                gC, gA, gD)
    g = _run_gpu(arrs, deg, cam, gpu, ts)
    _compare_forward(f, g, ts)
    dm, ds, dr, do, dsh = g["dev_arrs"]
    out = raster.backward(dm, ds, dr, do, dsh, deg, g["gcam"], ts, g["ws"],
                          torch.from_numpy(gC).to(gpu), torch.from_numpy(gD).to(gpu),
                          torch.from_numpy(gA).to(gpu))
    torch.cuda.synchronize()
    names = ["dL_dmeans2D", "dL_dcolors", "dL_dopacity", "dL_dmeans3D", "dL_dcov3D", "dL_dsh",
             "dL_dscales", "dL_drotations"]
    got = {k: v.cpu().numpy() for k, v in zip(names, out)}
    for k in ["dL_dmeans3D", "dL_dscales", "dL_drotations", "dL_dsh"]:
        e = _rel(got[k].reshape(ref[k].shape), ref[k])
        print(k, e, flush=True)
        assert e <= GRAD_TOL, (k, e)
    e = _rel(got["dL_dopacity"].reshape(-1), ref["dL_dopacity"])
    print("dL_dopacity", e, flush=True)
    assert e <= GRAD_TOL
    # the intermediate gradients against the oracle's float32 hand backward
    ob = oracle.backward(f, m, s, r, sh, *cam_args(cam), cam.focal_x, cam.focal_y, gC, gA, gD)
    for k in ["dL_dmeans2D", "dL_dcolors", "dL_dcov3D"]:
        e = _rel(got[k].reshape(ob[k].shape), ob[k])
        print(k, "vs oracle", e, flush=True)
        assert e <= GRAD_TOL, (k, e)
    return f


def test_dense_faint_scene_backward_vs_autograd(gpu):
    """A dense, faint scene (8,000 large Gaussians of opacity 0.03 at 64^2:
    up to ~500 contributors per pixel, tiles of 2,000 to 83,000 contributing
    pairs) against float64 autograd and the oracle's hand backward: long
    per-pixel walks, the heaviest tiles first in the backward's order, and
    tiles above 2^15 pairs walked by a whole workgroup (one wave per 8x8
    block, sums merged in LDS)."""
    arrs = list(scene_np(8000, 1, 41, 3.0))
    arrs[3] = np.full_like(arrs[3], 0.03)
    cam = orbit_views(64, 64)[1]
    f = _check_backward(gpu, arrs, 1, cam, 16, 41)
    cost = f["rec_count"].reshape(4, 16, 4, 16).sum(axis=(1, 3))
    print("tile costs", cost.ravel(), flush=True)
    assert f["rec_count"].max() >= 400 and cost.max() >= 1 << 16


def test_empty_and_culled(gpu):
    """P=0, every Gaussian behind the near plane, and far off-screen
    Gaussians (the reference's unsigned-wrap case, PD-3)."""
    cam = orbit_views(64, 64)[0]
    gcam = gpu_camera(cam, gpu)
    z = lambda *s: torch.zeros(*s, device=gpu)
    K, color, depth, alpha, radii, ws = raster.forward_checked(z(0, 3), z(0, 3), z(0, 4), z(0), z(0, 1, 3),
                                                               0, gcam, 16)
    assert K == 0
    assert torch.all(color == 1.0) and torch.all(alpha == 0.0)
    arrs = list(scene_np(500, 0, 5))
    behind = arrs[0].copy()
    behind[:] = cam.c2w[:3, 3] * 2.0  # beyond the camera, away from the target
    off = arrs[0].copy() * 0.01
    off[:, 0] += 50.0  # far to one side: projects far off-screen
    for means in (behind, off):
        a = [means] + arrs[1:]
        f = oracle_forward(a, 0, cam, 16)
        g = _run_gpu(a, 0, cam, gpu, 16)
        _compare_forward(f, g, 16)


def test_depth_keys_of_mixed_top_bytes(gpu):
    """A near plane at 1.0 puts the ndc depth of Gaussians nearer than it
    below zero and the others in [0, 0.5): the depth keys' top bytes differ,
    so the depth sort counts and runs its top pass.  Bit-exact against the
    oracle, whose keys order the same float bits as unsigned integers."""
    cam = orbit_camera(30.0, 10.0, 2.5, 128, 128, z_near=1.0)
    arrs = scene_np(3000, 1, 21)
    ndcz = []
    for k in range(2):
        c = cam if k == 0 else orbit_camera(200.0, -20.0, 2.2, 96, 128, z_near=1.0)
        print(f"case mixed top bytes view={k}", flush=True)
        f = oracle_forward(arrs, 1, c, 16)
        g = _run_gpu(arrs, 1, c, gpu, 16)
        _compare_forward(f, g, 16)
        assert g["K"] > 0


def test_dropin_signature(gpu):
    """rasterize_gaussians / rasterize_gaussians_backward with the reference's
    positional arguments (Renderer.py:54-66, :88-93)."""
    arrs = scene_np(1000, 3, 21, 2.0)
    m, s, r, o, sh = to_dev(arrs, gpu)
    cam = orbit_views(96, 96)[3]
    vm, pm, cp = (cam.viewmatrix.to(gpu), cam.projmatrix.to(gpu), cam.campos.to(gpu))
    K, color, depth, alpha, radii, gb, bb, ib = raster.rasterize_gaussians(
        sh, cam.focal_y, s, m, pm, 3, o.reshape(-1, 1), r, cp, vm, cam.focal_x,
        image_width=96, image_height=96)
    f = oracle_forward(arrs, 3, cam, 16)
    assert K == f["K"]
    assert color.shape == (3, 96, 96) and depth.shape == (1, 96, 96) and alpha.shape == (1, 96, 96)
    assert np.abs(color.cpu().numpy() - f["color"]).max() <= IMG_TOL
    gC = torch.randn(3, 96, 96, device=gpu)
    outs = raster.rasterize_gaussians_backward(
        torch.ones(3, device=gpu), m, radii, torch.empty(0, device=gpu), s, r, 1.0,
        torch.empty(0, device=gpu), vm, pm, cam.tanfovx, cam.tanfovy, gC, torch.zeros(1, 96, 96, device=gpu),
        torch.zeros(1, 96, 96, device=gpu), sh, 3, cp, gb, K, bb, ib, alpha, False)
    assert len(outs) == 8
    shapes = [(1000, 3), (1000, 3), (1000, 1), (1000, 3), (1000, 6), (1000, 16, 3), (1000, 3), (1000, 4)]
    assert [tuple(t.shape) for t in outs] == shapes
    # values: the oracle's float32 hand backward on the same inputs (itself
    # pinned against float64 autograd in the CPU tests)
    m_, s_, r_, o_, sh_ = arrs
    zero = np.zeros((96, 96), np.float32)
    ob = oracle.backward(f, m_, s_, r_, sh_, *cam_args(cam), cam.focal_x, cam.focal_y,
                         gC.cpu().numpy(), zero, zero)
    names = ["dL_dmeans2D", "dL_dcolors", "dL_dopacity", "dL_dmeans3D", "dL_dcov3D", "dL_dsh",
             "dL_dscales", "dL_drotations"]
    for k, t in zip(names, outs):
        got = t.cpu().numpy()
        e = _rel(got.reshape(ob[k].shape), ob[k])
        print(k, "drop-in vs oracle", e, flush=True)
        assert e <= GRAD_TOL, (k, e)


class _RendererFunc(torch.autograd.Function):
    """The autograd Function of Renderer.py:43-104 around the drop-in calls:
    the three buffers are saved in ctx and handed back to the backward."""

    @staticmethod
    def forward(ctx, means3D, sh, opacities, scales, rotations, vm, pm, cp, fx, fy, tanx, tany, W, H):
        K, color, depth, alpha, radii, gb, bb, ib = raster.rasterize_gaussians(
            sh, fy, scales, means3D, pm, 3, opacities, rotations, cp, vm, fx,
            image_width=W, image_height=H)
        ctx.K, ctx.cam = K, (vm, pm, cp, tanx, tany)
        ctx.save_for_backward(means3D, radii, sh, scales, rotations, gb, bb, ib, alpha)
        return color, radii, depth, alpha

    @staticmethod
    def backward(ctx, g_color, g_radii, g_depth, g_alpha):
        means3D, radii, sh, scales, rotations, gb, bb, ib, alpha = ctx.saved_tensors
        vm, pm, cp, tanx, tany = ctx.cam
        e = torch.empty(0, device=means3D.device)
        outs = raster.rasterize_gaussians_backward(
            torch.ones(3, device=means3D.device), means3D, radii, e, scales, rotations, 1.0, e, vm,
            pm, tanx, tany, g_color, g_depth, g_alpha, sh, 3, cp, gb, ctx.K, bb, ib, alpha, False)
        g_m2d, g_col, g_op, g_m3d, g_cov, g_sh, g_sc, g_rot = outs
        return (g_m3d, g_sh, g_op, g_sc, g_rot) + (None,) * 9


def test_dropin_buffers_are_freed(gpu):
    """A Renderer.py-style loop through the drop-in (GaussianTrainer.py:104-114:
    forward, loss, backward, drop everything): no forward state outlives its
    buffers and device memory returns to where it started."""
    arrs = scene_np(2000, 3, 31, 2.0)
    m, s, r, o, sh = to_dev(arrs, gpu)
    params = [t.clone().requires_grad_(True) for t in (m, sh, o.reshape(-1, 1), s, r)]
    cam = orbit_views(128, 128)[0]
    vm, pm, cp = (cam.viewmatrix.to(gpu), cam.projmatrix.to(gpu), cam.campos.to(gpu))
    args = (vm, pm, cp, cam.focal_x, cam.focal_y, cam.tanfovx, cam.tanfovy, 128, 128)

    def one():
        color, radii, depth, alpha = _RendererFunc.apply(*params, *args)
        loss = (color ** 2).mean()
        loss.backward()
        for p in params:
            p.grad = None

    one()  # the first call sizes the capacity policy and torch's pools
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    for i in range(50):
        one()
        if i == 0:
            torch.cuda.synchronize()
            first = torch.cuda.memory_allocated()
    torch.cuda.synchronize()
    print("live states", raster.dropin_live_states(), "allocated", base, first,
          torch.cuda.memory_allocated(), flush=True)
    assert raster.dropin_live_states() == 0
    assert torch.cuda.memory_allocated() == base
    # a forward whose outputs are kept holds exactly one state
    keep = raster.rasterize_gaussians(params[1].detach(), cam.focal_y, params[3].detach(),
                                      params[0].detach(), pm, 3, params[2].detach(),
                                      params[4].detach(), cp, vm, cam.focal_x,
                                      image_width=128, image_height=128)
    assert raster.dropin_live_states() == 1
    del keep
    assert raster.dropin_live_states() == 0


SORT_SIZES = [1000, 5000, 10000, 15000, 20000, 1_000_003]


def _sort_case(n, case, rng):
    """The six key distributions of sorting-optimizations/profile_2_bit_coarsened_vs_cub.cu:89-139."""
    if case == 0:
        return rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    if case == 1:
        return np.arange(n, dtype=np.uint64)
    if case == 2:
        return np.arange(n - 1, -1, -1, dtype=np.uint64)
    if case == 3:
        return rng.integers(0, max(n // 10, 1), n).astype(np.uint64)
    if case == 4:
        k = rng.integers(0, 1000, n).astype(np.uint64)
        k[::100] |= np.uint64(1) << np.uint64(63)
        return k
    return rng.integers(0, 100000, n).astype(np.uint64)


@pytest.mark.parametrize("n", SORT_SIZES)
def test_sort_pairs_six_distributions(gpu, n):
    import ctypes
    from word2world_amd import _lib
    L = _lib.load()
    for case in range(6):
        print(f"sort n={n} case={case}", flush=True)
        rng = np.random.default_rng(42 + case)
        keys = _sort_case(n, case, rng)
        vals = np.arange(n, dtype=np.uint32)
        kin = torch.from_numpy(keys.view(np.int64)).to(gpu)
        vin = torch.from_numpy(vals.view(np.int32)).to(gpu)
        kout = torch.empty_like(kin)
        vout = torch.empty_like(vin)
        tb = L.w2w_sort_temp_bytes(n, 64)
        temp = torch.empty(tb, dtype=torch.uint8, device=gpu)
        st = L.w2w_sort_pairs_u64(kin.data_ptr(), vin.data_ptr(), kout.data_ptr(), vout.data_ptr(),
                                  n, 64, temp.data_ptr(), tb, torch.cuda.current_stream().cuda_stream)
        assert st == 0
        torch.cuda.synchronize()
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(kout.cpu().numpy().view(np.uint64), keys[order])
        np.testing.assert_array_equal(vout.cpu().numpy().view(np.uint32), vals[order])


@pytest.mark.parametrize("n", [1, 1000, 4096, 4097, 15000, 16384, 16385, 1_000_001, 1_150_001])
def test_exclusive_scan(gpu, n):
    from word2world_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1729, n).astype(np.uint32)  # profile_ultra_optimized_vs_cub.cu:95-124 values
    xin = torch.from_numpy(x.view(np.int32)).to(gpu)
    out = torch.empty_like(xin)
    tot = torch.zeros(1, dtype=torch.int32, device=gpu)
    tb = L.w2w_scan_temp_bytes(n)
    temp = torch.empty(max(tb, 1), dtype=torch.uint8, device=gpu)
    st = L.w2w_exclusive_scan_u32(xin.data_ptr(), out.data_ptr(), tot.data_ptr(), n, temp.data_ptr(),
                                  tb, torch.cuda.current_stream().cuda_stream)
    assert st == 0
    torch.cuda.synchronize()
    c = np.cumsum(x.astype(np.uint64)).astype(np.uint32)
    excl = np.concatenate([[0], c[:-1]]).astype(np.uint32)
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), excl)
    assert int(tot.cpu().numpy().view(np.uint32)[0]) == int(c[-1])


def test_exclusive_scan_saturates(gpu):
    """A total of 2^30 or more (the look-back words hold 30 bits) reads as
    2^30 - 1: larger than every workspace capacity, so the instance count
    of an oversized batch is flagged as an overflow instead of wrapping."""
    from word2world_amd import _lib
    L = _lib.load()
    n = 300_000
    x = np.full(n, 4000, dtype=np.uint32)  # total 1.2e9 > 2^30
    xin = torch.from_numpy(x.view(np.int32)).to(gpu)
    out = torch.empty_like(xin)
    tot = torch.zeros(1, dtype=torch.int32, device=gpu)
    tb = L.w2w_scan_temp_bytes(n)
    temp = torch.empty(tb, dtype=torch.uint8, device=gpu)
    assert L.w2w_exclusive_scan_u32(xin.data_ptr(), out.data_ptr(), tot.data_ptr(), n, temp.data_ptr(), tb,
                                    torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    assert int(tot.cpu().numpy().view(np.uint32)[0]) == (1 << 30) - 1
    # below the limit the offsets are exact
    o = out.cpu().numpy().view(np.uint32)
    k = (1 << 30) // 4000 - 16384
    np.testing.assert_array_equal(o[:k], 4000 * np.arange(k, dtype=np.uint64))
