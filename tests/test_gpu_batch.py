"""Batched form: one launch per stage over several scenes and views must give
exactly the per-frame results of the single-view path (same kernels, frame
table indirection), and a scene must render identically alone and in a mix."""
import numpy as np
import pytest
import torch

from helpers import gpu_camera, scene_np
from word2world_amd import batch, raster
from word2world_amd.camera import orbit_camera

pytestmark = pytest.mark.gpu


def _setup(dev):
    specs = [(800, 0, 11, [(96, 80, 10.0), ]),
             (1500, 3, 12, [(64, 64, 40.0), (100, 120, 200.0)]),
             (400, 1, 13, [(48, 48, 0.0), (130, 97, 95.0), (64, 64, 300.0)])]
    scenes, arrays, cams = [], [], []
    for n, deg, seed, views in specs:
        arrs = scene_np(n, deg, seed, 2.0)
        arrays.append(arrs)
        fr, cs = [], []
        for (w, h, az) in views:
            cam = orbit_camera(az, 10.0, 2.5, w, h)
            cs.append(cam)
            fr.append(batch.frame_from_camera(cam))
        cams.append(cs)
        scenes.append(batch.SceneSpec(n, deg, arrs[4].shape[1], fr))
    cat = lambda i: torch.from_numpy(np.concatenate([a[i].reshape(a[i].shape[0], -1) for a in arrays])).to(dev)
    means, scales, rots = cat(0), cat(1), cat(2)
    opac = torch.from_numpy(np.concatenate([a[3].reshape(-1) for a in arrays])).to(dev)
    sh = torch.from_numpy(np.concatenate([a[4].reshape(-1) for a in arrays])).to(dev)
    return scenes, arrays, cams, (means.contiguous(), scales.contiguous(), rots.contiguous(), opac, sh)


def _run_batch(scenes, packed, dev, capacity=200000):
    caps = batch.caps_for(scenes, 16, capacity)
    ws = batch.BatchWorkspace(caps, dev)
    ws.load_table(scenes)
    batch.forward(ws, *packed)
    torch.cuda.synchronize()
    return ws


def test_batch_forward_equals_single_view(gpu):
    scenes, arrays, cams, packed = _setup(gpu)
    ws = _run_batch(scenes, packed, gpu)
    K_sum = 0
    views = ws.frame_views(scenes)
    fi = 0
    radii = ws.radii.cpu().numpy()
    item = 0
    for s, arrs, cs in zip(scenes, arrays, cams):
        d = [torch.from_numpy(np.ascontiguousarray(a)).to(gpu) for a in arrs]
        per_view_radii = []
        for cam in cs:
            print(f"frame {fi}: {cam.width}x{cam.height}", flush=True)
            K, color, depth, alpha, r, wsp = raster.forward_checked(*d, s.sh_degree, gpu_camera(cam, gpu), 16)
            K_sum += K
            bc, bd, ba = views[fi]
            assert torch.equal(bc, color) and torch.equal(bd, depth) and torch.equal(ba, alpha)
            per_view_radii.append(r.cpu().numpy())
            fi += 1
        V = len(cs)
        got = radii[item:item + s.n * V].reshape(s.n, V)
        for v in range(V):
            np.testing.assert_array_equal(got[:, v], per_view_radii[v])
        item += s.n * V
    assert int(ws.status()[0]) == K_sum


def test_batch_backward_equals_sum_of_views(gpu):
    scenes, arrays, cams, packed = _setup(gpu)
    ws = _run_batch(scenes, packed, gpu)
    torch.manual_seed(0)
    dLdc = torch.randn_like(ws.color)
    sh_floats = int(ws.totals[5])
    grads = batch.BatchGrads(ws.caps, sh_floats, gpu)
    batch.backward(ws, grads, packed[0], packed[1], packed[2], packed[4], dLdc)
    torch.cuda.synchronize()
    g_off, sh_off, pix, item = 0, 0, 0, 0
    for s, arrs, cs in zip(scenes, arrays, cams):
        d = [torch.from_numpy(np.ascontiguousarray(a)).to(gpu) for a in arrs]
        acc = None
        for v, cam in enumerate(cs):
            gcam = gpu_camera(cam, gpu)
            K, color, depth, alpha, r, wsp = raster.forward_checked(*d, s.sh_degree, gcam, 16)
            hw = cam.width * cam.height
            gC = dLdc[3 * pix:3 * (pix + hw)].view(3, cam.height, cam.width).contiguous()
            z = torch.zeros(cam.height, cam.width, device=gpu)
            out = raster.backward(*d, s.sh_degree, gcam, 16, wsp, gC, z, z)
            # per-item means2D of this view
            m2 = grads.means2D[item:item + s.n * len(cs)].view(s.n, len(cs), 3)[:, v]
            np.testing.assert_allclose(m2.cpu().numpy(), out[0].cpu().numpy(), rtol=1e-5, atol=1e-9)
            vals = [out[2].view(-1), out[3], out[4], out[5].reshape(-1), out[6], out[7]]
            acc = vals if acc is None else [a + b for a, b in zip(acc, vals)]
            pix += hw
        n, ns = s.n, s.n * s.n_coeffs * 3
        got = [grads.opacity[g_off:g_off + n], grads.means3D[g_off:g_off + n],
               grads.cov3D[g_off:g_off + n], grads.sh[sh_off:sh_off + ns],
               grads.scales[g_off:g_off + n], grads.rotations[g_off:g_off + n]]
        for name, a, b in zip(["opacity", "means3D", "cov3D", "sh", "scales", "rotations"], got, acc):
            a, b = a.cpu().numpy(), b.cpu().numpy()
            err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
            print(name, err, flush=True)
            assert err < 1e-5, (name, err)
        g_off += n
        sh_off += ns
        item += s.n * len(cs)


def test_scene_alone_and_in_the_mix(gpu):
    scenes, arrays, cams, packed = _setup(gpu)
    ws = _run_batch(scenes, packed, gpu)
    mix = [tuple(t.clone() for t in v) for v in ws.frame_views(scenes)]
    # scene 1 alone
    s1 = [scenes[1]]
    n0 = scenes[0].n
    n1 = scenes[1].n
    sh0 = scenes[0].n * scenes[0].n_coeffs * 3
    sh1 = scenes[1].n * scenes[1].n_coeffs * 3
    p1 = (packed[0][n0:n0 + n1].contiguous(), packed[1][n0:n0 + n1].contiguous(),
          packed[2][n0:n0 + n1].contiguous(), packed[3][n0:n0 + n1].contiguous(),
          packed[4][sh0:sh0 + sh1].contiguous())
    ws1 = _run_batch(s1, p1, gpu)
    alone = ws1.frame_views(s1)
    for (c, d, a), (mc, md, ma) in zip(alone, mix[1:3]):
        assert torch.equal(c, mc) and torch.equal(d, md) and torch.equal(a, ma)


def test_capacity_overflow_is_reported(gpu):
    scenes, arrays, cams, packed = _setup(gpu)
    ws = _run_batch(scenes, packed, gpu, capacity=1000)
    K, ovf = (int(v) for v in ws.status().cpu())
    assert ovf == 1 and K > 1000


def test_tight_emission_same_images_and_gradients(gpu):
    """Emitting only into the tiles a Gaussian's cull extent reaches drops
    instances no pixel can use: images, alpha, depth, final T and every
    gradient are bit-identical to the reference emission, with fewer
    instances."""
    import copy
    scenes, arrays, cams, packed = _setup(gpu)
    tight = copy.deepcopy(scenes)
    for s in tight:
        for f in s.frames:
            f.tight = True
    out = []
    for sc in (scenes, tight):
        ws = _run_batch(sc, packed, gpu)
        torch.manual_seed(0)
        dLdc = torch.randn_like(ws.color)
        grads = batch.BatchGrads(ws.caps, int(ws.totals[5]), gpu)
        batch.backward(ws, grads, packed[0], packed[1], packed[2], packed[4], dLdc)
        torch.cuda.synchronize()
        out.append((int(ws.status()[0]), [tuple(t.clone() for t in v) for v in ws.frame_views(sc)],
                    ws.final_T().clone(), grads))
    (K0, v0, T0, g0), (K1, v1, T1, g1) = out
    print("K reference", K0, "tight", K1, flush=True)
    assert K1 < K0
    for a, b in zip(v0, v1):
        assert all(torch.equal(x, y) for x, y in zip(a, b))
    assert torch.equal(T0, T1)
    for name in ("means2D", "colors", "opacity", "means3D", "cov3D", "sh", "scales", "rotations"):
        assert torch.equal(getattr(g0, name), getattr(g1, name)), name


def _dense_setup(dev, n=60000, res=96):
    """A scene dense enough that most tiles saturate early (the case the
    depth-sliced binning is for): the benchmark scene's distribution."""
    arrs = scene_np(n, 1, 5, 1.5)
    fr = [batch.frame_from_camera(orbit_camera(az, el, 2.5, res, res + 8), tight=True)
          for az, el in ((0.0, 0.0), (70.0, 20.0), (200.0, -25.0))]
    scenes = [batch.SceneSpec(n, 1, arrs[4].shape[1], fr)]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    packed = (t(arrs[0]), t(arrs[1]), t(arrs[2]), t(arrs[3].reshape(-1)), t(arrs[4].reshape(-1)))
    return scenes, packed


def _render_and_grads(scenes, packed, dev, tile, split, capacity=3_000_000, forwards=1, chunk=0):
    caps = batch.caps_for(scenes, tile, capacity, depth_split=split, chunk=chunk)
    ws = batch.BatchWorkspace(caps, dev)
    ws.load_table(scenes)
    for _ in range(forwards):  # later forwards use the earlier ones' tile hint
        batch.forward(ws, *packed)
    g = torch.Generator(device=dev).manual_seed(0)
    dLdc = torch.randn(ws.color.shape, generator=g, device=dev)
    dLdd = torch.randn(ws.depth.shape, generator=g, device=dev)
    dLda = torch.randn(ws.alpha.shape, generator=g, device=dev)
    grads = batch.BatchGrads(ws.caps, int(ws.totals[5]), dev)
    batch.backward(ws, grads, packed[0], packed[1], packed[2], packed[4], dLdc, dLdd, dLda)
    torch.cuda.synchronize()
    assert int(ws.status()[1]) == 0
    counts = [int(v) for v in ws.sliced_counts().cpu()]
    return ws, grads, counts


@pytest.mark.parametrize("dense", [False, True])
@pytest.mark.parametrize("tile,split,forwards", [(16, 0.3, 1), (16, 0.3, 2), (8, 0.3, 2), (16, 0.05, 1),
                                                 (8, 0.7, 2), (32, 0.5, 2), (16, 0.001, 2), (8, 0.999, 1)])
def test_depth_sliced_binning_is_exact(gpu, dense, tile, split, forwards):
    """Near lists first (the near slice, plus the far items on the tiles the
    previous forward left unfinished), far lists only on the other tiles
    with unsaturated pixels: every image, alpha, depth, final T, n_contrib
    and gradient (with depth and alpha gradients given) is bit-identical to
    one binning pass, with or without a hint.  Tile edges 8 and 16 here need
    two and one tile-sort passes (both buffer parities of the far sort)."""
    print(f"dense={dense} tile={tile} split={split} forwards={forwards}", flush=True)
    scenes, packed = _dense_setup(gpu) if dense else _setup(gpu)[0::3]
    ref_ws, ref_g, _ = _render_and_grads(scenes, packed, gpu, tile, 0.0)
    ws, g, (S, KA, KB, n_open, KH, KN) = _render_and_grads(scenes, packed, gpu, tile, split, forwards=forwards)
    K = int(ref_ws.status()[0])
    assert int(ws.status()[0]) == K
    P = int(ws.totals[1])
    assert S == min(P, -(-P * int(round(1000 * split)) // 1000))
    assert KN == KA + KH and 0 <= KA <= K and KN + KB <= K
    if forwards == 1:
        assert KH == 0  # no hint before the first forward
    print(f"K={K} near={KA} hinted={KH} far={KB} open blocks={n_open} (near items {S} of {P})", flush=True)
    for a, b in zip(ref_ws.frame_views(scenes), ws.frame_views(scenes)):
        assert all(torch.equal(x, y) for x, y in zip(a, b))
    assert torch.equal(ref_ws.final_T(), ws.final_T())
    n = int(ws.totals[4])
    o = ws.layout.n_contrib
    assert torch.equal(ref_ws.image[o:o + 4 * n], ws.image[o:o + 4 * n])
    for name in ("means2D", "colors", "opacity", "means3D", "cov3D", "sh", "scales", "rotations"):
        assert torch.equal(getattr(ref_g, name), getattr(g, name)), name
    if dense and split == 0.3 and forwards == 1:
        assert KN + KB < K  # tiles saturated by the near lists drop their far instances


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_blend_chunk_is_exact(gpu, tile):
    """The forward blend's chunk (64, 128 or 256 instances per step of a
    wave's walk) changes what is in flight, not the result: images, alpha,
    depth, final T, contributor counts and every gradient are bit-identical,
    in one binning pass and depth-sliced (near, hinted and far lists)."""
    scenes, packed = _dense_setup(gpu)
    for split in (0.0, 0.3):
        ref_ws, ref_g, _ = _render_and_grads(scenes, packed, gpu, tile, split, forwards=2, chunk=64)
        n = int(ref_ws.totals[4])
        o = ref_ws.layout.n_contrib
        for ch in (128, 256):
            print(f"tile={tile} split={split} chunk={ch}", flush=True)
            ws, g, _ = _render_and_grads(scenes, packed, gpu, tile, split, forwards=2, chunk=ch)
            for a, b in zip(ref_ws.frame_views(scenes), ws.frame_views(scenes)):
                assert all(torch.equal(x, y) for x, y in zip(a, b))
            assert torch.equal(ref_ws.final_T(), ws.final_T())
            assert torch.equal(ref_ws.image[o:o + 4 * n], ws.image[o:o + 4 * n])
            for name in ("means2D", "colors", "opacity", "means3D", "cov3D", "sh", "scales", "rotations"):
                assert torch.equal(getattr(ref_g, name), getattr(g, name)), (ch, name)
            del ws, g
