"""Step-level harness (row A17): fused Adam against torch.optim.Adam (the
library oracle for the trainer's optimiser), the flat parameter store, the
mean learning-rate schedule (densification: test_densify.py)."""
import math

import numpy as np
import pytest
import torch

from word2world_amd.scene import make_scene
from word2world_amd.train import GROUPS, LRS, GaussianStore, mean_lr


def test_mean_lr_schedule_closed_form():
    # misc_utils.py:11-27: exp(log(a)(1-t) + log(b) t), t = clip(step/300)
    assert mean_lr(0) == pytest.approx(0.01)
    assert mean_lr(300) == pytest.approx(0.0002)
    assert mean_lr(10_000) == pytest.approx(0.0002)
    assert mean_lr(150) == pytest.approx(math.sqrt(0.01 * 0.0002))


def test_store_views_roundtrip():
    sc = make_scene(50, 3, 0)
    st = GaussianStore(sc.params(), 3, "cpu")
    v = st.views()
    for k in GROUPS:
        np.testing.assert_array_equal(v[k].detach().numpy(), sc.params()[k].numpy())
    assert st.group_ends()[-1] == st.flat.numel() == 50 * (3 + 3 + 45 + 1 + 3 + 4)


@pytest.mark.gpu
def test_fused_adam_matches_torch(gpu):
    torch.manual_seed(0)
    sc = make_scene(1000, 3, 1)
    st = GaussianStore(sc.params(), 3, gpu)
    ref = {k: v.detach().clone().to(gpu).requires_grad_(True) for k, v in sc.params().items()}
    opt = torch.optim.Adam([{"params": [ref[k]], "lr": LRS[k], "name": k} for k in GROUPS],
                           lr=0.0, eps=1e-15)
    for it in range(1, 6):
        grads = {k: torch.randn_like(ref[k]) * 10 ** (-it) for k in GROUPS}
        lr_means = mean_lr(it)
        for gr in opt.param_groups:
            if gr["name"] == "means":
                gr["lr"] = lr_means
        for k in GROUPS:
            ref[k].grad = grads[k].clone()
        opt.step()
        st.flat.grad = torch.cat([grads[k].reshape(-1) for k in GROUPS])
        st.adam_step({"means": lr_means})
        torch.cuda.synchronize()
        v = st.views()
        for k in GROUPS:
            np.testing.assert_allclose(v[k].detach().cpu().numpy(), ref[k].detach().cpu().numpy(),
                                       rtol=1e-5, atol=1e-7, err_msg=k)


@pytest.mark.gpu
def test_plain_trainer_reduces_loss(gpu):
    from helpers import gpu_camera
    from word2world_amd import raster
    from word2world_amd.camera import orbit_views
    from word2world_amd.scene import activate
    from word2world_amd.train import PlainTrainer, TrainConfig

    cams_h = orbit_views(128, 128)
    cams = [gpu_camera(c, gpu) for c in cams_h]
    tgt_scene = make_scene(3000, 3, 1)
    tm, ts, tr, to, tsh = (t.to(gpu).contiguous() for t in activate(tgt_scene.params()))
    targets = []
    for c in cams:
        out = raster.forward_checked(tm, ts, tr, to, tsh, 3, c, 16)
        targets.append(out[1].clamp(0, 1))
    tr_ = PlainTrainer(make_scene(3000, 3, 0), cams, targets, TrainConfig(densify_every=5), gpu)
    losses = [float(tr_.step()) for _ in range(20)]
    assert not tr_.overflow()
    assert losses[-1] < losses[0]
    assert tr_.store.n > 0


@pytest.mark.gpu
def test_autotuner_picks_a_tile_and_tiles_agree(gpu):
    """The sweep times every (tile edge, chunk); the images at different
    edges agree (the per-pixel contributor sequence only changes where the
    reference's non-conservative rect rule drops a tile, a handful of
    pixels), and at one edge every chunk gives the same image, depth, alpha,
    contributor counts and final T bit for bit."""
    from word2world_amd import autotune, raster
    from word2world_amd.camera import orbit_camera
    from word2world_amd.scene import activate, make_scene
    sc = make_scene(20000, 3, 5)
    m, s, r, o, sh = (t.to(gpu).contiguous() for t in activate(sc.params()))
    c = orbit_camera(30.0, 10.0, 2.5, 256, 192)
    cam = raster.camera_from_tensors(c.viewmatrix, c.projmatrix, c.campos, 256, 192, c.focal_x, c.focal_y,
                                     None, device=gpu)
    tuner = autotune.Autotuner(reps=1)
    best = tuner.best(m, s, r, o, sh, 3, cam)
    t = tuner.timings[tuner.key(20000, 256, 192)]
    assert best in autotune.candidates() and set(t) == set(autotune.candidates())
    assert all(v > 0 for v in t.values()) and tuner.best_tile(m, s, r, o, sh, 3, cam) == best[0]
    imgs = {ts: raster.forward_checked(m, s, r, o, sh, 3, cam, ts)[1] for ts in (8, 16, 32)}
    for ts in (8, 32):
        d = (imgs[ts] - imgs[16]).abs()
        assert float((d > 1e-5).float().mean()) < 0.01 and float(d.max()) < 0.5
    for ts in (8, 16, 32):
        ref = None
        for ch in (64, 128, 256):
            K, color, depth, alpha, radii, ws = raster.forward_checked(m, s, r, o, sh, 3, cam, ts, chunk=ch)
            got = (color, depth, alpha, ws.n_contrib().clone(), ws.final_T().clone())
            if ref is None:
                ref = got
            else:
                assert all(torch.equal(a, b) for a, b in zip(ref, got)), (ts, ch)


@pytest.mark.gpu
@pytest.mark.parametrize("offset", [0, 1])
def test_adam_hyper_matches_per_element_reference(gpu, offset):
    """w2w_adam_step_hyper (device scalars, a second rate along each SH row)
    against torch's Adam update written out per group, bit for bit: the
    16-byte path (offset 0) and the scalar path (arrays offset by one float);
    n not a multiple of 4, group ends not multiples of 4, an empty group."""
    import ctypes
    from word2world_amd import _lib
    L = _lib.load()
    torch.manual_seed(0)
    G, M = 1001, 16
    sizes = [3 * G, 3 * M * G, 0, G, 3 * G, 4 * G]
    ends = np.cumsum(sizes).astype(np.int64)
    n = int(ends[-1])
    mk = lambda: torch.randn(n + 4, device=gpu)
    p, g, m, v = mk(), mk(), mk() * 0.1, mk().abs() * 0.01
    period = np.array([0, 3 * M, 0, 0, 0, 0], np.int32)
    split = np.array([0, 3, 0, 0, 0, 0], np.int32)
    step, b1, b2, eps = 7, 0.9, 0.999, 1e-15
    lrs = [0.01, 0.01, 0.3, 0.05, 0.005, 0.005]
    h = np.zeros(32, np.float32)
    for k, lr in enumerate(lrs):
        h[k] = -lr / (1 - b1 ** step)
    h[8 + 1] = -0.0005 / (1 - b1 ** step)
    h[16], h[17], h[18], h[19], h[20] = 1 / np.sqrt(1 - b2 ** step), 1 - b1, b2, 1 - b2, eps
    hyper = torch.from_numpy(h).to(gpu)
    sl = lambda t: t[offset:offset + n]
    # reference: the same float32 expressions, per group, in torch
    ns = torch.empty(n, device=gpu)
    start = 0
    for k, e in enumerate(ends):
        ns[start:e] = float(h[k])
        if period[k]:
            idx = torch.arange(int(e) - start, device=gpu)
            ns[start:e] = torch.where(idx % int(period[k]) >= int(split[k]), torch.tensor(float(h[8 + k]), device=gpu),
                                      ns[start:e])
        start = int(e)
    P, Gr, Mo, Vo = (sl(t).clone() for t in (p, g, m, v))
    Mo = Mo + torch.tensor(np.float32(1 - b1), device=gpu) * (Gr - Mo)
    Vo = Vo * torch.tensor(np.float32(b2), device=gpu) + (torch.tensor(np.float32(1 - b2), device=gpu) * Gr) * Gr
    den = torch.sqrt(Vo) * hyper[16] + hyper[20]
    P = P + ns * (Mo / den)
    st = L.w2w_adam_step_hyper(sl(p).data_ptr(), sl(g).data_ptr(), sl(m).data_ptr(), sl(v).data_ptr(), n,
                               len(sizes), ends.ctypes.data, period.ctypes.data, split.ctypes.data,
                               hyper.data_ptr(), None, None, torch.cuda.current_stream().cuda_stream)
    assert st == 0
    torch.cuda.synchronize()
    assert torch.equal(sl(m), Mo) and torch.equal(sl(v), Vo)
    assert torch.equal(sl(p), P)
    # gated: an overflowed step updates nothing and leaves its mark
    gate = torch.tensor([0, 1], dtype=torch.int32, device=gpu)
    sticky = torch.zeros(1, dtype=torch.int32, device=gpu)
    before = [sl(t).clone() for t in (p, m, v)]
    st = L.w2w_adam_step_hyper(sl(p).data_ptr(), sl(g).data_ptr(), sl(m).data_ptr(), sl(v).data_ptr(), n,
                               len(sizes), ends.ctypes.data, period.ctypes.data, split.ctypes.data,
                               hyper.data_ptr(), gate[1:].data_ptr(), sticky.data_ptr(),
                               torch.cuda.current_stream().cuda_stream)
    assert st == 0
    torch.cuda.synchronize()
    assert all(torch.equal(sl(t), b) for t, b in zip((p, m, v), before))
    assert int(sticky[0]) == 1
    # an open gate updates as before and leaves the mark alone
    sticky.zero_()
    st = L.w2w_adam_step_hyper(sl(p).data_ptr(), sl(g).data_ptr(), sl(m).data_ptr(), sl(v).data_ptr(), n,
                               len(sizes), ends.ctypes.data, period.ctypes.data, split.ctypes.data,
                               hyper.data_ptr(), gate[:1].data_ptr(), sticky.data_ptr(),
                               torch.cuda.current_stream().cuda_stream)
    assert st == 0
    torch.cuda.synchronize()
    assert not torch.equal(sl(p), before[0]) and int(sticky[0]) == 0


@pytest.mark.gpu
def test_adam_ranges_matches_per_element_reference(gpu):
    """w2w_adam_ranges (per-job row ranges of a resident store, each with its
    own hyper block and SH period) against the same update written out per
    range, bit for bit; ranges out of store order, one empty."""
    from word2world_amd import _lib
    from word2world_amd.serve import PARAM_RANGE
    L = _lib.load()
    torch.manual_seed(1)
    N = 50_000
    p, g = torch.randn(N, device=gpu), torch.randn(N, device=gpu)
    m, v = torch.randn(N, device=gpu) * 0.1, torch.randn(N, device=gpu).abs() * 0.01
    spec = [(30_000, 4_321, 0, 48, 3), (100, 0, 1, 0, 0), (5, 17_000, 1, 0, 0), (40_000, 9_999, 2, 9, 3)]
    rng = np.zeros(len(spec), PARAM_RANGE)
    off = 0
    for k, (start, ln, hy, per, spl) in enumerate(spec):
        rng[k] = (start, off, ln, hy, per, spl, 0)
        off += ln
    total = off
    h = np.zeros(12, np.float32)
    for k in range(3):
        h[4 * k:4 * k + 3] = (-0.01 * (k + 1), -0.0005 * (k + 1), 1.0 / np.sqrt(1 - 0.999 ** (k + 2)))
    hyper = torch.from_numpy(h).to(gpu)
    ranges = torch.from_numpy(rng.view(np.uint8)).to(gpu)
    b1, b2, eps = 0.9, 0.999, 1e-15
    # the entry point takes float beta1/beta2 and forms 1 - beta from them
    c1 = torch.tensor(np.float32(1.0 - float(np.float32(b1))), device=gpu)
    c2 = torch.tensor(np.float32(1.0 - float(np.float32(b2))), device=gpu)
    P, Mo, Vo = p.clone(), m.clone(), v.clone()
    for (start, ln, hy, per, spl) in spec:
        if ln == 0:
            continue
        sl = slice(start, start + ln)
        ns = torch.full((ln,), float(h[4 * hy]), device=gpu)
        if per:
            e = torch.arange(ln, device=gpu)
            ns = torch.where(e % per >= spl, torch.tensor(float(h[4 * hy + 1]), device=gpu), ns)
        gi = g[sl]
        Mo[sl] = Mo[sl] + c1 * (gi - Mo[sl])
        Vo[sl] = Vo[sl] * torch.tensor(np.float32(b2), device=gpu) + (c2 * gi) * gi
        den = torch.sqrt(Vo[sl]) * hyper[4 * hy + 2] + torch.tensor(np.float32(eps), device=gpu)
        P[sl] = P[sl] + ns * (Mo[sl] / den)
    st = L.w2w_adam_ranges(p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), ranges.data_ptr(),
                           len(spec), total, hyper.data_ptr(), b1, b2, eps, torch.cuda.current_stream().cuda_stream)
    assert st == 0
    torch.cuda.synchronize()
    assert torch.equal(m, Mo) and torch.equal(v, Vo) and torch.equal(p, P)
