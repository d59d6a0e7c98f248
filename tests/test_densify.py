"""Adaptive density control (clone / split / prune with the Adam-state remap,
Densification.py:33-187, OptimizerManager.py:23-45, gaussians_handler.py:185-194).

CPU tests pin the oracle (oracle/densify.py) on small cases whose results are
written out by hand: which rows survive and in what order, the zero moments
of new rows, and the children's means and log-scales in closed form.  GPU
tests run the HIP kernels (w2w_densify_plan / _apply through the store)
against the oracle: bit-exact for every copied value, moment and count;
children's means and log-scales within 2e-6 (expf/logf of the device
library against numpy's).  The decisions compare exp(svec) and
sigmoid(logit) with thresholds, where the device library's and numpy's
rounding may differ in the last place: random inputs are kept 1e-4 (in
log-space) clear of every threshold."""
import math

import numpy as np
import pytest
import torch

from oracle.densify import GROUPS, adc_cycle
from word2world_amd.scene import make_scene
from word2world_amd.train import GaussianStore

ls = math.log


def _params(n, K, svec, opacity_logit, means=None, quat=None, seed=5):
    sc = make_scene(n, 3, seed, n_coeffs=K + 1)
    p = {k: v.numpy().copy() for k, v in sc.params().items()}
    p["svec"] = np.asarray(svec, np.float32).reshape(n, 3)
    p["opacity"] = np.asarray(opacity_logit, np.float32).reshape(n, 1)
    if means is not None:
        p["means"] = np.asarray(means, np.float32).reshape(n, 3)
    if quat is not None:
        p["quaternion"] = np.asarray(quat, np.float32).reshape(n, 4)
    return p


def _moments(p, base):
    """Row r of every group holds base + r."""
    n = p["means"].shape[0]
    return {k: np.broadcast_to((base + np.arange(n, dtype=np.float32)).reshape((n,) + (1,) * (v.ndim - 1)),
                               v.shape).copy() for k, v in p.items()}


SIX_SVEC = [[ls(0.01)] * 3,               # 0: small, high grad -> clone
            [ls(5.0), ls(0.1), ls(0.1)],  # 1: large (> 4), high grad -> split; children 3.125 > 0.4 -> pruned
            [ls(0.01)] * 3,               # 2: small, low grad -> kept
            [ls(0.01)] * 3,               # 3: low opacity -> pruned
            [ls(0.2)] * 3,                # 4: medium, high grad -> neither
            [ls(0.5)] * 3]                # 5: max scale 0.5 > 0.4 -> pruned
SIX_OPAC = [2.0, 2.0, 2.0, -8.0, 2.0, 2.0]
SIX_ACCUM = [1.0, 1.0, 1e-6, 0.0, 1.0, 0.0]


def test_oracle_clone_split_prune_and_remap():
    p = _params(6, 15, SIX_SVEC, SIX_OPAC)
    P, M, V, cnt = adc_cycle(p, _moments(p, 1.0), _moments(p, 2.0), SIX_ACCUM, [1.0] * 6,
                             np.zeros((2, 3), np.float32), 2e-4, 0.005)
    # rows after clone: 0..5, clone(0); split removes 1 and appends two children;
    # prune removes 3, 5 and both children
    assert cnt == (1, 1, 4)
    for k in GROUPS:
        np.testing.assert_array_equal(P[k], p[k][[0, 2, 4, 0]])
    np.testing.assert_array_equal(M["means"][:, 0], [1.0, 3.0, 5.0, 0.0])
    np.testing.assert_array_equal(V["svec"][:, 2], [2.0, 4.0, 6.0, 0.0])


def test_oracle_children_closed_form():
    """Two split rows A (q = 90 deg about z, unnormalised) and B (q = 3 * identity)
    around a kept row C; children survive (prune scale 10).  Child k of the
    s-th split row uses sample row k * S + s; output [C, A0, B0, A1, B1]."""
    c45 = math.cos(math.pi / 4)
    svec = [[ls(1.0), ls(0.5), ls(2.0)], [ls(0.2)] * 3, [ls(2.0), ls(2.0), ls(0.5)]]
    means = [[1, 2, 3], [0, 0, 0], [-1, -2, -3]]
    quat = [[2 * c45, 0, 0, 2 * c45], [1, 0, 0, 0], [3, 0, 0, 0]]
    p = _params(3, 3, svec, [1.0, 1.0, 1.0], means, quat)
    z = np.array([[0.3, -0.4, 0.5], [1.0, 2.0, -1.0], [-0.2, 0.1, 0.0], [0.5, 0.5, 0.5]], np.float32)
    P, M, V, cnt = adc_cycle(p, _moments(p, 1.0), _moments(p, 2.0), [1.0, 0.0, 1.0], [1.0] * 3, z,
                             2e-4, 0.005, clone_scale_max=0.01, split_scale_min=1.5,
                             prune_scale_max=10.0)
    assert cnt == (0, 2, 0)
    # A: std (1, .5, 2); s = std * z; R_z(90) (x, y, z) = (-y, x, z)
    want_means = [[0, 0, 0],
                  [1 + 0.2, 2 + 0.3, 3 + 1.0],           # A, z0: s = (.3, -.2, 1)
                  [-1 + 2.0, -2 + 4.0, -3 - 0.5],        # B, z1: R = I, s = (2, 4, -.5)
                  [1 - 0.05, 2 - 0.2, 3 + 0.0],          # A, z2: s = (-.2, .05, 0)
                  [-1 + 1.0, -2 + 1.0, -3 + 0.25]]       # B, z3: s = (1, 1, .25)
    np.testing.assert_allclose(P["means"], np.array(want_means, np.float32), rtol=0, atol=2e-6)
    la, lb = [math.log(x / 1.6) for x in (1.0, 0.5, 2.0)], [math.log(x / 1.6) for x in (2.0, 2.0, 0.5)]
    np.testing.assert_allclose(P["svec"], np.array([svec[1], la, lb, la, lb], np.float32), atol=1e-6)
    np.testing.assert_array_equal(P["quaternion"], np.array(quat, np.float32)[[1, 0, 2, 0, 2]])
    np.testing.assert_array_equal(P["sh_ac"], p["sh_ac"][[1, 0, 2, 0, 2]])
    np.testing.assert_array_equal(M["opacity"][:, 0], [2.0, 0, 0, 0, 0])
    np.testing.assert_array_equal(V["quaternion"][:, 3], [3.0, 0, 0, 0, 0])


def test_oracle_clone_is_split_when_threshold_nonpositive():
    """thr = 0: a fresh clone (zero statistics) passes the gradient test too,
    so a row that is both small and large in some axis is cloned and both it
    and its clone are split: children [orig0, clone0, orig1, clone1]."""
    svec = [[ls(0.01), ls(5.0), ls(1.0)]]
    p = _params(1, 0, svec, [1.0], [[0, 0, 0]], [[1, 0, 0, 0]])
    z = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], np.float32)
    P, M, V, cnt = adc_cycle(p, _moments(p, 1.0), _moments(p, 2.0), [0.0], [0.0], z, 0.0, 0.005,
                             prune_scale_max=10.0)
    assert cnt == (1, 2, 0)
    std = np.array([0.01, 5.0, 1.0], np.float32)
    np.testing.assert_allclose(P["means"], std * z, rtol=1e-6)
    assert not M["means"].any() and not V["svec"].any()


def test_oracle_empty_and_all_pruned():
    p = _params(2, 3, [[ls(0.01)] * 3] * 2, [-9.0, -9.0])
    P, M, V, cnt = adc_cycle(p, _moments(p, 1.0), _moments(p, 2.0), [1.0, 1.0], [1.0, 1.0],
                             np.zeros((0, 3), np.float32), 2e-4, 0.005)
    assert cnt == (2, 0, 4) and P["means"].shape == (0, 3) and P["sh_ac"].shape == (0, 3, 3)
    P2, _, _, cnt2 = adc_cycle(P, M, V, [], [], np.zeros((0, 3), np.float32), 2e-4, 0.005)
    assert cnt2 == (0, 0, 0) and P2["svec"].shape == (0, 3)


def test_store_densify_refuses_cpu():
    sc = make_scene(4, 1, 0)
    st = GaussianStore(sc.params(), 1, "cpu")
    with pytest.raises(RuntimeError, match="device"):
        st.densify(2e-4, 0.005)


# ---------------------------------------------------------------- GPU

def _store(p, accum, count, dev, mom=True):
    st = GaussianStore({k: torch.from_numpy(np.ascontiguousarray(v)) for k, v in p.items()},
                       3, dev)
    if mom:
        for t, src in ((st.exp_avg, _moments(p, 1.0)), (st.exp_avg_sq, _moments(p, 2.0))):
            vv = st.views(t)
            for k in GROUPS:
                vv[k].copy_(torch.from_numpy(src[k]))
    st.grad_accum = torch.tensor(np.asarray(accum, np.float32), device=dev)
    st.grad_count = torch.tensor(np.asarray(count, np.float32), device=dev)
    return st


def _check_against_oracle(st, p, accum, count, z, kw, exact_rows=None):
    P, M, V, cnt = adc_cycle(p, _moments(p, 1.0), _moments(p, 2.0), accum, count, z.cpu().numpy(), **kw)
    got = st.densify(kw["grad_thresh"], kw["min_opacity"], normals=z,
                     **{k: v for k, v in kw.items() if k not in ("grad_thresh", "min_opacity")})
    assert got == cnt
    n = P["means"].shape[0]
    assert st.n == n
    for name, want, t in (("params", P, st.flat.detach()), ("exp_avg", M, st.exp_avg),
                          ("exp_avg_sq", V, st.exp_avg_sq)):
        g = {k: v.cpu().numpy() for k, v in st.views(t).items()}
        for k in GROUPS:
            want_k = want[k].reshape(g[k].shape)
            if name == "params" and k in ("means", "svec"):
                # rows appended by split carry computed values; the rest are copies
                n_copy = n - (0 if exact_rows is None else exact_rows)
                np.testing.assert_array_equal(g[k][:n_copy], want_k[:n_copy], err_msg=f"{name}.{k}")
                np.testing.assert_allclose(g[k][n_copy:], want_k[n_copy:], rtol=2e-6, atol=2e-6,
                                           err_msg=f"{name}.{k}")
            else:
                np.testing.assert_array_equal(g[k], want_k, err_msg=f"{name}.{k}")
    assert st.grad_accum.shape == (n,) and float(st.grad_accum.abs().sum()) == 0.0
    return cnt


def _clear(x, edges, margin=1e-4):
    """Moves values within `margin` of an edge off it (to edge + 2 margin)."""
    x = x.copy()
    for e in edges:
        close = np.abs(x - e) < margin
        x[close] = e + 2 * margin
    return x


def _random_case(n, K, seed, div=1.6):
    rng = np.random.default_rng(seed)
    svec = rng.uniform(ls(0.003), ls(12.0), (n, 3)).astype(np.float32)
    # log-space edges of every scale test: clone 0.04, split 4, prune 0.4 and
    # a child's prune edge 0.4 * div
    svec = _clear(svec, [ls(0.04), ls(4.0), ls(0.4), ls(0.4 * div)]).astype(np.float32)
    logit = rng.normal(0, 3, n).astype(np.float32)
    logit = _clear(logit, [ls(0.005 / 0.995)]).astype(np.float32)
    count = rng.integers(0, 5, n).astype(np.float32)
    accum = (rng.uniform(0, 2e-3, n) * count).astype(np.float32)
    g = accum / (count + np.float32(1e-5))
    bad = np.abs(np.log(np.maximum(g, 1e-30)) - ls(2e-4)) < 1e-4
    accum[bad] *= np.float32(1.01)
    sc = make_scene(n, 3, seed, n_coeffs=K + 1)
    p = {k: v.numpy().copy() for k, v in sc.params().items()}
    p["svec"], p["opacity"] = svec, logit.reshape(n, 1)
    return p, accum, count


@pytest.mark.gpu
def test_gpu_densify_literal_case(gpu):
    p = _params(6, 15, SIX_SVEC, SIX_OPAC)
    st = _store(p, SIX_ACCUM, [1.0] * 6, gpu)
    z = torch.zeros(2, 3, device=gpu)
    cnt = _check_against_oracle(st, p, SIX_ACCUM, [1.0] * 6, z,
                                dict(grad_thresh=2e-4, min_opacity=0.005))
    assert cnt == (1, 1, 4)
    np.testing.assert_array_equal(st.views()["means"].detach().cpu().numpy(), p["means"][[0, 2, 4, 0]])


@pytest.mark.gpu
@pytest.mark.parametrize("kw", [
    dict(grad_thresh=2e-4, min_opacity=0.005),                                  # trainer values
    dict(grad_thresh=2e-4, min_opacity=0.005, split_scale_min=1.0, prune_scale_max=8.0),
    dict(grad_thresh=2e-4, min_opacity=0.005, split_scale_min=1.0, prune_scale_max=8.0, split_factor=3),
])
def test_gpu_densify_matches_oracle(gpu, kw):
    n, K = 50_000, 15
    div = kw.get("split_factor", 2) * 0.8
    p, accum, count = _random_case(n, K, seed=11, div=div)
    # keep the custom thresholds clear too
    if "split_scale_min" in kw:
        p["svec"] = _clear(p["svec"], [ls(kw["split_scale_min"]), ls(kw["prune_scale_max"]),
                                       ls(kw["prune_scale_max"] * div)]).astype(np.float32)
    st = _store(p, accum, count, gpu)
    z = torch.randn(n * kw.get("split_factor", 2), 3, generator=torch.Generator(device=gpu).manual_seed(3),
                    device=gpu)
    # children are appended last; the rows before them are copies
    f = kw.get("split_factor", 2)
    child_rows = f * _kept_children_per_block(p, accum, count, kw, div)
    got = _check_against_oracle(st, p, accum, count, z, kw, exact_rows=child_rows)
    assert got[0] > 100 and got[1] > 100 and got[2] > 100, got
    if "split_scale_min" in kw:
        assert child_rows > 100


def _kept_children_per_block(p, accum, count, kw, div):
    """Split rows (originals, and clones when thr <= 0) whose children survive the prune."""
    f32 = np.float32
    e = np.exp(p["svec"])
    g = np.abs(np.asarray(accum, f32) / (np.asarray(count, f32) + f32(1e-5))) >= f32(kw["grad_thresh"])
    big = (e > f32(kw.get("split_scale_min", 4.0))).any(-1)
    op = f32(1) / (f32(1) + np.exp(-p["opacity"][:, 0]))
    ce = np.exp(np.log(e / f32(div)))
    keep_ch = ~((op < f32(kw["min_opacity"])) | (ce.max(-1) > f32(kw.get("prune_scale_max", 0.4))))
    gm0 = 0.0 >= kw["grad_thresh"]
    small = (e <= f32(kw.get("clone_scale_max", 0.04))).any(-1)
    split_c = g & small & gm0 & big
    return int((g & big & keep_ch).sum() + (split_c & keep_ch).sum())


@pytest.mark.gpu
def test_gpu_densify_threshold_zero_and_empty(gpu):
    svec = [[ls(0.01), ls(5.0), ls(1.0)], [ls(0.2)] * 3]
    p = _params(2, 15, svec, [1.0, 1.0], [[0, 0, 0], [1, 1, 1]], [[1, 0, 0, 0], [1, 0, 0, 0]])
    st = _store(p, [0.0, 0.0], [0.0, 0.0], gpu)
    z = torch.tensor([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=torch.float32, device=gpu)
    cnt = _check_against_oracle(st, p, [0.0, 0.0], [0.0, 0.0], z,
                                dict(grad_thresh=0.0, min_opacity=0.005, prune_scale_max=10.0),
                                exact_rows=4)
    assert cnt == (1, 2, 0) and st.n == 5
    # everything pruned, then a cycle over an empty store
    p2 = _params(3, 15, [[ls(0.01)] * 3] * 3, [-9.0] * 3)
    st2 = _store(p2, [1.0] * 3, [1.0] * 3, gpu)
    assert st2.densify(2e-4, 0.005) == (3, 0, 6) and st2.n == 0
    assert st2.densify(2e-4, 0.005) == (0, 0, 0) and st2.n == 0


@pytest.mark.gpu
def test_gpu_densify_too_few_samples(gpu):
    svec = [[ls(5.0)] * 3]
    p = _params(1, 15, svec, [1.0])
    st = _store(p, [1.0], [1.0], gpu)
    with pytest.raises(ValueError, match="normal samples"):
        st.densify(2e-4, 0.005, normals=torch.zeros(1, 3, device=gpu))


def _scene_from(p, deg=3):
    from word2world_amd.scene import GaussianScene
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    return GaussianScene(t(p["means"]), t(p["svec"]), t(p["quaternion"]), t(p["opacity"]),
                         t(p["sh_dc"]), t(p["sh_ac"]), deg)


def _batched_rows(st, t, lo, hi):
    """Rows [lo, hi) of a batched-store tensor as the oracle's six groups."""
    v = {k: st.view(k, t)[lo:hi].cpu().numpy() for k in ("means", "sh", "opacity", "svec", "quaternion")}
    return {"means": v["means"], "sh_dc": v["sh"][:, :1], "sh_ac": v["sh"][:, 1:],
            "opacity": v["opacity"].reshape(-1, 1), "svec": v["svec"], "quaternion": v["quaternion"]}


@pytest.mark.gpu
def test_gpu_densify_segments_of_a_batched_store(gpu):
    """Three scenes densified in one call: each scene's new rows stay together
    and equal the oracle's cycle on that scene alone."""
    from word2world_amd import densify as dens
    from word2world_amd.batched import BatchedStore
    kw = dict(grad_thresh=2e-4, min_opacity=0.005, split_scale_min=1.0, prune_scale_max=8.0)
    cases = [_random_case(n, 15, seed=20 + k) for k, n in enumerate((3000, 700, 5000))]
    for p, _, _ in cases:
        p["svec"] = _clear(p["svec"], [ls(1.0), ls(8.0), ls(12.8)]).astype(np.float32)
    st = BatchedStore([_scene_from(p) for p, _, _ in cases], gpu)
    gen = torch.Generator(device=gpu).manual_seed(7)
    st.exp_avg.copy_(torch.rand(st.exp_avg.shape, generator=gen, device=gpu))
    st.exp_avg_sq.copy_(torch.rand(st.exp_avg_sq.shape, generator=gen, device=gpu))
    accum = torch.from_numpy(np.concatenate([a for _, a, _ in cases])).to(gpu)
    count = torch.from_numpy(np.concatenate([c for _, _, c in cases])).to(gpu)
    z = torch.randn(2 * st.G, 3, generator=gen, device=gpu)
    old_ns = list(st.ns)
    ins = [(_batched_rows(st, st.flat, lo, lo + n), _batched_rows(st, st.exp_avg, lo, lo + n),
            _batched_rows(st, st.exp_avg_sq, lo, lo + n)) for lo, n in zip(np.cumsum([0] + old_ns[:-1]), old_ns)]
    cfg = dens.config(kw["grad_thresh"], kw["min_opacity"], split_scale_min=1.0, prune_scale_max=8.0)
    p2, m2, v2, ns, counts = dens.densify_segments([3, 3 * st.M, 1, 3, 4], (0, 2, 3, 4), st.flat,
                                                   st.exp_avg, st.exp_avg_sq, st.ns, accum, count, cfg,
                                                   normals=z)
    st.rebuild(p2, m2, v2, ns)
    zo, lo = 0, 0
    for k, ((pi, mi, vi), (p, accum_k, count_k)) in enumerate(zip(ins, cases)):
        S = counts[k][1]
        P, M, V, cnt = adc_cycle(pi, mi, vi, accum_k, count_k, z[zo:zo + 2 * S].cpu().numpy(), **kw)
        zo += 2 * S
        assert cnt == counts[k] and ns[k] == P["means"].shape[0]
        n = ns[k]
        kids = 2 * _kept_children_per_block(p, accum_k, count_k, kw, 1.6)
        for name, want, t in (("params", P, st.flat), ("exp_avg", M, st.exp_avg), ("exp_avg_sq", V, st.exp_avg_sq)):
            got = _batched_rows(st, t, lo, lo + n)
            for g in GROUPS:
                if name == "params" and g in ("means", "svec"):
                    np.testing.assert_array_equal(got[g][:n - kids], want[g][:n - kids], err_msg=f"{k}.{g}")
                    np.testing.assert_allclose(got[g][n - kids:], want[g][n - kids:], rtol=2e-6, atol=2e-6)
                else:
                    np.testing.assert_array_equal(got[g], want[g], err_msg=f"{k}.{name}.{g}")
        lo += n


@pytest.mark.gpu
def test_gpu_batched_trainer_statistics_and_densify(gpu):
    """The per-step statistics (gaussians_handler.py:116-118) against the
    backward's outputs, then densification inside the captured step loop."""
    from word2world_amd.batched import BatchedTrainer
    from word2world_amd.camera import orbit_camera
    scenes = [make_scene(4000, 3, 1), make_scene(2500, 3, 2)]
    cams = [[orbit_camera(a, 10.0, 2.5, 128, 128) for a in (0.0, 120.0)],
            [orbit_camera(a, -5.0, 2.5, 96, 96) for a in (30.0, 200.0, 300.0)]]
    tr = BatchedTrainer(scenes, cams, None, gpu, densify_every=1000)
    tr.step()
    torch.cuda.synchronize()
    m2d = tr.grads.means2D.cpu().numpy()
    radii = tr.ws.radii.cpu().numpy()
    acc = np.zeros(tr.store.G, np.float32)
    cnt = np.zeros(tr.store.G, np.float32)
    item, row = 0, 0
    for sc, cl in zip(scenes, cams):
        V = len(cl)
        for g in range(sc.n):
            for v in range(V):
                it = item + g * V + v
                if radii[it] > 0:
                    cnt[row + g] += np.float32(1)
                    x, y = m2d[it, 0], m2d[it, 1]
                    acc[row + g] += np.sqrt(x * x + y * y)
        item += sc.n * V
        row += sc.n
    assert cnt.sum() > 1000
    np.testing.assert_array_equal(tr.count.cpu().numpy(), cnt)
    np.testing.assert_array_equal(tr.accum.cpu().numpy(), acc)

    tr2 = BatchedTrainer(scenes, cams, None, gpu, densify_every=2, densify_grad=1e-9,
                         generator=torch.Generator(device=gpu).manual_seed(0))
    tr2.step()
    tr2.capture()
    ns0 = list(tr2.store.ns)
    for _ in range(6):  # densify at steps 2, 4, 6; re-captured at 3, 5, 7
        loss = tr2.step()
    torch.cuda.synchronize()
    assert len(tr2.densify_log) == 3 and tr2.graph is not None
    n = ns0
    for _, counts in tr2.densify_log:
        n = [a + c + 2 * s - s - p for a, (c, s, p) in zip(n, counts)]
    assert n == tr2.store.ns and tr2.store.ns != ns0
    assert [sp.n for sp in tr2.specs] == tr2.store.ns
    assert torch.isfinite(loss).all() and not tr2.overflow()
