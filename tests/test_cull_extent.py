"""The cull extent packed into each splat record must be conservative: a
pixel outside the box never passes the blend kernels' exponent pre-test
(power >= skip_below, evaluated in float exactly as the kernels do), so
dropping such pairs cannot change an image or a gradient.  Checked on the
host build of the same function, against brute force over points around and
outside the box, and against the exact ellipse extent (the box must contain
the ellipse's extreme points and stay close to them)."""
import ctypes

import numpy as np
import pytest

from word2world_amd import _lib


def _extents(conic, skip):
    L = _lib.load()
    conic = np.ascontiguousarray(conic, np.float32)
    skip = np.ascontiguousarray(skip, np.float32)
    out = np.zeros(len(skip), np.uint32)
    assert L.w2w_splat_extents(conic.ctypes.data, skip.ctypes.data, len(skip), out.ctypes.data) == 0
    h = out.view(np.float16).reshape(-1, 2).astype(np.float32)
    return h[:, 0], h[:, 1]


def _f32(x):
    return np.asarray(x, np.float64).astype(np.float32)


def _fma(a, b, c):
    return _f32(np.float64(a) * np.float64(b) + np.float64(c))


def _power(ca, cb, cc, mx, my, px, py):
    """power_of(ca, cb, cc, mx - px, my - py) with the kernels' float32 steps."""
    dx = _f32(np.float64(mx) - np.float64(px))
    dy = _f32(np.float64(my) - np.float64(py))
    q = _fma(_f32(np.float64(ca) * dx), dx, _f32(np.float64(_f32(np.float64(cc) * dy)) * dy))
    t = _f32(np.float64(_f32(np.float64(cb) * dx)) * dy)
    return _fma(np.float32(-0.5), q, -t)


def _splats(n, seed, dilate=0.3, max_aniso=1e3):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, np.pi, n)
    l1 = np.exp(rng.uniform(np.log(0.05), np.log(4e4), n))
    l2 = l1 / np.exp(rng.uniform(0, np.log(max_aniso), n))
    c, s = np.cos(th), np.sin(th)
    cxx = c * c * l1 + s * s * l2 + dilate
    cyy = s * s * l1 + c * c * l2 + dilate
    cxy = c * s * (l1 - l2)
    det = cxx * cyy - cxy * cxy
    conic = _f32(np.stack([cyy / det, -cxy / det, cxx / det], 1))
    o = rng.uniform(0.0045, 1.0, n)
    skip = _f32(-np.log(255.0 * o) - 1e-5)
    mean = _f32(np.stack([rng.uniform(-50, 1100, n), rng.uniform(-50, 1100, n)], 1))
    return conic, skip, mean


@pytest.mark.parametrize("seed,dilate,aniso", [(0, 0.3, 1e3), (1, 0.3, 1e6), (2, 0.0, 1e4)])
def test_no_pixel_outside_the_box_passes_the_pretest(seed, dilate, aniso):
    conic, skip, mean = _splats(400, seed, dilate, aniso)
    hx, hy = _extents(conic, skip)
    rng = np.random.default_rng(100 + seed)
    checked = 0
    for i in range(len(skip)):
        ca, cb, cc = conic[i]
        mx, my = mean[i]
        if not (np.isfinite(hx[i]) and np.isfinite(hy[i])):
            continue
        # points just outside each side of the box, along the whole side,
        # plus random points in a band around it
        t = rng.uniform(-1, 1, 256)
        e = rng.uniform(0, 3, 256) * np.float32(0.01) * (1 + np.float32(hx[i]) / 100)
        f = rng.uniform(0, 3, 256) * np.float32(0.01) * (1 + np.float32(hy[i]) / 100)
        xs = [mx + hx[i] + e, mx - hx[i] - e, mx + t * hx[i] * 1.5, mx + t * hx[i] * 1.5]
        ys = [my + t * hy[i] * 1.5, my + t * hy[i] * 1.5, my + hy[i] + f, my - hy[i] - f]
        for px, py in zip(xs, ys):
            px, py = _f32(px), _f32(py)
            outside = (np.abs(np.float64(mx) - px) > hx[i]) | (np.abs(np.float64(my) - py) > hy[i])
            pw = _power(ca, cb, cc, mx, my, px, py)
            passes = ~((pw > 0) | (pw < skip[i]))
            assert not np.any(passes & outside), (i, conic[i], skip[i], hx[i], hy[i])
            checked += int(outside.sum())
    assert checked > 100_000


def test_box_contains_the_ellipse_and_is_tight():
    conic, skip, _ = _splats(2000, 7)
    hx, hy = _extents(conic, skip)
    a, b, c = (conic[:, k].astype(np.float64) for k in range(3))
    det = a * c - b * b
    q0 = -2.0 * skip.astype(np.float64)
    ex, ey = np.sqrt(q0 * c / det), np.sqrt(q0 * a / det)
    fin = np.isfinite(hx) & np.isfinite(hy)
    assert fin.mean() > 0.99
    # contains the exact extreme points (a swapped axis or a wrong det fails)
    assert np.all(hx[fin] >= ex[fin]) and np.all(hy[fin] >= ey[fin])
    # and is not much larger
    assert np.all(hx[fin] <= ex[fin] * 1.01 + 0.01)
    assert np.all(hy[fin] <= ey[fin] * 1.01 + 0.01)
    # the extreme point of the ellipse in x: d = q0 * A^-1 e_x / sqrt(q0 (A^-1)_xx)
    dx = np.sqrt(q0 * c / det)
    dy = -b / c * dx
    Q = a * dx * dx + 2 * b * dx * dy + c * dy * dy
    np.testing.assert_allclose(Q, q0, rtol=1e-9)


def test_degenerate_conics_never_cull():
    conic = _f32([[1.0, 1.0, 1.0], [0.0, 0.0, 1.0], [-1.0, 0.0, 1.0], [1.0, 0.0, 1.0], [1e-30, 0, 1e30]])
    skip = _f32([-2.0, -2.0, -2.0, 1.0, -2.0])
    hx, hy = _extents(conic, skip)
    assert np.all(np.isinf(hx)) and np.all(np.isinf(hy))


def _touches(spl, rect):
    L = _lib.load()
    spl = np.ascontiguousarray(spl, np.float32)
    rect = np.ascontiguousarray(rect, np.float32)
    out = np.zeros(len(spl), np.uint8)
    assert L.w2w_block_touches(spl.ctypes.data, rect.ctypes.data, len(spl), out.ctypes.data) == 0
    return out.astype(bool)


@pytest.mark.parametrize("seed,dilate,aniso", [(10, 0.3, 1e3), (11, 0.3, 1e5), (12, 0.0, 1e4)])
def test_block_test_is_conservative_and_tighter_than_the_box(seed, dilate, aniso):
    """A dropped (splat, 8x8 block) pair has no pixel centre of the block that
    passes the float pre-test; brute force over all 64 centres.  Also the
    exact test keeps no more than the box test and drops a good share of
    what the box keeps."""
    conic, skip, mean = _splats(3000, seed, dilate, aniso)
    rng = np.random.default_rng(seed)
    hx, hy = _extents(conic, skip)
    # one 8x8 block per splat, placed near the box so both outcomes occur
    ox = mean[:, 0] + rng.uniform(-1.2, 1.2, len(skip)) * np.minimum(hx, 400) - 4
    oy = mean[:, 1] + rng.uniform(-1.2, 1.2, len(skip)) * np.minimum(hy, 400) - 4
    x0, y0 = np.floor(ox).astype(np.float32), np.floor(oy).astype(np.float32)
    rect = np.stack([x0, x0 + 7, y0, y0 + 7], 1)
    spl = np.concatenate([mean, conic, skip[:, None]], 1)
    keep = _touches(spl, rect)
    box = ~(((x0 - mean[:, 0]) > hx) | ((mean[:, 0] - x0 - 7) > hx) | ((y0 - mean[:, 1]) > hy) |
            ((mean[:, 1] - y0 - 7) > hy))
    dropped = np.nonzero(~keep)[0]
    assert len(dropped) > 300 and np.all(keep <= box | ~np.isfinite(hx))
    assert (box & ~keep).sum() > 0.05 * box.sum()
    ii, jj = np.meshgrid(np.arange(8, dtype=np.float32), np.arange(8, dtype=np.float32))
    for i in dropped:
        ca, cb, cc = conic[i]
        px = x0[i] + ii.reshape(-1)
        py = y0[i] + jj.reshape(-1)
        pw = _power(ca, cb, cc, mean[i, 0], mean[i, 1], px, py)
        assert not np.any(~((pw > 0) | (pw < skip[i]))), i


def test_block_test_keeps_centres_inside_and_degenerate_cases():
    spl = _f32([[4.5, 4.5, 1.0, 0.0, 1.0, -2.0],      # centre inside the block
                [100.0, 100.0, 1.0, 1.0, 1.0, -2.0],  # singular conic
                [100.0, 100.0, 1.0, 0.0, 1.0, 1.0],   # non-negative skip
                [np.nan, 0.0, 1.0, 0.0, 1.0, -2.0],   # NaN position
                [100.0, 100.0, 1.0, 0.0, 1.0, -2.0]]) # far away, round: dropped
    rect = _f32([[0, 7, 0, 7]] * 5)
    assert _touches(spl, rect).tolist() == [True, True, True, True, False]
