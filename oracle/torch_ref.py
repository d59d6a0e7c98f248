"""Differentiable float64 restatement of the forward, for gradient checks.

TEST INFRASTRUCTURE ONLY (see oracle/oracle.py).

The rasterizer's discrete decisions -- which Gaussians pass the near and
determinant culls, which instances each pixel blends before termination --
are taken from the C oracle's float32 run (``rec_ids``/``rec_count``).  Every
continuous quantity is recomputed here in float64 with torch ops, so
``torch.autograd`` gives the exact derivative of the forward as defined
(projection.cu:7-154 for the geometry, gaussianForwardPass.cu:172-264 for SH,
pixelGaussiansDev.cuh:20-79 for the blend, PD-9/PD-10).  This is the
"reference autograd gradient" the hand-written backward is checked against;
float64 central finite differences of the same function check it in turn.
"""
from __future__ import annotations

import numpy as np
import torch

SH_C0 = 0.28209479177387814
SH_C1 = 0.4886025119029199
SH_C2A = 1.0925484305920792
SH_C2E = 0.5462742152960396
SH_C26 = 0.3
SH_K = [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
        -0.4570457994644658, 1.445305721320277, -0.5900435899266435]


def sh_basis(D: int, d: torch.Tensor) -> torch.Tensor:
    """Real SH basis with the reference constants; [N, (D+1)^2]."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    out = [torch.full_like(x, SH_C0)]
    if D >= 1:
        out += [-SH_C1 * y, SH_C1 * z, -SH_C1 * x]
    if D >= 2:
        xx, yy, zz = x * x, y * y, z * z
        out += [SH_C2A * x * y, -SH_C2A * y * z, SH_C26 * (2 * zz - xx - yy), -SH_C2A * x * z,
                SH_C2E * (xx - yy)]
    if D >= 3:
        k = SH_K
        out += [k[0] * y * (3 * xx - yy), k[1] * x * y * z, k[2] * y * (4 * zz - xx - yy),
                k[3] * z * (2 * zz - 3 * xx - 3 * yy), k[4] * x * (4 * zz - xx - yy),
                k[5] * z * (xx - yy), k[6] * x * (xx - 3 * yy)]
    return torch.stack(out, dim=1)


def _mat(flat) -> torch.Tensor:
    """Column-major 16-float storage -> [4,4] matrix (row, col)."""
    return torch.as_tensor(np.asarray(flat, np.float64)).reshape(4, 4).T.contiguous()


def preprocess64(means, scales, rots, sh, D, view, proj, campos, W, H, fx, fy):
    """Per-Gaussian continuous quantities in float64 (no culling applied)."""
    Vm, Fm = _mat(view), _mat(proj)
    N = means.shape[0]
    ph = torch.cat([means, torch.ones(N, 1, dtype=means.dtype)], dim=1)
    u = ph @ Vm.T
    t = u[:, :3] / u[:, 3:4]
    vz = t[:, 2]
    limx = W / (2.0 * fx)
    limy = H / (2.0 * fy)
    rx, ry = t[:, 0] / vz, t[:, 1] / vz
    sq = (rx / limx) ** 2 + (ry / limy) ** 2
    scl = torch.where(sq > 1.0, 1.0 / torch.sqrt(sq), torch.ones_like(sq))
    txc, tyc = rx * scl * vz, ry * scl * vz
    right, up, fwd = Vm[0, :3], Vm[1, :3], Vm[2, :3]
    zi = 1.0 / vz
    jx = fx * (right[None, :] * zi[:, None] - (txc * zi * zi)[:, None] * fwd[None, :])
    jy = fy * (up[None, :] * zi[:, None] - (tyc * zi * zi)[:, None] * fwd[None, :])
    w, x, y, z = rots[:, 0], rots[:, 1], rots[:, 2], rots[:, 3]
    R = torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
        2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
        2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], dim=1).reshape(N, 3, 3)
    Mm = R * scales[:, None, :]
    Sig = Mm @ Mm.transpose(1, 2)
    v = (Sig @ jx[:, :, None])[:, :, 0]
    wv = (Sig @ jy[:, :, None])[:, :, 0]
    cxx = (jx * v).sum(1)
    cxy = (jx * wv).sum(1)
    cyy = (jy * wv).sum(1)
    det = cxx * cyy - cxy * cxy
    conic = torch.stack([cyy / det, -cxy / det, cxx / det], dim=1)
    clip = ph @ Fm.T
    invW = 1.0 / (clip[:, 3] + 1e-5)
    ndc = clip[:, :3] * invW[:, None]
    px = (ndc[:, 0] + 1.0) * 0.5 * W - 0.5
    py = (ndc[:, 1] + 1.0) * 0.5 * H - 0.5
    cp = torch.as_tensor(np.asarray(campos, np.float64))
    dvec = means - cp[None, :]
    dirs = dvec / torch.sqrt((dvec * dvec).sum(1, keepdim=True))
    nb = (D + 1) ** 2
    Y = sh_basis(D, dirs)
    rgb = (Y[:, :, None] * sh[:, :nb, :]).sum(1) + 0.5
    cov3d = torch.stack([Sig[:, 0, 0], Sig[:, 0, 1], Sig[:, 0, 2], Sig[:, 1, 1], Sig[:, 1, 2],
                         Sig[:, 2, 2]], dim=1)
    return dict(m2d=torch.stack([px, py], 1), ndc=ndc, depth=ndc[:, 2], conic=conic, rgb=rgb,
                cov3d=cov3d, det=det)


def blend_fixed(pre, opac, rec_count, rec_ids, W, H, bg):
    """Composite each pixel over its recorded contributor list, front to back."""
    HW = W * H
    L = int(rec_count.max()) if rec_count.size else 0
    L = max(L, 1)
    ids = torch.as_tensor(rec_ids[:, :L].astype(np.int64))
    valid = torch.as_tensor(np.arange(L)[None, :] < rec_count[:, None])
    ids = torch.where(valid, ids, torch.zeros_like(ids))
    pix = torch.arange(HW)
    pxf = (pix % W).double()[:, None]
    pyf = (pix // W).double()[:, None]
    m2d, conic = pre["m2d"], pre["conic"]
    dx = m2d[ids, 0] - pxf
    dy = m2d[ids, 1] - pyf
    a, b, c = conic[ids, 0], conic[ids, 1], conic[ids, 2]
    power = -0.5 * (a * dx * dx + c * dy * dy) - b * dx * dy
    G = torch.exp(power)
    alpha = torch.clamp(opac[ids] * G, max=0.99)
    alpha = torch.where(valid, alpha, torch.zeros_like(alpha))
    one_m = 1.0 - alpha
    Tcum = torch.cumprod(torch.cat([torch.ones(HW, 1, dtype=alpha.dtype), one_m], 1), 1)
    Ti = Tcum[:, :-1]
    Tf = Tcum[:, -1]
    wgt = alpha * Ti
    col = (wgt[:, :, None] * pre["rgb"][ids]).sum(1) + Tf[:, None] * torch.as_tensor(
        np.asarray(bg, np.float64))[None, :]
    A = wgt.sum(1)
    Dd = (wgt * pre["depth"][ids]).sum(1)
    return col.T.reshape(3, H, W), A.reshape(H, W), Dd.reshape(H, W), Tf.reshape(H, W)


def render_fixed(means, scales, rots, opac, sh, D, view, proj, campos, bg, W, H, fx, fy,
                 rec_count, rec_ids):
    pre = preprocess64(means, scales, rots, sh, D, view, proj, campos, W, H, fx, fy)
    return blend_fixed(pre, opac, rec_count, rec_ids, W, H, bg) + (pre,)


def grads_fixed(fwd: dict, means, scales, rots, opac, sh, view, proj, campos, bg, fx, fy,
                dL_dcolor, dL_dalpha, dL_ddepth):
    """Autograd gradients of <dL/dout, out> w.r.t. every input, float64."""
    W, H, D = fwd["W"], fwd["H"], fwd["D"]
    ts = [torch.tensor(np.asarray(a, np.float64), requires_grad=True)
          for a in (means, scales, rots, np.asarray(opac).reshape(-1), sh)]
    m, s, r, o, shc = ts
    col, A, Dd, Tf, pre = render_fixed(m, s, r, o, shc, D, view, proj, campos, bg, W, H, fx, fy,
                                       fwd["rec_count"], fwd["rec_ids"])
    # means2D: the NDC xy of every Gaussian, as a leaf the loss reaches through pre
    loss = (col * torch.as_tensor(np.asarray(dL_dcolor, np.float64))).sum() + \
           (A * torch.as_tensor(np.asarray(dL_dalpha, np.float64).reshape(H, W))).sum() + \
           (Dd * torch.as_tensor(np.asarray(dL_ddepth, np.float64).reshape(H, W))).sum()
    gm, gs, gr, go, gsh = torch.autograd.grad(loss, ts)
    return dict(dL_dmeans3D=gm.numpy(), dL_dscales=gs.numpy(), dL_drotations=gr.numpy(),
                dL_dopacity=go.numpy(), dL_dsh=gsh.numpy(), color=col.detach().numpy(),
                alpha=A.detach().numpy(), depth=Dd.detach().numpy(), final_T=Tf.detach().numpy())
