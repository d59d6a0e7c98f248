"""Host side of the rasterizer: workspaces, the autograd Function and a
drop-in for the reference's ``_C`` module.

Mirrors the reference's operator interface:

* ``rasterize_gaussians`` / ``rasterize_gaussians_backward`` keep the
  positional signatures of ``_C`` (ext.cpp:5-7; rasterize_points.h:8-48) and
  their return tuples, so ``Renderer.py``-style callers work unchanged
  (``Renderer.py:54-66`` and ``:88-93``).  Extra keyword arguments carry what
  the reference hard-coded: image size (800x800), background and tile size.
* ``RasterizeGaussians`` is the autograd Function of ``Renderer.py:43-104``:
  inputs ``means3D, means2D, sh, opacities, scales, rotations`` (activated),
  outputs ``color, radii, depth, alpha``; ``means2D.grad`` receives
  dL/d(NDC xy) as ``accum_grads`` does in the reference.

All compute is the HIP library (word2world_amd/csrc); nothing here falls back
to the CPU.  Work is issued on torch's current stream.
"""
from __future__ import annotations

import ctypes
import weakref
from dataclasses import dataclass

import torch

from . import _lib


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def _check_f32(name: str, t: torch.Tensor) -> torch.Tensor:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != torch.float32:
        raise ValueError(f"{name} must be float32, got {t.dtype}")
    return t.contiguous()


class Workspace:
    """Caller-owned geometry / binning / image state of one view (the
    reference's geomBuffer, binningBuffer and imgBuffer)."""

    def __init__(self, P: int, W: int, H: int, tile: int, capacity: int, device):
        self.P, self.W, self.H, self.tile, self.capacity = P, W, H, tile, int(capacity)
        self.layout = _lib.layout(P, W, H, tile, self.capacity)
        L = self.layout
        self.geom = torch.empty(max(L.geom_bytes, 1), dtype=torch.uint8, device=device)
        self.binning = torch.empty(max(L.binning_bytes, 1), dtype=torch.uint8, device=device)
        self.image = torch.empty(max(L.image_bytes, 1), dtype=torch.uint8, device=device)
        self._struct = _lib.W2WWorkspace(self.geom.data_ptr(), L.geom_bytes, self.binning.data_ptr(),
                                         L.binning_bytes, self.image.data_ptr(), L.image_bytes,
                                         self.capacity)

    @classmethod
    def from_buffers(cls, geom, binning, image, P, W, H, tile, capacity) -> "Workspace":
        """A view of existing buffers (no allocation); their sizes must match
        the layout of (P, W, H, tile, capacity)."""
        ws = cls.__new__(cls)
        ws.P, ws.W, ws.H, ws.tile, ws.capacity = P, W, H, tile, int(capacity)
        ws.layout = L = _lib.layout(P, W, H, tile, ws.capacity)
        for name, t, need in (("geomBuffer", geom, L.geom_bytes), ("binningBuffer", binning, L.binning_bytes),
                              ("imgBuffer", image, L.image_bytes)):
            if t.dtype != torch.uint8 or not t.is_cuda or t.numel() < need:
                raise ValueError(f"{name} is not a workspace buffer of this size")
        ws.geom, ws.binning, ws.image = geom, binning, image
        ws._struct = _lib.W2WWorkspace(geom.data_ptr(), L.geom_bytes, binning.data_ptr(), L.binning_bytes,
                                       image.data_ptr(), L.image_bytes, ws.capacity)
        return ws

    def fits(self, P, W, H, tile, capacity) -> bool:
        return (self.P, self.W, self.H, self.tile) == (P, W, H, tile) and self.capacity >= capacity

    @property
    def struct(self) -> _lib.W2WWorkspace:
        return self._struct

    # ---- device-side views (no synchronisation) ----
    def status(self) -> torch.Tensor:
        """int32[2]: (K, overflow)."""
        return self.geom[:8].view(torch.int32)

    def _view(self, buf, off, n, dtype, shape=None):
        nbytes = n * torch.empty((), dtype=dtype).element_size()
        t = buf[off:off + nbytes].view(dtype)
        return t.view(shape) if shape is not None else t

    def radii(self):
        return self._view(self.geom, self.layout.radii, self.P, torch.int32)

    def counts(self):
        return self._view(self.geom, self.layout.counts, self.P, torch.int32)

    def rect(self):
        return self._view(self.geom, self.layout.rect, 2 * self.P, torch.int32, (self.P, 2))

    def perm(self):
        return self._view(self.geom, self.layout.perm, self.P, torch.int32)

    def offsets(self):
        return self._view(self.geom, self.layout.offsets, self.P, torch.int32)

    def splats(self):
        return self._view(self.geom, self.layout.splats, 16 * self.P, torch.float32, (self.P, 16))

    def sorted_ids(self, n: int):
        return self._view(self.binning, self.layout.sorted_ids, n, torch.int32)

    def sorted_tiles(self, n: int):
        """The first n sorted tile keys, as int32 (stored 16-bit when the view
        has at most 65536 tiles)."""
        if self.layout.tile_key_bytes == 2:
            return self._view(self.binning, self.layout.sorted_tiles, n, torch.int16).to(torch.int32) & 0xFFFF
        return self._view(self.binning, self.layout.sorted_tiles, n, torch.int32)

    def ranges(self):
        T = ((self.W + self.tile - 1) // self.tile) * ((self.H + self.tile - 1) // self.tile)
        return self._view(self.image, self.layout.ranges, 2 * T, torch.int32, (T, 2))

    def n_contrib(self):
        return self._view(self.image, self.layout.n_contrib, self.W * self.H, torch.int32)

    def final_T(self):
        return self._view(self.image, self.layout.final_T, self.W * self.H, torch.float32)


@dataclass
class Camera:
    """Device copies of the per-view camera inputs."""
    viewmatrix: torch.Tensor   # [4,4] (column-major W2C storage, as Renderer.py passes)
    projmatrix: torch.Tensor   # [4,4]
    campos: torch.Tensor       # [3]
    background: torch.Tensor   # [3]
    width: int
    height: int
    focal_x: float
    focal_y: float


def make_frame(means, scales, rotations, opacities, sh, sh_degree: int, cam: Camera, tile: int,
               chunk: int = 0):
    P = int(means.shape[0])
    M = int(sh.shape[1]) if sh.numel() else 1
    f = _lib.W2WFrame(P, M, int(sh_degree), int(cam.width), int(cam.height), int(tile), int(chunk), 0,
                      float(cam.focal_x), float(cam.focal_y), _ptr(means), _ptr(scales),
                      _ptr(rotations), _ptr(opacities), _ptr(sh), _ptr(cam.viewmatrix),
                      _ptr(cam.projmatrix), _ptr(cam.campos), _ptr(cam.background))
    return f


class CapacityPolicy:
    """Instance-capacity guesses per (P, W, H, tile); grows on overflow."""

    def __init__(self, slack: float = 1.25, minimum: int = 1 << 16):
        self.slack, self.minimum = slack, minimum
        self._last: dict = {}

    def guess(self, key) -> int:
        P = key[0]
        return self._last.get(key, max(self.minimum, 8 * P))

    def observe(self, key, K: int) -> int:
        cap = max(self.minimum, int(K * self.slack) + 1024)
        self._last[key] = max(cap, self._last.get(key, 0))
        return self._last[key]


POLICY = CapacityPolicy()


def _inputs(means, scales, rotations, opacities, sh):
    means = _check_f32("means3D", means)
    scales = _check_f32("scales", scales)
    rotations = _check_f32("rotations", rotations)
    opacities = _check_f32("opacities", opacities).reshape(-1)
    sh = _check_f32("sh", sh)
    P = means.shape[0]
    if means.shape != (P, 3) or scales.shape != (P, 3) or rotations.shape != (P, 4) \
            or opacities.shape != (P,) or (P and (sh.dim() != 3 or sh.shape[0] != P or sh.shape[2] != 3)):
        raise ValueError("bad Gaussian tensor shapes")
    return means, scales, rotations, opacities, sh


def forward(means, scales, rotations, opacities, sh, sh_degree: int, cam: Camera, tile: int = 16,
            workspace: Workspace | None = None, capacity: int | None = None, chunk: int = 0):
    """One view.  Returns (color [3,H,W], depth [H,W], alpha [H,W], radii [P],
    workspace).  K and the overflow flag stay on the device
    (``workspace.status()``); nothing here synchronises.  chunk: instances
    per forward-blend chunk (64, 128 or 256; 0 = 64), a speed knob only."""
    means, scales, rotations, opacities, sh = _inputs(means, scales, rotations, opacities, sh)
    P, W, H = int(means.shape[0]), int(cam.width), int(cam.height)
    key = (P, W, H, tile)
    cap = int(capacity) if capacity is not None else POLICY.guess(key)
    if workspace is None or not workspace.fits(P, W, H, tile, cap):
        workspace = Workspace(P, W, H, tile, cap, means.device)
    dev = means.device
    color = torch.empty(3, H, W, dtype=torch.float32, device=dev)
    depth = torch.empty(H, W, dtype=torch.float32, device=dev)
    alpha = torch.empty(H, W, dtype=torch.float32, device=dev)
    radii = torch.empty(P, dtype=torch.int32, device=dev)
    frame = make_frame(means, scales, rotations, opacities, sh, sh_degree, cam, tile, chunk)
    st = _lib.load().w2w_forward(ctypes.byref(frame), ctypes.byref(workspace.struct), color.data_ptr(),
                                 depth.data_ptr(), alpha.data_ptr(), _ptr(radii) if P else None,
                                 _stream())
    _lib.check(st, "w2w_forward")
    return color, depth, alpha, radii, workspace


def forward_checked(means, scales, rotations, opacities, sh, sh_degree: int, cam: Camera,
                    tile: int = 16, workspace: Workspace | None = None, chunk: int = 0):
    """``forward`` that reads K back (one synchronisation) and re-runs with a
    larger workspace when the instance capacity overflowed."""
    P = int(means.shape[0])
    key = (P, int(cam.width), int(cam.height), tile)
    out = forward(means, scales, rotations, opacities, sh, sh_degree, cam, tile, workspace, chunk=chunk)
    K, ovf = (int(v) for v in out[4].status().cpu())
    if ovf:
        cap = POLICY.observe(key, K)
        out = forward(means, scales, rotations, opacities, sh, sh_degree, cam, tile, None, cap, chunk)
        K2, ovf2 = (int(v) for v in out[4].status().cpu())
        assert not ovf2 and K2 == K, "instance count changed between identical runs"
    else:
        POLICY.observe(key, K)
    return (K,) + out


def backward(means, scales, rotations, opacities, sh, sh_degree: int, cam: Camera, tile: int,
             workspace: Workspace, dL_dcolor, dL_ddepth, dL_dalpha):
    """Gradients in the reference's tuple order (rasterize_points.cu:188)."""
    means, scales, rotations, opacities, sh = _inputs(means, scales, rotations, opacities, sh)
    P = int(means.shape[0])
    dev = means.device
    M = int(sh.shape[1]) if sh.numel() else 1
    z = lambda *s: torch.empty(*s, dtype=torch.float32, device=dev)
    g_m2d, g_col, g_op, g_m3d, g_cov = z(P, 3), z(P, 3), z(P, 1), z(P, 3), z(P, 6)
    g_sh, g_sc, g_rot = z(P, M, 3), z(P, 3), z(P, 4)
    dL_dcolor = _check_f32("dL_dout_color", dL_dcolor)
    dL_ddepth = _check_f32("dL_dout_depth", dL_ddepth)
    dL_dalpha = _check_f32("dL_dout_alpha", dL_dalpha)
    frame = make_frame(means, scales, rotations, opacities, sh, sh_degree, cam, tile)
    st = _lib.load().w2w_backward(ctypes.byref(frame), ctypes.byref(workspace.struct),
                                  dL_dcolor.data_ptr(), dL_ddepth.data_ptr(), dL_dalpha.data_ptr(),
                                  g_m2d.data_ptr(), g_col.data_ptr(), g_op.data_ptr(),
                                  g_m3d.data_ptr(), g_cov.data_ptr(), g_sh.data_ptr(),
                                  g_sc.data_ptr(), g_rot.data_ptr(), _stream())
    _lib.check(st, "w2w_backward")
    return g_m2d, g_col, g_op, g_m3d, g_cov, g_sh, g_sc, g_rot


def sorted_keys(means, scales, rotations, opacities, sh, sh_degree, cam, tile, workspace, K: int):
    """uint64 keys (as int64) of the sorted instances, for parity checks."""
    out = torch.empty(max(K, 1), dtype=torch.int64, device=means.device)
    frame = make_frame(*_inputs(means, scales, rotations, opacities, sh), sh_degree, cam, tile)
    _lib.check(_lib.load().w2w_sorted_keys(ctypes.byref(frame), ctypes.byref(workspace.struct),
                                           out.data_ptr(), _stream()), "w2w_sorted_keys")
    return out[:K]


class RasterizeGaussians(torch.autograd.Function):
    """``Renderer.py:43-104`` on the HIP rasterizer."""

    @staticmethod
    def forward(ctx, means3D, means2D, sh, opacities, scales, rotations, cam: Camera, sh_degree: int,
                tile: int, workspace_cb):
        color, depth, alpha, radii, ws = forward(means3D, scales, rotations, opacities, sh,
                                                 sh_degree, cam, tile)
        if workspace_cb is not None:
            workspace_cb(ws)
        ctx.cam, ctx.sh_degree, ctx.tile, ctx.ws = cam, sh_degree, tile, ws
        ctx.save_for_backward(means3D, sh, opacities, scales, rotations)
        ctx.mark_non_differentiable(radii)
        return color, radii, depth, alpha

    @staticmethod
    def backward(ctx, g_color, g_radii, g_depth, g_alpha):
        means3D, sh, opacities, scales, rotations = ctx.saved_tensors
        H, W = ctx.cam.height, ctx.cam.width
        dev = means3D.device
        g_color = torch.zeros(3, H, W, device=dev) if g_color is None else g_color
        g_depth = torch.zeros(H, W, device=dev) if g_depth is None else g_depth
        g_alpha = torch.zeros(H, W, device=dev) if g_alpha is None else g_alpha
        g_m2d, _, g_op, g_m3d, _, g_sh, g_sc, g_rot = backward(
            means3D, scales, rotations, opacities, sh, ctx.sh_degree, ctx.cam, ctx.tile, ctx.ws,
            g_color, g_depth.reshape(H, W), g_alpha.reshape(H, W))
        ctx.ws = None
        return (g_m3d, g_m2d, g_sh, g_op.view_as(opacities), g_sc, g_rot, None, None, None, None)


def rasterize(means3D, means2D, sh, opacities, scales, rotations, cam: Camera, sh_degree: int,
              tile: int = 16, workspace_cb=None):
    """Differentiable render of one view: (color [3,H,W], radii, depth [H,W],
    alpha [H,W]).  ``means2D`` is a zeros [P,3] tensor whose ``.grad`` becomes
    dL/d(NDC xy) (Renderer.py:126-130)."""
    return RasterizeGaussians.apply(means3D, means2D, sh, opacities, scales, rotations, cam,
                                    sh_degree, tile, workspace_cb)


def camera_from_tensors(viewmatrix, projmatrix, campos, width, height, focal_x, focal_y,
                        background=None, device=None) -> Camera:
    dev = device if device is not None else viewmatrix.device
    f = lambda t: torch.as_tensor(t, dtype=torch.float32).to(dev).contiguous()
    bg = f(background if background is not None else [1.0, 1.0, 1.0])
    return Camera(f(viewmatrix), f(projmatrix), f(campos).reshape(3), bg, int(width), int(height),
                  float(focal_x), float(focal_y))


# ---------------------------------------------------------------------------
# Drop-in for the reference's `_C` module (ext.cpp:5-7)
# ---------------------------------------------------------------------------
class _DropInState:
    """What the backward needs besides the three buffers: the sizes that carve
    them and the forward's camera.  It is kept alive only by an attribute of
    the binning buffer it describes, so it lives exactly as long as that
    tensor (the reference's buffers die with the autograd ctx that saved
    them, rasterize_points.cu:24-30, Renderer.py:78-79).  It holds no strong
    reference to any of the buffers."""

    __slots__ = ("P", "W", "H", "tile", "capacity", "cam", "ptrs", "__weakref__")

    def __init__(self, ws: Workspace, cam: Camera):
        self.P, self.W, self.H, self.tile, self.capacity = ws.P, ws.W, ws.H, ws.tile, ws.capacity
        self.cam = cam
        self.ptrs = (ws.geom.data_ptr(), ws.binning.data_ptr(), ws.image.data_ptr())


# binning-buffer address -> state, for buffers whose Python object was
# re-created (the attribute is the primary link).  Weak values: an entry
# disappears with the binning buffer that owns its state.
_DROPIN_STATE: "weakref.WeakValueDictionary[int, _DropInState]" = weakref.WeakValueDictionary()
_STATE_ATTR = "_w2w_dropin_state"


def _attach_state(ws: Workspace, cam: Camera) -> None:
    st = _DropInState(ws, cam)
    setattr(ws.binning, _STATE_ATTR, st)
    _DROPIN_STATE[ws.binning.data_ptr()] = st


def _find_state(geomBuffer, binningBuffer, imgBuffer) -> _DropInState:
    st = getattr(binningBuffer, _STATE_ATTR, None)
    if st is None:
        st = _DROPIN_STATE.get(binningBuffer.data_ptr())
    if st is None or st.ptrs != (geomBuffer.data_ptr(), binningBuffer.data_ptr(), imgBuffer.data_ptr()):
        raise RuntimeError("geomBuffer / binningBuffer / imgBuffer do not come from one "
                           "rasterize_gaussians call of this module")
    return st


def dropin_live_states() -> int:
    """Number of drop-in forward states still alive (for leak checks)."""
    return len(_DROPIN_STATE)


def rasterize_gaussians(sh, focal_y, scales, means3D, projmatrix, sh_degree, opacities, rotations,
                        campos, viewmatrix, focal_x, *, image_width: int = 800,
                        image_height: int = 800, background=None, tile: int = 16):
    """``_C.rasterize_gaussians`` (rasterize_points.cu:31-100): returns
    ``(num_rendered, color[3,H,W], depth[1,H,W], alpha[1,H,W], radii[P],
    geomBuffer, binningBuffer, imgBuffer)``.  Reads K back, as the reference
    does (prepareSort.cu:382-387).  The three buffers own the view's state:
    nothing else keeps it, so dropping them frees it."""
    cam = camera_from_tensors(viewmatrix, projmatrix, campos, image_width, image_height, focal_x,
                              focal_y, background, device=means3D.device)
    K, color, depth, alpha, radii, ws = forward_checked(means3D, scales, rotations, opacities, sh,
                                                        sh_degree, cam, tile)
    _attach_state(ws, cam)
    return (K, color, depth.unsqueeze(0), alpha.unsqueeze(0), radii, ws.geom, ws.binning, ws.image)


def rasterize_gaussians_backward(background, means3D, radii, colors_precomp, scales, rotations,
                                 scale_modifier, cov3D_precomp, viewmatrix, projmatrix, tan_fovx,
                                 tan_fovy, dL_dout_color, dL_dout_depth, dL_dout_alpha, sh, degree,
                                 campos, geomBuffer, R, binningBuffer, imgBuffer, alpha, debug):
    """``_C.rasterize_gaussians_backward`` (rasterize_points.cu:101-189):
    returns ``(dL_dmeans2D, dL_dcolors, dL_dopacity, dL_dmeans3D, dL_dcov3D,
    dL_dsh, dL_dscales, dL_drotations)``.  The opacities the forward used are
    read back from the forward's state; H, W come from dL_dout_color as in the
    reference."""
    st = _find_state(geomBuffer, binningBuffer, imgBuffer)
    ws = Workspace.from_buffers(geomBuffer, binningBuffer, imgBuffer, st.P, st.W, st.H, st.tile,
                                st.capacity)
    cam = st.cam
    H, W = int(dL_dout_color.shape[1]), int(dL_dout_color.shape[2])
    if (H, W) != (cam.height, cam.width):
        raise ValueError("dL_dout_color does not match the forward's image size")
    opac = ws.splats()[:, 5].contiguous()
    return backward(means3D, scales, rotations, opac, sh, degree, cam, ws.tile, ws,
                    dL_dout_color, dL_dout_depth.reshape(H, W), dL_dout_alpha.reshape(H, W))
