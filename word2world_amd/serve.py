"""Served form: a request loop with continuous batching over HBM-resident
scenes, against a plain one-request-at-a-time server.

Requests are renders (forward only, 1-8 views) or optimisation jobs (N steps
of fwd+bwd+Adam over 4 views of a resident scene, L2 to target renders).
Scene parameters and Adam moments stay in one HBM arena (group-major flat
store with a first-fit row allocator); a scene is admitted when the arena can
hold it (the HBM budget) and leaves when its request completes.

Every tick the continuous server gathers all work that is ready -- each newly
admitted render and one step of every active optimisation job, across scenes
of any size and resolution -- into one batch and runs it as one launch
sequence: activation over the batch's rows, the batched forward, the loss,
the batched backward (gradients land in the arena rows), activation backward
and one Adam launch over every job's parameter ranges, each job with its own
step count.  Requests that arrive while a tick runs are admitted at the next
tick as memory and batch slots free up.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, batch, raster
from .ply import save_ply
from .scene import GaussianScene, activate, load_scene
from .train import LRS, PlainTrainer, TrainConfig, mean_lr

GROUPS = ("means", "sh", "opacity", "svec", "quaternion")
ROW_RANGE = np.dtype([("start", np.int64), ("off", np.int64), ("len", np.int64)])
PARAM_RANGE = np.dtype([("start", np.int64), ("off", np.int64), ("len", np.int64), ("hyper", np.int32),
                        ("period", np.uint32), ("split", np.uint32), ("pad", np.uint32)])


class Arena:
    """Resident scenes: parameters, gradients and Adam moments in one flat
    store of `rows` Gaussians, group-major."""

    def __init__(self, rows: int, M: int, device):
        self.rows, self.M, self.device = rows, M, torch.device(device)
        self.sizes = [3 * rows, 3 * M * rows, rows, 3 * rows, 4 * rows]
        self.starts = np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(np.int64)
        n = int(sum(self.sizes))
        z = lambda: torch.zeros(n, dtype=torch.float32, device=device)
        self.flat, self.grad, self.exp_avg, self.exp_avg_sq = z(), z(), z(), z()
        f32 = dict(dtype=torch.float32, device=device)
        self.scales = torch.zeros(rows, 3, **f32)
        self.opac = torch.zeros(rows, **f32)
        self.rots = torch.zeros(rows, 4, **f32)
        self.g_scales = torch.zeros(rows, 3, **f32)
        self.g_opac = torch.zeros(rows, **f32)
        self.g_rots = torch.zeros(rows, 4, **f32)
        self.g_cov = torch.zeros(rows, 6, **f32)
        self.free_list = [(0, rows)]

    @staticmethod
    def bytes_per_row(M: int) -> int:
        return 4 * 4 * (3 + 3 * M + 1 + 3 + 4) + 4 * (3 + 1 + 4) * 2 + 4 * 6

    def group(self, t, name):
        k = GROUPS.index(name)
        return t[int(self.starts[k]):int(self.starts[k]) + self.sizes[k]]

    def alloc(self, n: int):
        for i, (s, ln) in enumerate(self.free_list):
            if ln >= n:
                if ln == n:
                    self.free_list.pop(i)
                else:
                    self.free_list[i] = (s + n, ln - n)
                return s
        return None

    def free(self, start: int, n: int):
        fl = sorted(self.free_list + [(start, n)])
        merged = []
        for s, ln in fl:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((s, ln))
        self.free_list = merged

    def load(self, base: int, params: dict):
        """params: device tensors of one scene (means, sh [n,M,3], opacity,
        svec, quaternion); Adam moments of the rows are zeroed."""
        n = params["means"].shape[0]
        for name in GROUPS:
            w = {"means": 3, "sh": 3 * self.M, "opacity": 1, "svec": 3, "quaternion": 4}[name]
            sl = slice(base * w, (base + n) * w)
            self.group(self.flat, name)[sl] = params[name].reshape(-1)
            self.group(self.exp_avg, name)[sl] = 0.0
            self.group(self.exp_avg_sq, name)[sl] = 0.0

    def sh_float_offset(self, base: int) -> int:
        return int(self.starts[1]) + base * 3 * self.M

    def width(self, name):
        return {"means": 3, "sh": 3 * self.M, "opacity": 1, "svec": 3, "quaternion": 4}[name]


@dataclass
class Request:
    rid: int
    kind: str                 # "render" | "optimize"
    arrival: float            # seconds after the stream start
    scene: int                # index into the scene pool
    cameras: list             # word2world_amd.camera.Camera
    steps: int = 0
    target_offs: list = field(default_factory=list)
    # runtime
    base: int = -1
    steps_done: int = 0
    admitted: float = -1.0
    done: float = -1.0
    result: object = None     # rendered images or final parameters (keep_results)


class ScenePool:
    """Scenes the requests refer to, kept on the device, with target images
    (renders of a second seed's scene) for optimisation jobs."""

    def __init__(self, scenes: list, target_scenes: list, device):
        self.device = torch.device(device)
        self.scenes = scenes
        self.params = []
        for sc in scenes:
            p = sc.params()
            self.params.append({"means": p["means"].to(device).contiguous(),
                                "sh": torch.cat((p["sh_dc"], p["sh_ac"]), 1).to(device).contiguous(),
                                "opacity": p["opacity"].reshape(-1).to(device).contiguous(),
                                "svec": p["svec"].to(device).contiguous(),
                                "quaternion": p["quaternion"].to(device).contiguous()})
        self.targets = target_scenes
        self.target_buf = None
        self._target_index = {}

    @classmethod
    def from_ply(cls, paths: list, target_paths: list, device) -> "ScenePool":
        """A pool of scenes (and target scenes) read from Gaussian PLY files."""
        return cls([load_scene(p) for p in paths], [load_scene(p) for p in target_paths], device)

    def build_targets(self, requests: list, tile: int = 16):
        """Render each optimisation request's target views once, into one
        device buffer; sets request.target_offs (float offsets)."""
        need = []
        for r in requests:
            if r.kind == "optimize":
                for c in r.cameras:
                    need.append((r, c))
        total = sum(3 * c.width * c.height for _, c in need)
        self.target_buf = torch.zeros(max(total, 1), dtype=torch.float32, device=self.device)
        o = 0
        cache = {}
        for r, c in need:
            if r.scene not in cache:
                m, s, q, op, sh = (t.to(self.device).contiguous() for t in activate(self.targets[r.scene].params()))
                cache = {r.scene: (m, s, q, op, sh)}
            m, s, q, op, sh = cache[r.scene]
            cam = raster.camera_from_tensors(c.viewmatrix, c.projmatrix, c.campos, c.width, c.height,
                                             c.focal_x, c.focal_y, None, device=self.device)
            K, color, *_ = raster.forward_checked(m, s, q, op, sh, self.scenes[r.scene].sh_degree, cam, tile)
            n = color.numel()
            self.target_buf[o:o + n] = color.clamp(0, 1).reshape(-1)
            r.target_offs.append(o)
            o += n


def make_requests(n: int, rate: float, pool_sizes: list, seed: int = 0, opt_frac: float = 0.3,
                  opt_steps: int = 50, resolutions=(256, 512, 1024)) -> list:
    """Poisson arrivals; 70 % renders of 1-8 views, 30 % optimisation jobs."""
    from .camera import orbit_camera
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        scene = int(rng.integers(0, len(pool_sizes)))
        res = int(rng.choice(np.array(resolutions)))
        az0 = float(rng.uniform(0, 360))
        if rng.uniform() < opt_frac:
            el = rng.uniform(-30, 30, 4)
            cams = [orbit_camera(az0 + 90.0 * v, float(el[v]), 2.5, res, res) for v in range(4)]
            out.append(Request(k, "optimize", t, scene, cams, opt_steps))
        else:
            nv = int(rng.integers(1, 9))
            el = rng.uniform(-30, 30, nv)
            cams = [orbit_camera(az0 + 360.0 * v / nv, float(el[v]), 2.5, res, res) for v in range(nv)]
            out.append(Request(k, "render", t, scene, cams, 0))
    return out


def _est_instances(n: int, cams) -> int:
    # rough upper estimate of tile instances for admission (grown on overflow)
    return int(sum(48 * n * max(c.width * c.height / (1024.0 * 1024.0), 0.0625) for c in cams)) + 4096


def workspace_bytes(caps: _lib.W2WCaps, max_jobs: int = 256) -> int:
    """Device bytes of a continuous server's per-batch state for `caps`: the
    batched workspace (geometry, binning, image buffers and table), the
    per-frame outputs and loss gradient, the per-item gradients and the
    job range tables."""
    L = _lib.batch_layout(caps)
    table = _lib.load().w2w_batch_table_bytes(caps.scenes, caps.frames)
    items, pix, G = max(int(caps.items), 1), max(int(caps.pixels), 1), max(int(caps.gaussians), 1)
    per_batch = (L.geom_bytes + L.binning_bytes + L.image_bytes + table
                 + 4 * (3 * pix + pix + pix) + 4 * items      # colour, depth, alpha, radii
                 + 4 * 3 * pix                                 # dL/dcolour
                 + 4 * (6 * items + 3 * G + 6 * G + 3 * G + 4 * G + G)  # BatchGrads
                 + 4 * (max(int(caps.frames), 1) + max(int(caps.tiles), 1)))  # losses
    jobs = 4 * 4 * 5 * max_jobs + max_jobs * ROW_RANGE.itemsize + 5 * max_jobs * PARAM_RANGE.itemsize
    return int(per_batch + jobs)


class ContinuousServer:
    @classmethod
    def for_budget(cls, pool: ScenePool, hbm_bytes: int, caps: _lib.W2WCaps, device, tile: int = 16,
                   max_jobs: int = 256, keep_results: bool = False) -> "ContinuousServer":
        """A server whose per-batch workspace and resident-scene arena
        together fit `hbm_bytes`: the workspace is sized by `caps`, the rest
        becomes arena rows (parameters, gradients, Adam moments and
        activations of resident scenes), so admission by arena rows is
        admission by the HBM budget."""
        M = pool.params[0]["sh"].shape[1]
        ws = workspace_bytes(caps, max_jobs)
        rows = (int(hbm_bytes) - ws) // Arena.bytes_per_row(M)
        if rows < 1:
            raise ValueError(f"an HBM budget of {hbm_bytes} bytes does not hold the batch workspace "
                             f"({ws} bytes for these caps) and any scene")
        return cls(pool, int(rows), caps, device, tile, max_jobs, keep_results)

    def __init__(self, pool: ScenePool, arena_rows: int, caps: _lib.W2WCaps, device, tile: int = 16,
                 max_jobs: int = 256, keep_results: bool = False):
        if int(caps.tile) != tile:
            raise ValueError(f"caps.tile {caps.tile} differs from tile {tile}")
        self.keep_results = keep_results
        self.pool, self.device, self.tile = pool, torch.device(device), tile
        self.M = pool.params[0]["sh"].shape[1]
        self.arena = Arena(arena_rows, self.M, device)
        self.caps = caps
        self.ws = batch.BatchWorkspace(caps, device)
        self.grads = batch.BatchGrads(caps, 0, device)
        pix = int(caps.pixels)
        self.dL_dcolor = torch.zeros(3 * pix, dtype=torch.float32, device=device)
        self.loss = torch.zeros(max(int(caps.frames), 1), dtype=torch.float32, device=device)
        self.tile_loss = torch.zeros(max(int(caps.tiles), 1), dtype=torch.float32, device=device)
        self.max_jobs = max_jobs
        # per job: one row range, five parameter ranges (one per group), each
        # with a 4-float hyper block
        self.hyper = torch.zeros(4 * 5 * max_jobs, dtype=torch.float32, device=device)
        self.rows_dev = torch.zeros(max_jobs * ROW_RANGE.itemsize, dtype=torch.uint8, device=device)
        self.prm_dev = torch.zeros(5 * max_jobs * PARAM_RANGE.itemsize, dtype=torch.uint8, device=device)
        self.k_scale = 1.0
        self.ticks = 0
        self.hbm_bytes = workspace_bytes(caps, max_jobs) + arena_rows * Arena.bytes_per_row(self.M)

    def _fits_alone(self, r: Request) -> str | None:
        """Why request r can never be served by this server (None: it can)."""
        n = self.pool.scenes[r.scene].n
        nv, ts, c = len(r.cameras), self.tile, self.caps
        px = sum(cm.width * cm.height for cm in r.cameras)
        tl = sum(((cm.width + ts - 1) // ts) * ((cm.height + ts - 1) // ts) for cm in r.cameras)
        if n > self.arena.rows:
            return f"its scene has {n} Gaussians, the arena holds {self.arena.rows}"
        for what, need, cap in (("items", n * nv, c.items), ("Gaussians", n, c.gaussians),
                                ("frames", nv, c.frames), ("pixels", px, c.pixels), ("tiles", tl, c.tiles)):
            if need > cap:
                return f"it needs {need} {what}, the workspace holds {cap}"
        return None

    # -- admission --
    def _admit(self, queue: list, active: list, now: float):
        while queue and queue[0].arrival <= now:
            r = queue[0]
            n = self.pool.scenes[r.scene].n
            why = self._fits_alone(r)
            if why is not None:
                raise ValueError(f"request {r.rid} cannot be served: {why}")
            if len(active) >= self.max_jobs:
                break
            base = self.arena.alloc(n)
            if base is None:
                break  # the HBM budget is full: wait for completions
            self.arena.load(base, self.pool.params[r.scene])
            r.base = base
            r.admitted = now
            active.append(queue.pop(0))

    def _batch(self, active: list):
        """Work of this tick that fits the workspace capacities."""
        items = frames = pixels = tiles = 0
        est_k = 0
        chosen = []
        ts = self.tile
        for r in active:
            n = self.pool.scenes[r.scene].n
            nv = len(r.cameras)
            px = sum(c.width * c.height for c in r.cameras)
            tl = sum(((c.width + ts - 1) // ts) * ((c.height + ts - 1) // ts) for c in r.cameras)
            ek = int(_est_instances(n, r.cameras) * self.k_scale)
            if (items + n * nv > self.caps.items or frames + nv > self.caps.frames or
                    pixels + px > self.caps.pixels or tiles + tl > self.caps.tiles or
                    len(chosen) + 1 > self.caps.scenes or (chosen and est_k + ek > self.caps.capacity)):
                continue
            chosen.append(r)
            items += n * nv
            frames += nv
            pixels += px
            tiles += tl
            est_k += ek
        return chosen

    def _run(self, work: list) -> bool:
        """One launch sequence over `work`; False when K overflowed (nothing
        was updated)."""
        L = _lib.load()
        A = self.arena
        s = torch.cuda.current_stream().cuda_stream
        specs, bases, shoffs, toffs = [], [], [], []
        rows = np.zeros(len(work), ROW_RANGE)
        total = 0
        for k, r in enumerate(work):
            sc = self.pool.scenes[r.scene]
            bw = r.kind == "optimize"
            specs.append(batch.SceneSpec(sc.n, sc.sh_degree, self.M,
                                         [batch.frame_from_camera(c, backward=bw, tight=True) for c in r.cameras]))
            bases.append(r.base)
            shoffs.append(A.sh_float_offset(r.base) - int(A.starts[1]))
            toffs.extend(r.target_offs if bw else [0] * len(r.cameras))
            rows[k] = (r.base, total, sc.n)
            total += sc.n
        self.ws.load_table(specs, bases, shoffs, toffs)
        rows_b = rows.view(np.uint8)
        self.rows_dev[:rows_b.size].copy_(torch.from_numpy(rows_b.copy()), non_blocking=False)
        _lib.check(L.w2w_activate_rows(A.group(A.flat, "svec").data_ptr(), A.group(A.flat, "opacity").data_ptr(),
                                       A.group(A.flat, "quaternion").data_ptr(), self.rows_dev.data_ptr(),
                                       len(work), total, A.scales.data_ptr(), A.opac.data_ptr(),
                                       A.rots.data_ptr(), s), "w2w_activate_rows")
        sh = A.group(A.flat, "sh")
        means = A.group(A.flat, "means")
        _lib.check(L.w2w_forward_batch(self.ws.table.data_ptr(), ctypes.byref(self.caps),
                                       ctypes.byref(self.ws.struct), means.data_ptr(), A.scales.data_ptr(),
                                       A.rots.data_ptr(), A.opac.data_ptr(), sh.data_ptr(),
                                       self.ws.color.data_ptr(), self.ws.depth.data_ptr(),
                                       self.ws.alpha.data_ptr(), None, s), "w2w_forward_batch")
        K, ovf = (int(v) for v in self.ws.status().cpu())
        if ovf:
            return False
        opt = [r for r in work if r.kind == "optimize"]
        if opt:
            _lib.check(L.w2w_l2_loss(self.ws.table.data_ptr(), ctypes.byref(self.caps),
                                     self.ws.color.data_ptr(), self.pool.target_buf.data_ptr(),
                                     self.dL_dcolor.data_ptr(), self.loss.data_ptr(),
                                     self.tile_loss.data_ptr(), s), "w2w_l2_loss")
            g = self.grads
            _lib.check(L.w2w_backward_batch(
                self.ws.table.data_ptr(), ctypes.byref(self.caps), ctypes.byref(self.ws.struct),
                means.data_ptr(), A.scales.data_ptr(), A.rots.data_ptr(), sh.data_ptr(),
                self.dL_dcolor.data_ptr(), None, None, g.means2D.data_ptr(), g.colors.data_ptr(),
                A.g_opac.data_ptr(), A.group(A.grad, "means").data_ptr(), A.g_cov.data_ptr(),
                A.group(A.grad, "sh").data_ptr(), A.g_scales.data_ptr(), A.g_rots.data_ptr(), s),
                "w2w_backward_batch")
            _lib.check(L.w2w_activate_backward_rows(
                A.scales.data_ptr(), A.opac.data_ptr(), A.group(A.flat, "quaternion").data_ptr(),
                A.g_scales.data_ptr(), A.g_opac.data_ptr(), A.g_rots.data_ptr(), self.rows_dev.data_ptr(),
                len(work), total, A.group(A.grad, "svec").data_ptr(), A.group(A.grad, "opacity").data_ptr(),
                A.group(A.grad, "quaternion").data_ptr(), s), "w2w_activate_backward_rows")
            # one Adam launch over every job's five groups
            prm = np.zeros(5 * len(opt), PARAM_RANGE)
            hyp = np.zeros(4 * 5 * len(opt), np.float32)
            off = 0
            for j, r in enumerate(opt):
                t = r.steps_done + 1
                bc1, bc2 = 1.0 - 0.9 ** t, 1.0 - 0.999 ** t
                lrs = {"means": mean_lr(t), "sh": LRS["sh_dc"], "opacity": LRS["opacity"],
                       "svec": LRS["svec"], "quaternion": LRS["quaternion"]}
                n = self.pool.scenes[r.scene].n
                for gk, name in enumerate(GROUPS):
                    w = A.width(name)
                    hidx = 5 * j + gk
                    start = int(A.starts[gk]) + r.base * w
                    period, split = (3 * self.M, 3) if name == "sh" else (0, 0)
                    prm[5 * j + gk] = (start, off, n * w, hidx, period, split, 0)
                    off += n * w
                    hyp[4 * hidx:4 * hidx + 3] = (-lrs[name] / bc1,
                                                  -LRS["sh_ac"] / bc1 if name == "sh" else 0.0,
                                                  1.0 / np.sqrt(bc2))
            self.hyper[:hyp.size].copy_(torch.from_numpy(hyp), non_blocking=False)
            pb = prm.view(np.uint8)
            self.prm_dev[:pb.size].copy_(torch.from_numpy(pb.copy()), non_blocking=False)
            _lib.check(L.w2w_adam_ranges(A.flat.data_ptr(), A.grad.data_ptr(), A.exp_avg.data_ptr(),
                                         A.exp_avg_sq.data_ptr(), self.prm_dev.data_ptr(), len(prm),
                                         off, self.hyper.data_ptr(), 0.9, 0.999, 1e-15, s),
                       "w2w_adam_ranges")
        return True

    def _collect(self, work: list):
        views = self.ws.frame_views([batch.SceneSpec(self.pool.scenes[r.scene].n, 0, self.M,
                                                     [batch.frame_from_camera(c) for c in r.cameras])
                                     for r in work])
        k = 0
        for r in work:
            nv = len(r.cameras)
            if r.kind == "render":
                r.result = [v[0].detach().cpu().clone() for v in views[k:k + nv]]
            elif r.steps_done + 1 >= r.steps:
                n = self.pool.scenes[r.scene].n
                A = self.arena
                r.result = {name: A.group(A.flat, name)[r.base * A.width(name):(r.base + n) * A.width(name)]
                            .detach().cpu().clone() for name in GROUPS}
            k += nv

    def serve(self, requests: list, log=None) -> list:
        """Run the stream in real time until every request completed."""
        queue = sorted(requests, key=lambda r: r.arrival)
        active, finished = [], []
        t0 = time.perf_counter()
        while queue or active:
            now = time.perf_counter() - t0
            self._admit(queue, active, now)
            if not active:
                # idle until the next arrival
                time.sleep(max(0.0, min(queue[0].arrival - now, 0.01)))
                continue
            work = self._batch(active) or active[:1]  # a lone request always fits the caps
            if not self._run(work):
                if len(work) == 1:
                    r = work[0]
                    raise RuntimeError(f"request {r.rid} alone needs more tile instances than the "
                                       f"workspace capacity ({self.caps.capacity})")
                self.k_scale *= 1.5  # fewer requests per batch from now on
                continue
            torch.cuda.synchronize()
            self.ticks += 1
            now = time.perf_counter() - t0
            if self.keep_results:
                self._collect(work)
            for r in work:
                if r.kind == "render":
                    r.done = now
                else:
                    r.steps_done += 1
                    if r.steps_done >= r.steps:
                        r.done = now
                if r.done >= 0:
                    active.remove(r)
                    self.arena.free(r.base, self.pool.scenes[r.scene].n)
                    finished.append(r)
            if log and self.ticks % 20 == 0:
                log(f"tick {self.ticks}: t={now:.2f}s active={len(active)} queued={len(queue)} "
                    f"finished={len(finished)}")
        return finished


class PlainServer:
    """One request at a time, FIFO, on the single-view path (the reference's
    way: GaussianTrainer renders and steps one scene, one view at a time)."""

    def __init__(self, pool: ScenePool, device, tile: int = 16):
        self.pool, self.device, self.tile = pool, torch.device(device), tile

    def serve(self, requests: list, log=None) -> list:
        queue = sorted(requests, key=lambda r: r.arrival)
        finished = []
        t0 = time.perf_counter()
        for r in queue:
            now = time.perf_counter() - t0
            if now < r.arrival:
                time.sleep(r.arrival - now)
            r.admitted = time.perf_counter() - t0
            sc = self.pool.scenes[r.scene]
            cams = [raster.camera_from_tensors(c.viewmatrix, c.projmatrix, c.campos, c.width, c.height,
                                               c.focal_x, c.focal_y, None, device=self.device)
                    for c in r.cameras]
            if r.kind == "render":
                m, s, q, o, sh = (t.to(self.device).contiguous() for t in activate(sc.params()))
                for cam in cams:
                    raster.forward_checked(m, s, q, o, sh, sc.sh_degree, cam, self.tile)
            else:
                tg = [self.pool.target_buf[o:o + 3 * c.width * c.height].view(3, c.height, c.width)
                      for o, c in zip(r.target_offs, r.cameras)]
                tr = PlainTrainer(sc, cams, tg, TrainConfig(sh_degree=sc.sh_degree, tile=self.tile),
                                  self.device)
                for _ in range(r.steps):
                    tr.step()
            torch.cuda.synchronize()
            r.done = time.perf_counter() - t0
            finished.append(r)
            if log and len(finished) % 20 == 0:
                log(f"plain: {len(finished)}/{len(queue)} t={r.done:.2f}s")
        return finished


def save_result_ply(path: str, result: dict, M: int) -> None:
    """Writes an optimisation request's final parameters (`Request.result` of
    a ContinuousServer with keep_results) as a Gaussian PLY."""
    n = result["means"].numel() // 3
    sh = result["sh"].reshape(n, M, 3)
    save_ply(path, result["means"].reshape(n, 3), sh[:, :1], sh[:, 1:], result["opacity"].reshape(n),
             result["svec"].reshape(n, 3), result["quaternion"].reshape(n, 4))


def latency_stats(finished: list) -> dict:
    lat = np.array([r.done - r.arrival for r in finished])
    span = max(r.done for r in finished) - min(r.arrival for r in finished)
    views = sum(len(r.cameras) * (r.steps if r.kind == "optimize" else 1) for r in finished)
    return {"requests": len(finished), "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
            "makespan_s": float(span), "requests_per_s": len(finished) / span,
            "views_per_s": views / span}
