#!/usr/bin/env python
"""Benchmark: fwd+bwd views/sec of the HIP Gaussian rasterizer.

Default workload (the large-scene one-GPU configuration): 1,000,000
Gaussians, SH degree 3, FP32, 4 orbital views per step at 1024x1024 with
elevations drawn uniform in [-30, 30] degrees (radius 2.5, fovy 49.1), 16x16
tiles.  One step = render the 4 views, L2 loss against images rendered from a
second seed's scene, backward, Adam (eps 1e-15, the trainer's per-group
learning rates), all through the batched trainer (one launch per stage over
the step's views).  Inputs are synthetic, generated on the CPU from fixed
seeds and resident in HBM before timing starts.  The drop-in per-view path
(autograd Function + torch Adam) is timed beside it.

Other workloads: --workload batched64 (64 scenes, batched vs plain) and
--workload served (continuous batching vs a FIFO plain server).

Prints ONE JSON line.  `value` = views/s over the timed steps; `roofline` is
the dominant kernel, timed live with HIP events on its launch stream;
`cpu_baseline` is the CPU oracle (C, one thread per view) on the CPU-path
configuration (4,096 Gaussians, SH 0, 4 views at 256^2), median of five
steps, with the GPU's rate on that configuration beside it and a 1M-scene
sample scaled to the full scene (labelled extrapolated).  The default line
also carries the 64-scene batch (`batched64`: batched vs plain, step
latency) and the served stream (`served`: continuous batching vs a FIFO
plain server, p50/p99 latency).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
METRIC = "fwd+bwd views/sec at 1M Gaussians 1024^2; p50/p99 latency, served vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (scene1m: up to and including the first densification "
                         "cycle; batched64: 10)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--gaussians", type=int, default=1_000_000)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--views", type=int, default=4)
    ap.add_argument("--tile", type=int, default=0,
                    help="tile edge 8, 16 or 32; 0 = pick tile and chunk by the autotuner")
    ap.add_argument("--chunk", type=int, default=0,
                    help="instances per forward-blend chunk, 64, 128 or 256 (with --tile given; "
                         "0 = 64)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-gaussians", type=int, default=250_000)
    ap.add_argument("--workload", choices=["scene1m", "batched64", "served", "voxel"], default="scene1m")
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--requests", type=int, default=120)
    ap.add_argument("--rate", type=float, default=40.0)
    ap.add_argument("--opt-steps", type=int, default=50)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--densify-every", type=int, default=100,
                    help="scene1m: densification cycle every this many steps (0 = never)")
    ap.add_argument("--depth-split", type=float, default=-1.0,
                    help="depth-sliced binning: fraction of the items (nearest first) binned and "
                         "blended first, the rest only on tiles with unsaturated pixels (same "
                         "outputs); 0 = one binning pass, negative = pick by timing a few steps")
    ap.add_argument("--no-extra-workloads", action="store_true",
                    help="scene1m: skip the 64-scene and served legs that the default line carries")
    ap.add_argument("--cpu-steps", type=int, default=5,
                    help="steps of the CPU-path configuration the CPU baseline takes the median of")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="after the benchmark, run a plain device copy of this many bytes three "
                         "times (a reference kernel for hardware byte counters)")
    return ap.parse_args()


HBM_STAGES_NULL = ("fill",)  # buffer clears: no algorithmic bytes
SPLIT_CANDIDATES = (0.0, 0.2, 0.25, 0.3, 0.35, 0.4)


def stage_cost(name, P, M, V, K, T, HW, stats_view, tpasses, dpasses=4, fused_sh=True):
    """Algorithmic bytes (or flops) of a stage over a whole step (all its
    launches): V views, K tile instances sorted (over both slices when the
    binning is depth-sliced), T tiles and HW pixels summed over the views,
    tile sorts of `tpasses` passes.  Adam, activation, the SH backward and
    the statistics are per Gaussian or per item.  fused_sh: the SH backward
    takes the SH coefficients' Adam step (their parameter and both moments
    read and written, no gradient), which the Adam stage then leaves out."""
    p_in = 44 + 12 * M
    I = V * P  # items
    table = {
        "activate": ("hbm", 64 * P),            # svec, quaternion, opacity in; activated out
        "activate_bwd": ("hbm", 96 * P),        # three gradients and the parameters in, gradients out
        "preprocess": ("hbm", I * (p_in + 49)),
        "depth_sort_hist": ("hbm", 4 * I),
        "depth_sort_pass": ("hbm", 16 * I * dpasses),
        "scan": ("hbm", 12 * I),
        "emit": ("hbm", 20 * I + 8 * K),
        "tile_sort_hist": ("hbm", None),        # clears look-back words; emission built the histogram
        "tile_sort_pass": ("hbm", 16 * K * tpasses),
        "ranges": ("hbm", 4 * K + 8 * T),
        "loss": ("hbm", 36 * HW),               # render and target in, dL/dcolor out
        "preprocess_bwd": ("hbm", I * (2 * p_in + 36)),
        "sh_bwd": ("hbm", P * ((72 if fused_sh else 24) * M + 24) + 12 * I),
        "densify": ("hbm", 16 * I + 16 * P),
        "adam": ("hbm", 28 * (11 + (0 if fused_sh else 3 * M)) * P),
    }
    if name == "blend":
        return "valu", V * (18 * stats_view[0] + 12 * stats_view[1])
    if name == "blend_bwd":
        return "valu", V * (18 * stats_view[3] + 62 * stats_view[4])
    return table.get(name, ("hbm", None))


# kernels of the far slice (depth-sliced binning) belong to these stages
STAGE_OF = {"blend_far": "blend", "emit_far": "emit", "count_open": "emit"}


def stage_table(per, counts, P, M, V, K, T, HW, stats_view, tpasses):
    """Per-stage roofline of one step and the path ratio sum(ideal) /
    sum(measured), ideal = bytes / HBM peak or flops / FP32 peak."""
    folded, fcounts = {}, {}
    for name, ms in per.items():
        st = STAGE_OF.get(name, name)
        folded[st] = folded.get(st, 0.0) + ms
        fcounts[st] = fcounts.get(st, 0) + counts[name]
    per, counts = folded, fcounts
    stages, ideal, measured = {}, 0.0, 0.0
    for name, ms in sorted(per.items(), key=lambda kv: -kv[1]):
        bound, work = stage_cost(name, P, M, V, K, T, HW, stats_view, tpasses)
        measured += ms
        row = {"ms": ms, "launches": counts[name], "bound": bound}
        if work:
            peak, unit, scale = ((FP32_PEAK_TFLOPS, "TFLOP/s", 1e12) if bound == "valu"
                                 else (HBM_PEAK_GBS, "GB/s", 1e9))
            achieved = work / (ms * 1e-3) / scale
            ideal += work / (peak * scale) * 1e3
            row.update({"achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak})
        stages[name] = row
    return stages, {"ideal_ms": ideal, "measured_ms": measured, "frac": ideal / measured if measured else None}


def main():
    args = parse()
    if args.tile not in (0, 8, 16, 32) or args.chunk not in (0, 64, 128, 256):
        print("--tile must be 0 (autotune), 8, 16 or 32; --chunk 0, 64, 128 or 256", file=sys.stderr)
        sys.exit(2)
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); the rasterizer's "
              "multi-GPU sharding is a design note only", file=sys.stderr)
        sys.exit(2)
    if args.workload == "voxel":
        return bench_voxel(args)
    if args.workload in ("batched64", "served"):
        args.steps = 10 if args.steps is None else args.steps
        return bench_batched(args) if args.workload == "batched64" else bench_served(args)
    return bench_scene(args)


def plain_step_fn(params, cams, targets, P, ts, dev):
    """The drop-in path: the reference trainer's loop (one view per launch
    sequence through the autograd Function, torch Adam)."""
    import torch
    from word2world_amd import raster
    from word2world_amd.scene import activate
    lrs = {"means": 0.01, "sh_dc": 0.01, "sh_ac": 0.0005, "opacity": 0.05, "svec": 0.005,
           "quaternion": 0.005}  # optimizer_config.json
    opt = torch.optim.Adam([{"params": [params[k]], "lr": lr, "name": k} for k, lr in lrs.items()],
                           lr=0.0, eps=1e-15)

    def step(i):
        for g in opt.param_groups:
            if g["name"] == "means":
                g["lr"] = mean_lr(i)
        opt.zero_grad(set_to_none=True)
        m, s, r, o, sh = activate(params)
        sh = sh.contiguous()
        loss = 0.0
        for v, cam in enumerate(cams):
            m2d = torch.zeros(P, 3, device=dev, requires_grad=True)
            img, radii, depth, alpha = raster.rasterize(m, m2d, sh, o, s, r, cam, 3, ts)
            loss = loss + ((img.clamp(0, 1) - targets[v]) ** 2).mean()
        loss.backward()
        opt.step()
        return loss
    return step


def mean_lr(step):  # misc_utils.py:11-27 with (0.01, 0.0002, 300)
    t = min(max(step / 300, 0.0), 1.0)
    return float(np.exp(np.log(0.01) * (1 - t) + np.log(0.0002) * t))


def bench_scene(args):
    """1M Gaussians, 4 views at 1024^2 per step: fwd+bwd+Adam through the
    batched trainer (all views of the step in one launch per stage, flat
    parameter store, fused Adam), with the drop-in per-view path timed beside
    it."""
    import torch

    from word2world_amd import _lib, raster
    from word2world_amd.batched import BatchedTrainer
    from word2world_amd.camera import orbit_camera
    from word2world_amd.scene import activate, make_scene

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    L = _lib.load()
    P, R, V, ts = args.gaussians, args.res, args.views, args.tile
    rng = np.random.default_rng(0)
    elev = rng.uniform(-30.0, 30.0, V)
    cams_host = [orbit_camera(k * 360.0 / V, float(elev[k]), 2.5, R, R) for k in range(V)]
    cams = [raster.camera_from_tensors(c.viewmatrix, c.projmatrix, c.campos, R, R, c.focal_x,
                                       c.focal_y, [1.0, 1.0, 1.0], device=dev) for c in cams_host]
    log(f"scene: {P} Gaussians SH3, {V} views at {R}^2, tile {ts or 'auto'}")
    scene = make_scene(P, 3, seed=0)
    target_scene = make_scene(P, 3, seed=1)
    M = 16
    with torch.no_grad():
        tm, ts_, tr, to, tsh = (t.to(dev).contiguous() for t in activate(target_scene.params()))
        targets = []
        for cam in cams:
            K, color, _, _, _, _ = raster.forward_checked(tm, ts_, tr, to, tsh, 3, cam, 16)
            targets.append(color.clamp(0, 1))
        del tm, ts_, tr, to, tsh
    torch.cuda.empty_cache()

    tune, chunk = None, args.chunk
    if args.tile == 0:
        # time the trainer step itself (all views in one launch per stage, at
        # the middle depth-split candidate): a single view's fwd+bwd ranks the
        # chunks differently from the four-view launch
        from word2world_amd import autotune
        split0 = args.depth_split if args.depth_split >= 0 else 0.3
        t = autotune.sweep_trainer(lambda ts_, ch: BatchedTrainer([scene], [cams_host], [targets], dev, ts_,
                                                                  depth_split=split0, chunk=ch),
                                   warmup=10, steps=10)
        ts, chunk = min(t, key=t.get)
        tune = {autotune.label(k): v for k, v in t.items()}
        log("autotune (ms per trainer step): " + ", ".join(f"{k}: {v:.2f}" for k, v in tune.items())
            + f" -> tile {ts}, chunk {chunk}")
        torch.cuda.empty_cache()
    split_ms = None
    if args.depth_split < 0:
        # depth split: time a few trainer steps per candidate (the outputs are
        # the same for every choice)
        split_ms = {}
        for frac in SPLIT_CANDIDATES:
            t = BatchedTrainer([scene], [cams_host], [targets], dev, ts, depth_split=frac, chunk=chunk)
            for _ in range(30):  # past the first steps, where the scene changes fastest
                t.step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(10):
                t.step()
            torch.cuda.synchronize()
            split_ms[frac] = (time.perf_counter() - t0) / 10 * 1e3
            del t
            torch.cuda.empty_cache()
        args.depth_split = min(split_ms, key=split_ms.get)
        log("depth split (ms per step): " + ", ".join(f"{k}: {v:.3f}" for k, v in split_ms.items())
            + f" -> {args.depth_split}")
    bt = BatchedTrainer([scene], [cams_host], [targets], dev, ts, densify_every=args.densify_every, chunk=chunk,
                        densify_grad=2e-4, densify_min_opacity=0.005,
                        generator=torch.Generator(device=dev).manual_seed(0), depth_split=args.depth_split)
    if args.steps is None:
        # end the timed region with the first densification cycle (warmup
        # and the profiled step come first)
        args.steps = max(1, args.densify_every - args.warmup - 1) if args.densify_every else 20
    for i in range(args.warmup):
        bt.step()
        torch.cuda.synchronize()
        K, ovf = (int(v) for v in bt.ws.status().cpu())
        log(f"warmup step {i}: K={K} ({K / V:.0f} per view) overflow={ovf}")
        if ovf:
            raise RuntimeError("instance capacity overflowed")

    # which kernel dominates: one profiled step over every kernel id
    L.w2w_prof_enable(4096, 0xFFFFFFFF)
    L.w2w_prof_reset()
    bt.step()
    torch.cuda.synchronize()
    per = {}
    for name, ms in _lib.prof_read():
        per[name] = per.get(name, 0.0) + ms
    L.w2w_prof_enable(0, 0)
    dominant = max(per, key=per.get)
    log("per-kernel ms in one step: " + ", ".join(f"{k}={v:.3f}" for k, v in
                                                  sorted(per.items(), key=lambda kv: -kv[1])))
    kid = {v: k for k, v in _lib.KERNEL_IDS.items()}[dominant]

    # timed region: events only around the dominant kernel's launches
    L.w2w_prof_enable(8192, 1 << kid)
    L.w2w_prof_reset()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record()
        bt.step()
        ev[i][1].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if bt.overflow() or bt.skipped_overflow() or bt.grow_log:
        # any overflowed step inside the timed region did incomplete work
        raise RuntimeError("instance capacity overflowed inside the timed region")
    dens = {"every": args.densify_every, "cycles_timed": sum(1 for it, _ in bt.densify_log
                                                             if it > bt.it - args.steps),
            "counts": [(it, c[0]) for it, c in bt.densify_log], "gaussians_after": int(bt.store.G)}
    log(f"densification: {dens}")
    step_ms = np.array([a.elapsed_time(b) for a, b in ev])
    dom = [ms for name, ms in _lib.prof_read() if name == dominant]
    L.w2w_prof_enable(0, 0)
    wall = t1 - t0
    views_per_s = args.steps * V / wall
    log(f"timed {args.steps} steps: {wall * 1e3 / args.steps:.2f} ms/step, {views_per_s:.1f} views/s")

    # work counters for the roofline flops: one extra, untimed step right
    # after the timed region (the scene as the last timed steps saw it)
    stats = torch.zeros(6, dtype=torch.int64, device=dev)
    L.w2w_stats_enable(stats.data_ptr())
    bt.step()
    torch.cuda.synchronize()
    L.w2w_stats_enable(None)
    stats_np = stats.cpu().numpy().astype(np.float64) / V  # per view
    K_total = int(bt.ws.status()[0])

    # every stage of one more (untimed) step after the timed region, where
    # the scene has trained as far as the timed steps
    L.w2w_prof_enable(4096, 0xFFFFFFFF)
    L.w2w_prof_reset()
    bt.step()
    torch.cuda.synchronize()
    per_end, cnt_end = {}, {}
    for name, ms in _lib.prof_read():
        per_end[name] = per_end.get(name, 0.0) + ms
        cnt_end[name] = cnt_end.get(name, 0) + 1
    L.w2w_prof_enable(0, 0)
    K_end = int(bt.ws.status()[0])
    _, KA, KB, _, KH, KN = (int(v) for v in bt.ws.sliced_counts().cpu())
    K_sorted = KN + KB if bt.caps.depth_split else K_end
    T_step = V * ((R + ts - 1) // ts) ** 2
    tpasses = (max(1, int(np.ceil(np.log2(T_step)))) + 7) // 8
    stages, path = stage_table(per_end, cnt_end, int(bt.store.G), M, V, K_sorted, T_step, V * R * R, stats_np,
                               tpasses)
    log(f"path: ideal {path['ideal_ms']:.3f} ms of measured {path['measured_ms']:.3f} ms per step")

    # roofline of the dominant kernel (one launch covers all V views)
    K_view = K_total / V
    bound, work_step = stage_cost(dominant, int(bt.store.G), M, V, K_sorted, T_step, V * R * R, stats_np, tpasses)
    launches_per_step = max(1, round(len(dom) / args.steps))
    work_per_launch = (work_step or 0.0) / launches_per_step
    avg_ms = float(np.mean(dom)) if dom else float("nan")
    if bound == "valu":
        achieved = work_per_launch / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None}
    else:
        achieved = work_per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    roof.update({"kernel": dominant, "avg_launch_ms": avg_ms, "launches": len(dom),
                 "algorithmic_per_launch": work_per_launch})
    # HBM bytes per launch of this kernel, from the FETCH_SIZE / WRITE_SIZE
    # counter runs of this same command (scaled by a copy kernel of known size)
    tpath = os.path.join(ROOT, "profiles", f"{dominant}_scene1m_traffic.json")
    if os.path.exists(tpath) and args.gaussians == 1_000_000 and args.res == 1024 and V == 4:
        t = json.load(open(tpath))
        if t.get("kernel") == dominant and "bytes_per_launch" in t:
            roof["traffic"] = t["bytes_per_launch"]
            roof["traffic_source"] = os.path.relpath(tpath, ROOT)
    del bt
    torch.cuda.empty_cache()

    # the drop-in per-view path, for comparison
    plain = None
    if args.plain_steps > 0:
        params = {k: v.to(dev).requires_grad_(True) for k, v in scene.params().items()}
        step = plain_step_fn(params, cams, targets, P, 16, dev)  # the reference's tile edge
        step(0)
        torch.cuda.synchronize()
        tp0 = time.perf_counter()
        for i in range(args.plain_steps):
            step(1 + i)
        torch.cuda.synchronize()
        tp = (time.perf_counter() - tp0) / args.plain_steps
        plain = {"views_per_s": V / tp, "ms_per_step": tp * 1e3, "steps": args.plain_steps,
                 "path": "drop-in rasterize_gaussians autograd Function, one view per launch "
                         "sequence, torch.optim.Adam, tile 16"}
        log(f"plain drop-in path: {tp * 1e3:.2f} ms/step, {V / tp:.1f} views/s")
        del params
        torch.cuda.empty_cache()

    if args.copy_bytes > 0:
        src = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)  # an elementwise kernel: reads and writes copy_bytes
        torch.cuda.synchronize()
        del src, dst

    cpu = None
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_small(args, dev)
        cpu["extrapolated_1m"] = cpu_baseline(args, cams_host, P)

    # the 64-scene batch and the served stream, as the metric names them
    extra = {}
    if not args.no_extra_workloads:
        import copy
        torch.cuda.empty_cache()
        for name, fn, kw in (("batched64", bench_batched, dict(depth_split=0.0)),
                             ("served", bench_served, {})):
            a = copy.copy(args)
            a.tile, a.chunk, a.steps, a.warmup = 0, 0, 10, 3
            for k, v in kw.items():
                setattr(a, k, v)
            r = fn(a, emit=False)
            extra[name] = {k: r[k] for k in r if k not in ("metric", "n_gpus", "higher_is_better", "scaling",
                                                          "vs_baseline", "dtype", "data")}
            torch.cuda.empty_cache()

    out = {
        "metric": METRIC, "value": views_per_s, "unit": "views/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{P} Gaussians SH3 (seeded synthetic), {V} orbital views at {R}x{R} "
                               "(elevation U[-30,30]), fwd+bwd+Adam per step, L2 to a second "
                               "seed's renders; all views of a step in one launch per stage",
                   "gaussians": P, "views_per_step": V, "resolution": R, "tile": ts, "chunk": chunk or 64,
                   "densify_every": args.densify_every},
        "densify": dens,
        "autotune_ms": tune,
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)), "p99": float(np.percentile(step_ms, 99))},
        "per_view": {"K": K_view, "evaluated_pairs_fwd": stats_np[0], "contributing_pairs_fwd": stats_np[1],
                     "instances_loaded_fwd": stats_np[2], "evaluated_pairs_bwd": stats_np[3],
                     "contributing_pairs_bwd": stats_np[4], "instances_loaded_bwd": stats_np[5]},
        "binning": {"depth_split": args.depth_split, "instances_sorted_per_view": K_sorted / V,
                    "near_per_view": KA / V, "hinted_far_per_view": KH / V, "far_per_view": KB / V,
                    "split_ms_per_step": split_ms},
        "kernel_ms_one_step": per,
        "kernel_ms_last_step": per_end,
        "stages": stages,
        "path_roofline": path,
        "plain": plain,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


def _targets(jobs, dev, tile):
    import torch
    from word2world_amd import raster
    from word2world_amd.scene import activate
    out = []
    with torch.no_grad():
        for j in jobs:
            m, s, r, o, sh = (t.to(dev).contiguous() for t in activate(j.target_scene.params()))
            tl = []
            for c in j.cameras:
                cam = raster.camera_from_tensors(c.viewmatrix, c.projmatrix, c.campos, c.width, c.height,
                                                 c.focal_x, c.focal_y, None, device=dev)
                tl.append(raster.forward_checked(m, s, r, o, sh, j.target_scene.sh_degree, cam,
                                                 tile)[1].clamp(0, 1))
            out.append(tl)
    return out


def bench_batched(args, emit: bool = True):
    """64 scenes (10k-500k Gaussians, log-uniform), 4 views each at a
    resolution drawn from {256,512,1024}^2, fwd+bwd+Adam per step: the batched
    trainer (one launch sequence, HIP-graph replay) against the plain path
    (one scene and one view at a time)."""
    import torch
    from word2world_amd import _lib, raster
    from word2world_amd.batched import BatchedTrainer
    from word2world_amd.train import PlainTrainer, TrainConfig
    from word2world_amd.workloads import batched_64

    dev = torch.device("cuda:0")
    jobs = batched_64(args.scenes)
    G = sum(j.scene.n for j in jobs)
    V = sum(len(j.cameras) for j in jobs)
    log(f"batched: {len(jobs)} scenes, {G} Gaussians, {V} views")
    targets = _targets(jobs, dev, 16)
    split = max(args.depth_split, 0.0)  # this workload is not swept: 0 unless given
    scenes, cameras = [j.scene for j in jobs], [j.cameras for j in jobs]
    tile, chunk, tune = args.tile, args.chunk, None
    if tile == 0:
        # one (tile, chunk) serves every scene of the batch: time the trainer itself
        from word2world_amd import autotune
        t = autotune.sweep_trainer(lambda ts, ch: BatchedTrainer(scenes, cameras, targets, dev, ts,
                                                                 depth_split=split, chunk=ch),
                                   warmup=2, steps=3)
        tile, chunk = min(t, key=t.get)
        tune = {autotune.label(k): v for k, v in t.items()}
        log("batched autotune (ms per step): " + ", ".join(f"{k}: {v:.1f}" for k, v in tune.items())
            + f" -> tile {tile}, chunk {chunk}")
    bt = BatchedTrainer(scenes, cameras, targets, dev, tile, depth_split=split, chunk=chunk)
    for i in range(max(args.warmup, 1)):
        bt.step()
        torch.cuda.synchronize()
        K, ovf = (int(v) for v in bt.ws.status().cpu())
        log(f"warmup {i}: K={K} overflow={ovf}")
        if ovf:
            raise RuntimeError("batched workspace capacity too small")
    if not args.no_graph:
        bt.capture()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record()
        bt.step()
        ev[i][1].record()
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / args.steps
    lat_b = np.array([a.elapsed_time(b) for a, b in ev])
    if int(bt.ws.status()[1]):
        raise RuntimeError("overflow in the timed region")
    log(f"batched: {tb * 1e3:.1f} ms/step, {V / tb:.1f} views/s")
    del bt
    torch.cuda.empty_cache()
    # plain: one scene at a time, one view at a time
    cams = [[raster.camera_from_tensors(c.viewmatrix, c.projmatrix, c.campos, c.width, c.height, c.focal_x,
                                        c.focal_y, None, device=dev) for c in j.cameras] for j in jobs]
    plains = [PlainTrainer(j.scene, cs, t, TrainConfig(tile=tile), dev)
              for j, cs, t in zip(jobs, cams, targets)]
    for p in plains:
        # a step whose workspace overflowed sized the next one; its update is
        # kept (this is warm-up, not a timed step)
        p.step()
        if p.observe_capacity():
            p.step()
            if p.observe_capacity():
                raise RuntimeError("plain trainer workspace overflowed twice")
    torch.cuda.synchronize()
    psteps = max(1, min(args.steps, 3))
    plain_ms = []
    t0 = time.perf_counter()
    for _ in range(psteps):
        t_s = time.perf_counter()
        for p in plains:
            p.step()
        torch.cuda.synchronize()
        plain_ms.append((time.perf_counter() - t_s) * 1e3)
    tp = (time.perf_counter() - t0) / psteps
    log(f"plain: {tp * 1e3:.1f} ms/step, {V / tp:.1f} views/s")
    plain_ms = np.array(plain_ms)
    out = {"metric": "fwd+bwd views/sec, 64-scene mixed batch (batched vs plain)", "value": V / tb,
           "unit": "views/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": tb * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"{len(jobs)} scenes, Gaussians log-uniform [10k,500k] (total {G}), SH3, "
                                  "4 views each at {256,512,1024}^2, fwd+bwd+Adam",
                      "graph": not args.no_graph, "tile": tile, "chunk": chunk or 64, "depth_split": split},
           "autotune_ms": tune,
           "plain_views_per_s": V / tp, "speedup_vs_plain": tp / tb,
           "latency_ms": {"p50": float(np.percentile(lat_b, 50)), "p99": float(np.percentile(lat_b, 99))},
           "plain_latency_ms": {"p50": float(np.percentile(plain_ms, 50)),
                                "p99": float(np.percentile(plain_ms, 99))}}
    del plains
    torch.cuda.empty_cache()
    if emit:
        print(json.dumps(out), flush=True)
    return out


def bench_voxel(args):
    """Occupancy field of the export path (extract_fields defaults,
    utils_for_mesh.py:146: 128^3 grid, 16 blocks, relax 1.5) for the 1M-
    Gaussian scene, against the CPU oracle (the reference's per-point loop over
    every Gaussian) on a bounded sample."""
    import torch
    from word2world_amd.scene import make_scene
    from word2world_amd.voxel import prepare_gaussians_for_extraction, voxel_field

    dev = torch.device("cuda:0")
    res, nb, relax = 128, 16, 1.5
    sc = make_scene(args.gaussians, 3, seed=0)
    m, covs, op, _, _ = prepare_gaussians_for_extraction(sc.means.to(dev), torch.exp(sc.svec).to(dev),
                                                         sc.quaternion.to(dev), torch.sigmoid(sc.opacity).to(dev),
                                                         0.005)
    for _ in range(args.warmup):
        voxel_field(m, covs, op, res, nb, relax)
    torch.cuda.synchronize()
    steps = args.steps or 5
    t0 = time.perf_counter()
    for _ in range(steps):
        occ = voxel_field(m, covs, op, res, nb, relax)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    log(f"voxel field {res}^3 over {int(m.shape[0])} Gaussians: {ms:.2f} ms")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, ROOT)
        from oracle import oracle
        k = 4000  # Gaussians of the sample; 32^3 points
        t1 = time.perf_counter()
        oracle.voxel_field(m[:k].cpu().numpy(), covs[:k].cpu().numpy(), op[:k].cpu().numpy(), 32, nb, relax)
        dt = time.perf_counter() - t1
        # the oracle's work is points x Gaussians: scale to the full field
        full_s = dt * (res ** 3 / 32 ** 3) * (int(m.shape[0]) / k)
        cpu = {"value": 1.0 / full_s, "unit": "fields/s", "cores": 1, "kind": "port",
               "sample": f"32^3 points x the first {k} Gaussians, {dt:.2f} s, scaled by points x Gaussians"}
    out = {"metric": "occupancy fields/s (128^3, 16 blocks, relax 1.5)", "value": 1e3 / ms, "unit": "fields/s",
           "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"occupancy field of {int(m.shape[0])} Gaussians on a {res}^3 grid",
                      "resolution": res, "num_blocks": nb, "relax_ratio": relax},
           "occ_max": float(occ.max()), "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_served(args, emit: bool = True):
    """Poisson request stream: 70 % renders (1-8 views), 30 % optimisation jobs
    (fwd+bwd+Adam steps on a resident scene), scenes log-uniform [5k, 1M];
    continuous batching against the plain FIFO server."""
    import copy
    import torch
    from word2world_amd import serve
    from word2world_amd.scene import make_scene
    from word2world_amd.workloads import log_uniform_counts

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(0)
    sizes = [int(n) for n in log_uniform_counts(16, 5_000, 1_000_000, rng)]
    pool = serve.ScenePool([make_scene(n, 3, 500 + k) for k, n in enumerate(sizes)],
                           [make_scene(n, 3, 600 + k) for k, n in enumerate(sizes)], dev)
    reqs = serve.make_requests(args.requests, args.rate, sizes, seed=1, opt_steps=args.opt_steps)
    pool.build_targets(reqs)
    tile, chunk, tune = args.tile, args.chunk, None
    if tile == 0:
        # requests of every size and resolution share the server's choice,
        # and the server runs them as one batch per launch: tune on a batch
        # like the ones it forms (the eight largest pool scenes, 4 views each
        # at 256, 512 and 1024^2 in turn; fwd+bwd+Adam steps)
        from word2world_amd import autotune
        from word2world_amd.batched import BatchedTrainer
        from word2world_amd.workloads import SceneJob, orbit4
        pick = [int(k) for k in np.argsort(sizes)[::-1][:8]]
        jobs = [SceneJob(pool.scenes[k], pool.targets[k], orbit4((256, 512, 1024)[j % 3], azimuth0=37.0 * j))
                for j, k in enumerate(pick)]
        tt = _targets(jobs, dev, 16)
        t = autotune.sweep_trainer(lambda ts_, ch: BatchedTrainer([j.scene for j in jobs], [j.cameras for j in jobs],
                                                                  tt, dev, ts_, depth_split=0.0, chunk=ch),
                                   warmup=3, steps=5)
        tile, chunk = min(t, key=t.get)
        tune = {autotune.label(key): v for key, v in t.items()}
        log(f"served autotune on {len(jobs)} pool scenes ({sum(sizes[k] for k in pick)} Gaussians, 4 views each; "
            "ms per step): " + ", ".join(f"{key}: {v:.2f}" for key, v in tune.items())
            + f" -> tile {tile}, chunk {chunk}")
        del jobs, tt
        torch.cuda.empty_cache()
    c = serve._lib.W2WCaps()
    c.scenes, c.frames, c.tile, c.depth_split = 128, 512, tile, 0
    c.chunk = chunk
    c.items, c.gaussians = 48_000_000, 12_000_000
    c.tiles, c.pixels, c.capacity = 2_000_000, 160_000_000, 400_000_000
    srv = serve.ContinuousServer(pool, 12_000_000, c, dev, tile)
    # warm up the kernels on a copy of the first requests
    srv.serve(copy.deepcopy(reqs[:4]))
    done_c = srv.serve(copy.deepcopy(reqs), log=log)
    st_c = serve.latency_stats(done_c)
    log(f"continuous: {st_c}")
    del srv
    torch.cuda.empty_cache()
    done_p = serve.PlainServer(pool, dev, tile).serve(copy.deepcopy(reqs), log=log)
    st_p = serve.latency_stats(done_p)
    log(f"plain: {st_p}")
    out = {"metric": "served fwd+bwd views/sec; p50/p99 latency, served vs plain", "value": st_c["views_per_s"],
           "unit": "views/s", "n_gpus": 1, "steps": len(reqs), "warmup": 4, "ms_per_step": None,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic",
           "config": {"workload": f"{len(reqs)} requests, Poisson {args.rate}/s, 30% {args.opt_steps}-step "
                                  "optimisation jobs (4 views), 70% renders of 1-8 views, scenes "
                                  "log-uniform [5k,1M] SH3, views at {256,512,1024}^2",
                      "tile": tile, "chunk": chunk or 64},
           "autotune_ms": tune,
           "served": st_c, "plain": st_p,
           "speedup_vs_plain": st_c["views_per_s"] / st_p["views_per_s"] if st_p["views_per_s"] else None}
    del pool
    torch.cuda.empty_cache()
    if emit:
        print(json.dumps(out), flush=True)
    return out


def cpu_baseline_small(args, dev):
    """The CPU-path configuration (4,096 Gaussians, SH 0, 4 views at 256^2,
    one fwd+bwd step): the CPU oracle in C, one host thread per view, median
    of --cpu-steps steps after one warm-up step.  The GPU's rate on the same
    configuration (batched trainer, fwd+bwd+Adam, graph replay) is timed
    beside it."""
    from concurrent.futures import ThreadPoolExecutor

    import torch
    from oracle import oracle
    from word2world_amd import workloads
    from word2world_amd.batched import BatchedTrainer
    from word2world_amd.scene import activate

    job = workloads.cpu_small()
    m, s, r, o, sh = [t.numpy().copy() for t in activate(job.scene.params())]
    o = o.reshape(-1)
    bg = np.ones(3, np.float32)
    V, R = len(job.cameras), job.cameras[0].width
    rng = np.random.default_rng(0)
    gC = [rng.standard_normal((3, R, R)).astype(np.float32) * 1e-3 for _ in range(V)]
    z = np.zeros((R, R), np.float32)

    def one_view(k):
        cam = job.cameras[k]
        f = oracle.forward(m, s, r, o, sh, 0, cam.view_flat(), cam.proj_flat(), cam.campos_np(), bg, R, R, 16,
                           cam.focal_x, cam.focal_y)
        oracle.backward(f, m, s, r, sh, cam.view_flat(), cam.proj_flat(), cam.campos_np(), bg, cam.focal_x,
                        cam.focal_y, gC[k], z, z)
        return f["K"]

    one_view(0)  # loads the library
    steps = []
    with ThreadPoolExecutor(V) as ex:
        list(ex.map(one_view, range(V)))
        for _ in range(max(args.cpu_steps, 1)):
            t0 = time.perf_counter()
            Ks = list(ex.map(one_view, range(V)))
            steps.append(time.perf_counter() - t0)
    med = float(np.median(steps))
    log(f"cpu oracle, 4096/SH0/4x256^2: step times {[round(t, 3) for t in steps]} s, median {med:.3f} s")
    targets = _targets([job], dev, 16)
    # a fixed capacity well above the ~18k instances per view this scene
    # needs, so that no timed step can overflow as it trains
    bt = BatchedTrainer([job.scene], [job.cameras], targets, dev, 16, capacity=1 << 20)
    for _ in range(3):
        bt.step()
    bt.capture()
    bt.step()
    torch.cuda.synchronize()
    n = 100
    t0 = time.perf_counter()
    for _ in range(n):
        bt.step()
    torch.cuda.synchronize()
    g = (time.perf_counter() - t0) / n
    if bt.overflow() or bt.skipped_overflow():
        raise RuntimeError("CPU-path configuration overflowed on the GPU")
    del bt
    return {"value": V / med, "unit": "views/s", "cores": V, "kind": "port",
            "sample": f"the CPU-path configuration in full: 4,096 Gaussians, SH0, {V} views at {R}x{R}, one "
                      f"fwd+bwd step (K per view {Ks}); CPU oracle in C, one thread per view; median of "
                      f"{len(steps)} steps after a warm-up step",
            "step_s": steps, "median_step_s": med, "cpu": _cpu_model(),
            "gpu_same_config": {"value": V / g, "unit": "views/s", "ms_per_step": g * 1e3,
                                "path": "batched trainer fwd+bwd+Adam, graph replay, tile 16"}}


def cpu_baseline(args, cams_host, P):
    """The CPU oracle (single-threaded C restatement) on one view of the
    benchmark scene, restricted to a sample of its Gaussians: forward +
    backward, timed on this host and scaled to the full scene (an
    extrapolation: the work is not exactly linear in the Gaussian count)."""
    from oracle import oracle
    from word2world_amd.scene import activate, make_scene

    n = min(args.cpu_sample_gaussians, P)
    sc = make_scene(P, 3, seed=0)  # the benchmark scene; its Gaussians are i.i.d.
    m, s, r, o, sh = [t.numpy()[:n].copy() for t in activate(sc.params())]
    o = o.reshape(-1)
    cam = cams_host[0]
    bg = np.ones(3, np.float32)
    R = cam.width
    t0 = time.perf_counter()
    f = oracle.forward(m, s, r, o, sh, 3, cam.view_flat(), cam.proj_flat(), cam.campos_np(), bg, R, R,
                       args.tile or 16, cam.focal_x, cam.focal_y)
    t1 = time.perf_counter()
    rng = np.random.default_rng(0)
    gC = rng.standard_normal((3, R, R)).astype(np.float32) * 1e-3
    z = np.zeros((R, R), np.float32)
    oracle.backward(f, m, s, r, sh, cam.view_flat(), cam.proj_flat(), cam.campos_np(), bg,
                    cam.focal_x, cam.focal_y, gC, z, z)
    t2 = time.perf_counter()
    log(f"cpu oracle: forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s (K={f['K']})")
    frac = n / P
    return {"value": frac / (t2 - t0), "unit": "views/s", "cores": 1, "kind": "port", "extrapolated": True,
            "sample": f"fwd+bwd of view 0 of the benchmark scene restricted to its first {n} of {P} "
                      f"(i.i.d.) Gaussians, {R}x{R}, SH3, CPU oracle in C on 1 thread: forward "
                      f"{t1 - t0:.2f} s, backward {t2 - t1:.2f} s, K={f['K']}; views/s scaled by "
                      f"{frac:g} (work grows linearly with the Gaussian count)",
            "cpu": _cpu_model()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
