"""The benchmark workloads, generated from fixed seeds.

* ``cpu_small``: 4,096 Gaussians, SH 0, 4 orbital views at 256^2 (elevation 0,
  azimuth 0/90/180/270, radius 2.5, fovy 49.1), L2 to a second seed's renders.
* ``scene_100k``: 100k Gaussians, SH 3, 4 views at 512^2.
* ``scene_1m``: 1M Gaussians, SH 3, 4 views at 1024^2, elevations U[-30, 30].
* ``batched_64``: 64 independent scenes, Gaussian counts log-uniform in
  [10k, 500k], SH 3, 4 views each at a resolution drawn per scene from
  {256, 512, 1024}^2.
* served requests (see word2world_amd.serve): Poisson arrivals, scene sizes
  log-uniform in [5k, 1M], 70 % renders of 1-8 views at 256-1024^2, 30 %
  50-step optimisation jobs.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .camera import orbit_camera
from .scene import GaussianScene, make_scene


@dataclass
class SceneJob:
    scene: GaussianScene
    target_scene: GaussianScene
    cameras: list


def orbit4(res: int, elevations=(0.0, 0.0, 0.0, 0.0), azimuth0: float = 0.0):
    return [orbit_camera(azimuth0 + 90.0 * k, float(elevations[k]), 2.5, res, res) for k in range(4)]


def single_scene(n: int, deg: int, res: int, seed: int = 0, random_elevation: bool = False) -> SceneJob:
    rng = np.random.default_rng(seed)
    el = rng.uniform(-30.0, 30.0, 4) if random_elevation else np.zeros(4)
    return SceneJob(make_scene(n, deg, seed), make_scene(n, deg, seed + 1), orbit4(res, el))


def cpu_small() -> SceneJob:
    return single_scene(4096, 0, 256, 0)


def scene_100k() -> SceneJob:
    return single_scene(100_000, 3, 512, 0)


def scene_1m() -> SceneJob:
    return single_scene(1_000_000, 3, 1024, 0, random_elevation=True)


def log_uniform_counts(k: int, lo: float, hi: float, rng) -> np.ndarray:
    return np.exp(rng.uniform(np.log(lo), np.log(hi), k)).astype(np.int64)


def batched_64(n_scenes: int = 64, lo: float = 10_000, hi: float = 500_000, seed: int = 0,
               resolutions=(256, 512, 1024)) -> list:
    rng = np.random.default_rng(seed)
    counts = log_uniform_counts(n_scenes, lo, hi, rng)
    res = rng.choice(np.array(resolutions), n_scenes)
    jobs = []
    for k in range(n_scenes):
        el = rng.uniform(-30.0, 30.0, 4)
        az0 = float(rng.uniform(0.0, 360.0))
        jobs.append(SceneJob(make_scene(int(counts[k]), 3, 1000 + 2 * k),
                             make_scene(int(counts[k]), 3, 1001 + 2 * k),
                             orbit4(int(res[k]), el, az0)))
    return jobs
