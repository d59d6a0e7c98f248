"""Served form: continuous batching admits requests as slots and HBM free up,
and each request's result is the same alone and in the mix, bit for bit
(renders and optimisation jobs alike)."""
import copy

import numpy as np
import pytest
import torch

from word2world_amd import batch, serve
from word2world_amd.scene import make_scene

pytestmark = pytest.mark.gpu


def _pool(dev):
    sizes = [3000, 800, 5000]
    scenes = [make_scene(n, 3, 300 + k) for k, n in enumerate(sizes)]
    targets = [make_scene(n, 3, 400 + k) for k, n in enumerate(sizes)]
    return serve.ScenePool(scenes, targets, dev), sizes


def _caps(pool):
    c = serve._lib.W2WCaps()
    c.scenes, c.frames, c.tile, c.depth_split = 16, 64, 16, 0
    c.items, c.gaussians, c.tiles, c.pixels, c.capacity = 200_000, 40_000, 8192, 64 * 96 * 96, 4_000_000
    return c


def test_continuous_server_same_alone_and_in_mix(gpu):
    pool, sizes = _pool(gpu)
    reqs = serve.make_requests(10, 400.0, sizes, seed=3, opt_steps=4, resolutions=(64, 96))
    pool.build_targets(reqs)
    srv = serve.ContinuousServer(pool, 20_000, _caps(pool), gpu, keep_results=True)
    done = srv.serve(copy.deepcopy(reqs))
    assert len(done) == len(reqs)
    assert all(r.done >= r.arrival for r in done)
    mixed = {r.rid: r for r in done}
    assert srv.ticks < sum(1 if r.kind == "render" else r.steps for r in reqs), "no batching happened"
    # each request alone
    for r in reqs:
        alone = serve.ContinuousServer(pool, 20_000, _caps(pool), gpu, keep_results=True)
        rr = copy.deepcopy(r)
        rr.arrival = 0.0
        out = alone.serve([rr])[0]
        m = mixed[r.rid]
        if r.kind == "render":
            for a, b in zip(out.result, m.result):
                assert torch.equal(a, b), r.rid
        else:
            # every sum on the gradient path has a fixed order (no atomics), so
            # the job's parameters are bit-identical alone and in the mix
            for name in serve.GROUPS:
                assert torch.equal(out.result[name], m.result[name]), (r.rid, name)


def test_admission_respects_hbm_budget(gpu):
    pool, sizes = _pool(gpu)
    reqs = serve.make_requests(8, 1000.0, sizes, seed=5, opt_steps=2, resolutions=(64,))
    pool.build_targets(reqs)
    # the arena holds only as many rows as the largest scene: at no time may
    # the resident scenes exceed it, and everything still completes
    budget = max(sizes)
    srv = serve.ContinuousServer(pool, budget, _caps(pool), gpu)
    done = srv.serve(reqs)
    assert len(done) == len(reqs)
    events = sorted([(r.admitted, 1, sizes[r.scene]) for r in done] +
                    [(r.done, 0, -sizes[r.scene]) for r in done])
    resident = 0
    for _, _, d in events:
        resident += d
        assert resident <= budget


def test_plain_server_completes(gpu):
    pool, sizes = _pool(gpu)
    reqs = serve.make_requests(4, 1000.0, sizes, seed=7, opt_steps=2, resolutions=(64,))
    pool.build_targets(reqs)
    done = serve.PlainServer(pool, gpu).serve(reqs)
    st = serve.latency_stats(done)
    assert st["requests"] == 4 and st["p99_s"] >= st["p50_s"] > 0


def test_budget_counts_the_workspace(gpu):
    """for_budget: the batch workspace and the arena together stay within the
    HBM budget, and the device memory the server takes matches its own
    count."""
    pool, sizes = _pool(gpu)
    caps = _caps(pool)
    ws = serve.workspace_bytes(caps)
    budget = ws + 12_000 * serve.Arena.bytes_per_row(16)
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    srv = serve.ContinuousServer.for_budget(pool, budget, caps, gpu)
    torch.cuda.synchronize()
    took = torch.cuda.memory_allocated() - before
    print("budget", budget, "server count", srv.hbm_bytes, "allocated", took, flush=True)
    assert srv.arena.rows == 12_000 and srv.hbm_bytes <= budget
    assert took <= srv.hbm_bytes + (1 << 20) and took >= 0.9 * srv.hbm_bytes
    with pytest.raises(ValueError):
        serve.ContinuousServer.for_budget(pool, ws // 2, caps, gpu)


def test_requests_that_cannot_fit_raise(gpu):
    """A request larger than the workspace or the arena is refused at
    admission; a lone request that overflows the instance capacity raises
    instead of being retried for ever."""
    pool, sizes = _pool(gpu)
    reqs = serve.make_requests(3, 1000.0, sizes, seed=9, opt_steps=1, resolutions=(96,))
    pool.build_targets(reqs)
    small = _caps(pool)
    small.pixels = 96 * 96 - 1
    with pytest.raises(ValueError, match="cannot be served"):
        serve.ContinuousServer(pool, 20_000, small, gpu).serve(copy.deepcopy(reqs))
    with pytest.raises(ValueError, match="arena"):
        serve.ContinuousServer(pool, 100, _caps(pool), gpu).serve(copy.deepcopy(reqs))
    tight = _caps(pool)
    tight.capacity = 64
    with pytest.raises(RuntimeError, match="alone needs more tile instances"):
        serve.ContinuousServer(pool, 20_000, tight, gpu).serve(copy.deepcopy(reqs))
