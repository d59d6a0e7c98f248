"""Lifetime of the drop-in's forward state (CPU).

The reference's geomBuffer / binningBuffer / imgBuffer are tensors that die
with the autograd ctx that saved them (rasterize_points.cu:24-30,
Renderer.py:78-79).  The shim keeps what the backward needs on the binning
buffer itself and indexes it weakly, so nothing outlives the buffers.  These
checks use CPU tensors in place of the workspace buffers; the GPU test
test_dropin_buffers_are_freed runs the real loop."""
import types

import pytest
import torch

from word2world_amd import raster


def _fake_ws(nbytes=64):
    return types.SimpleNamespace(P=10, W=32, H=32, tile=16, capacity=100,
                                 geom=torch.zeros(nbytes, dtype=torch.uint8),
                                 binning=torch.zeros(nbytes, dtype=torch.uint8),
                                 image=torch.zeros(nbytes, dtype=torch.uint8))


def test_state_dies_with_binning_buffer():
    before = raster.dropin_live_states()
    ws = _fake_ws()
    raster._attach_state(ws, cam="camera")
    assert raster.dropin_live_states() == before + 1
    st = raster._find_state(ws.geom, ws.binning, ws.image)
    assert (st.P, st.W, st.H, st.tile, st.capacity, st.cam) == (10, 32, 32, 16, 100, "camera")
    del st
    geom, image = ws.geom, ws.image
    del ws  # the binning buffer goes with it
    assert raster.dropin_live_states() == before
    del geom, image


def test_mismatched_buffers_rejected():
    a, b = _fake_ws(), _fake_ws()
    raster._attach_state(a, cam=None)
    raster._attach_state(b, cam=None)
    with pytest.raises(RuntimeError):
        raster._find_state(b.geom, a.binning, a.image)
    with pytest.raises(RuntimeError):
        raster._find_state(a.geom, a.binning, b.image)
    with pytest.raises(RuntimeError):
        raster._find_state(a.geom, torch.zeros(8, dtype=torch.uint8), a.image)


def test_state_survives_save_for_backward():
    """The Renderer.py pattern: buffers made in forward, saved with
    save_for_backward, the forward's own references dropped; the backward
    finds the state, and it is gone once the graph is freed."""
    before = raster.dropin_live_states()
    seen = {}

    class F(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            ws = _fake_ws()
            raster._attach_state(ws, cam="c")
            ctx.save_for_backward(ws.geom, ws.binning, ws.image)
            return x * 2

        @staticmethod
        def backward(ctx, g):
            gb, bb, ib = ctx.saved_tensors
            seen["cam"] = raster._find_state(gb, bb, ib).cam
            return g * 2

    x = torch.ones(4, requires_grad=True)
    y = F.apply(x)
    assert raster.dropin_live_states() == before + 1
    y.sum().backward()
    assert seen["cam"] == "c"
    del y
    assert raster.dropin_live_states() == before
