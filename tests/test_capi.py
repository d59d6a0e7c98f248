"""The C-ABI library loads without a GPU and exports every function that
include/w2w_raster.h declares; layout queries are pure host logic."""
import os
import re

import pytest

from word2world_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "w2w_raster.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(w2w_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = _declared()
    assert "w2w_forward" in names and "w2w_backward" in names
    for n in names:
        assert hasattr(L, n), n
    assert set(_lib.EXPORTS) <= set(names)
    assert L.w2w_version() == 9


def test_layout_is_pure_host_logic():
    L1 = _lib.layout(4096, 256, 256, 16, 100000)
    assert L1.splats % 256 == 0 and L1.sorted_ids % 256 == 0
    L2 = _lib.layout(4096, 256, 256, 16, 200000)
    assert L2.binning_bytes > L1.binning_bytes and L2.geom_bytes == L1.geom_bytes
    assert _lib.layout(0, 64, 64, 8, 0).geom_bytes > 0
    # 16-bit tile coordinates and 30-bit look-back words bound the sizes
    with pytest.raises(RuntimeError):
        _lib.layout(10, 64, 64, 12, 100)  # unsupported tile size
    with pytest.raises(RuntimeError):
        _lib.layout(10, 64, 64, 16, 1 << 30)


def test_sort_and_scan_temp_sizes():
    L = _lib.load()
    assert L.w2w_sort_temp_bytes(0, 64) > 0
    assert L.w2w_sort_temp_bytes(1 << 20, 64) >= 12 * (1 << 20)
    assert L.w2w_sort_temp_bytes(100, 0) == 0
    assert L.w2w_scan_temp_bytes(1 << 20) >= 4 * ((1 << 20) // 16384 + 2)  # a word per tile of 16384, ticket, total


def test_batch_encode_rejects_batches_beyond_caps():
    """w2w_batch_encode is host logic: a batch larger than the workspace caps
    (scenes, frames, items, Gaussians, tiles, pixels) is refused with the
    capacity status instead of being handed to kernels sized by the caps."""
    import copy

    import numpy as np

    from word2world_amd import batch

    def frame(w, h):
        return batch.FrameSpec(np.eye(4, dtype=np.float32).reshape(-1), np.eye(4, dtype=np.float32).reshape(-1),
                               np.zeros(3, np.float32), w, h, 100.0, 100.0)

    scenes = [batch.SceneSpec(1000, 3, 16, [frame(64, 64), frame(128, 96)]),
              batch.SceneSpec(500, 3, 16, [frame(32, 32)])]
    caps = batch.caps_for(scenes, 16, 1 << 16)
    buf, totals = batch.encode(scenes, caps)
    assert list(totals[:5]) == [1500, 2500, 3, 16 + 48 + 4, 64 * 64 + 128 * 96 + 32 * 32]
    for field in ("scenes", "frames", "items", "gaussians", "tiles", "pixels"):
        small = copy.copy(caps)
        setattr(small, field, getattr(caps, field) - 1)
        with pytest.raises(RuntimeError, match="status 3"):
            batch.encode(scenes, small)
    other_tile = copy.copy(caps)
    other_tile.tile = 12
    with pytest.raises(RuntimeError, match="status 1"):
        batch.encode(scenes, other_tile)
