"""The tile planner (pyaldvc_amd/tiling.py): geometry of boxes and plans, following the reference's
tests/test_tiling.py (box crop/shift, pads, plan coverage and fit)."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from pyaldvc_amd.synthetic import node_lattice
from pyaldvc_amd.tiling import TileBox, plan_tiles, tile_pad, whole_box_tile

SHAPE = (96, 112, 128)


def _para(**kw):
    base = dict(winsize=(16, 16, 16), winstepsize=(8, 8, 8), search_radius=(5, 5, 5), tile_local=0)
    base.update(kw)
    return SimpleNamespace(**base)


def _grid(shape=SHAPE, w=16, step=8):
    coords = node_lattice(shape, (w, w, w), (step, step, step))
    g = tuple((s - w) // step + 1 for s in shape)
    return coords, g


def test_a_box_crops_and_shifts_consistently():
    box = TileBox(origin=(4, 8, 12), shape=(20, 24, 28), open_faces=(True, True, False, True, True, False))
    vol = torch.arange(int(np.prod(SHAPE)), dtype=torch.float32).reshape(SHAPE)
    crop = box.crop(vol)
    assert tuple(crop.shape) == (20, 24, 28) and crop.is_contiguous()
    assert crop[0, 0, 0] == vol[4, 8, 12] and crop[-1, -1, -1] == vol[23, 31, 39]
    coords = np.array([[20, 15, 10], [30, 20, 12]], dtype=np.int64)
    assert np.array_equal(box.shift(coords), coords - np.array([12, 8, 4]))
    for (x, y, z), (xs, ys, zs) in zip(coords, box.shift(coords)):
        assert crop[zs, ys, xs] == vol[z, y, x]
    assert box.has_open_face and not box.is_whole
    whole = whole_box_tile(SHAPE)
    assert whole.is_whole and whole.crop(vol) is vol
    ph = torch.zeros((1, 1, 1))
    assert box.crop(ph) is ph and box.crop(None) is None


def test_pads_by_hand():
    para = _para()
    assert tile_pad(para, (8, 8, 8)) == (11, 11, 11)  # half + 3 for the stencil
    # deformed: 11 + ceil(|u|) + step 8 + ceil(0.25 * 8) + radius 5 + 2
    disp = np.array([[1.2, -3.5, 0.0], [0.4, 2.0, -7.9], [np.nan, 100.0, 0.0]])
    assert tile_pad(para, (8, 8, 8), disp=disp, deformed=True) == (11 + 2 + 8 + 2 + 5 + 2, 11 + 4 + 8 + 2 + 5 + 2,
                                                                   11 + 8 + 8 + 2 + 5 + 2)  # fmt: skip
    assert tile_pad(_para(prefilter_sigma=1.0), (8, 8, 8)) == (15, 15, 15)


def test_untiled_plan_is_one_whole_box():
    coords, g = _grid()
    plan = plan_tiles(g, coords, SHAPE, _para(), (8, 8, 8))
    assert plan.is_whole and len(plan) == 1 and plan.tiles[0][0].size == coords.shape[0]


@pytest.mark.parametrize("edge", [64, 80, 100])
def test_plan_covers_every_node_once_and_boxes_hold_their_windows(edge):
    coords, g = _grid()
    para = _para(tile_local=edge)
    disp = np.full((coords.shape[0], 3), 1.5)
    plan = plan_tiles(g, coords, SHAPE, para, (8, 8, 8), disp=disp)
    assert len(plan) > 1
    allnodes = np.sort(np.concatenate([n for n, _b in plan]))
    assert np.array_equal(allnodes, np.arange(coords.shape[0]))
    pad = tuple(max(a, b) for a, b in zip(plan.ref_pad, plan.def_pad))
    for nodes, box in plan:
        c = coords[nodes]
        z0, y0, x0 = box.origin
        nz, ny, nx = box.shape
        for ax, lo, n, size in ((0, x0, nx, SHAPE[2]), (1, y0, ny, SHAPE[1]), (2, z0, nz, SHAPE[0])):
            assert lo == max(0, c[:, ax].min() - pad[ax]) and lo + n == min(size, c[:, ax].max() + pad[ax] + 1)
        # an open face is one the tile cut, not the volume
        assert box.open_faces == (z0 > 0, z0 + nz < SHAPE[0], y0 > 0, y0 + ny < SHAPE[1], x0 > 0, x0 + nx < SHAPE[2])


def test_too_small_an_edge_falls_back_to_the_whole_volume():
    coords, g = _grid()
    plan = plan_tiles(g, coords, SHAPE, _para(tile_local=20), (8, 8, 8))
    assert plan.is_whole and "smaller than one subset" in plan.fallback


@pytest.mark.parametrize("limit", [400_000, 700_000])
def test_volumes_beyond_the_limit_are_always_tiled_below_it(limit):
    coords, g = _grid()
    plan = plan_tiles(g, coords, SHAPE, _para(), (8, 8, 8), limit=limit)
    assert len(plan) > 1 and plan.max_voxels <= limit
    assert np.array_equal(np.sort(np.concatenate([n for n, _b in plan])), np.arange(coords.shape[0]))
    # a requested edge whose boxes would exceed the limit is overridden
    plan2 = plan_tiles(g, coords, SHAPE, _para(tile_local=100), (8, 8, 8), limit=limit)
    assert plan2.max_voxels <= limit
    # restricted to some nodes: boxes only around those
    sub = np.arange(0, coords.shape[0], 7)
    plan3 = plan_tiles(g, coords, SHAPE, _para(), (8, 8, 8), nodes=sub, limit=limit)
    assert np.array_equal(np.sort(np.concatenate([n for n, _b in plan3])), sub)
