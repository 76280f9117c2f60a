"""The tile planner (pyaldvc_amd/tiling.py): box geometry, the halo rule by hand, and plans that
cover every node once with boxes holding their nodes' windows (the properties the reference's
tests/test_tiling.py checks)."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from pyaldvc_amd.synthetic import node_lattice
from pyaldvc_amd.tiling import Box, halo, local_split, plan_tiles

SHAPE = (96, 112, 128)


def _para(**kw):
    base = dict(winsize=(16, 16, 16), winstepsize=(8, 8, 8), search_radius=(5, 5, 5), tile_local=0)
    base.update(kw)
    return SimpleNamespace(**base)


def _grid(shape=SHAPE, w=16, step=8):
    coords = node_lattice(shape, (w, w, w), (step, step, step))
    g = tuple((s - w) // step + 1 for s in shape)
    return coords, g


def test_a_box_crops_and_moves_coordinates_consistently():
    box = Box(lo=(4, 8, 12), hi=(24, 32, 40), cut=(True, True, False, True, True, False))
    assert box.shape == (20, 24, 28) and box.voxels == 20 * 24 * 28 and box.origin == (4, 8, 12)
    vol = torch.arange(int(np.prod(SHAPE)), dtype=torch.float32).reshape(SHAPE)
    crop = box.crop(vol)
    assert tuple(crop.shape) == (20, 24, 28) and crop.is_contiguous()
    assert crop[0, 0, 0] == vol[4, 8, 12] and crop[-1, -1, -1] == vol[23, 31, 39]
    coords = np.array([[20, 15, 10], [30, 20, 12]], dtype=np.int64)
    assert np.array_equal(box.local(coords), coords - np.array([12, 8, 4]))
    for (x, y, z), (xl, yl, zl) in zip(coords, box.local(coords)):
        assert crop[zl, yl, xl] == vol[z, y, x]
    assert box.open and not box.whole
    whole = Box.of_volume(SHAPE)
    assert whole.whole and not whole.open and whole.crop(vol) is vol
    ph = torch.zeros((1, 1, 1))
    assert box.crop(ph) is ph and box.crop(None) is None


def test_halo_by_hand():
    para = _para()
    assert halo(para, (8, 8, 8)) == (11, 11, 11)  # half + 3 for the stencil
    # deformed: 11 + ceil(|u|) + step 8 + ceil(0.25 * 8) + radius 5 + 2 (non-finite rows ignored)
    disp = np.array([[1.2, -3.5, 0.0], [0.4, 2.0, -7.9], [np.nan, 100.0, 0.0]])
    assert halo(para, (8, 8, 8), disp=disp, deformed=True) == (11 + 2 + 8 + 2 + 5 + 2, 11 + 4 + 8 + 2 + 5 + 2,
                                                               11 + 8 + 8 + 2 + 5 + 2)  # fmt: skip
    assert halo(_para(prefilter_sigma=1.0), (8, 8, 8)) == (15, 15, 15)  # + round(4 sigma)
    assert halo(_para(tile_disp_margin=3), (8, 6, 4), deformed=True) == (11 + 3 + 2 + 5 + 2, 9 + 3 + 2 + 5 + 2,
                                                                        7 + 3 + 1 + 5 + 2)  # fmt: skip


def test_untiled_plan_is_one_whole_box():
    coords, g = _grid()
    plan = plan_tiles(g, coords, SHAPE, _para(), (8, 8, 8))
    assert plan.whole and len(plan) == 1 and plan.parts[0][0].size == coords.shape[0]


@pytest.mark.parametrize("edge", [64, 80, 100])
def test_plan_covers_every_node_once_and_boxes_hold_their_windows(edge):
    coords, g = _grid()
    para = _para(tile_local=edge)
    disp = np.full((coords.shape[0], 3), 1.5)
    plan = plan_tiles(g, coords, SHAPE, para, (8, 8, 8), disp=disp)
    assert len(plan) > 1
    allnodes = np.sort(np.concatenate([n for n, _b in plan]))
    assert np.array_equal(allnodes, np.arange(coords.shape[0]))
    pad = tuple(max(a, b) for a, b in zip(plan.ref_halo, plan.def_halo))
    for nodes, box in plan:
        c = coords[nodes]
        z0, y0, x0 = box.lo
        nz, ny, nx = box.shape
        for ax, lo, n, size in ((0, x0, nx, SHAPE[2]), (1, y0, ny, SHAPE[1]), (2, z0, nz, SHAPE[0])):
            assert lo == max(0, c[:, ax].min() - pad[ax]) and lo + n == min(size, c[:, ax].max() + pad[ax] + 1)
        # a cut face is one the tile made, not the volume
        assert box.cut == (z0 > 0, z0 + nz < SHAPE[0], y0 > 0, y0 + ny < SHAPE[1], x0 > 0, x0 + nx < SHAPE[2])
        # a block is a box of the lattice: its nodes are every lattice node inside its index ranges
        gz, gy, gx = g
        iz, rem = np.divmod(nodes, gy * gx)
        iy, ix = np.divmod(rem, gx)
        assert nodes.size == (np.ptp(iz) + 1) * (np.ptp(iy) + 1) * (np.ptp(ix) + 1)


def test_too_small_an_edge_falls_back_to_the_whole_volume():
    coords, g = _grid()
    plan = plan_tiles(g, coords, SHAPE, _para(tile_local=20), (8, 8, 8))
    assert plan.whole and "cannot hold a subset" in plan.note


@pytest.mark.parametrize("limit", [400_000, 700_000])
def test_volumes_beyond_the_limit_are_always_tiled_below_it(limit):
    coords, g = _grid()
    plan = plan_tiles(g, coords, SHAPE, _para(), (8, 8, 8), limit=limit)
    assert len(plan) > 1 and plan.largest <= limit
    assert np.array_equal(np.sort(np.concatenate([n for n, _b in plan])), np.arange(coords.shape[0]))
    # a requested edge whose boxes would exceed the limit is overridden
    plan2 = plan_tiles(g, coords, SHAPE, _para(tile_local=100), (8, 8, 8), limit=limit)
    assert plan2.largest <= limit
    # restricted to some nodes: boxes only around those
    sub = np.arange(0, coords.shape[0], 7)
    plan3 = plan_tiles(g, coords, SHAPE, _para(), (8, 8, 8), nodes=sub, limit=limit)
    assert np.array_equal(np.sort(np.concatenate([n for n, _b in plan3])), sub)


def test_local_split_renumbers_rows_in_node_order():
    split_index = torch.tensor([-1, 2, -1, 0, 1, -1])
    keep = torch.arange(12, dtype=torch.uint8).reshape(3, 4)
    out = local_split(split_index, keep, torch.tensor([5, 3, 1, 0]))
    assert out["split_index"].tolist() == [-1, 0, 1, -1]
    assert out["split_keep"].tolist() == [keep[0].tolist(), keep[2].tolist()]
    none = local_split(split_index, keep, torch.tensor([0, 2]))
    assert none["split_index"].tolist() == [-1, -1] and tuple(none["split_keep"].shape) == (1, 4)
    assert local_split(None, None, torch.tensor([0])) == {}
