"""Host-side pieces of the mesh masking and cutting (pyaldvc_amd/mesh_cut.py) that need no GPU: the
nodes at the ends of cut edges, the hex8 element table, and the element-extent check that runs
before any device work."""

from __future__ import annotations

import numpy as np
import pytest

from pyaldvc_amd import mesh_cut as MC


def test_cut_edge_nodes_by_hand():
    grid = (2, 3, 4)  # nz, ny, nx
    n = int(np.prod(grid))
    edge_ok = np.ones((n, 3), dtype=bool)
    edge_ok[5, 0] = False  # node 5 = (z 0, y 1, x 1): its +x edge joins node 6
    edge_ok[2, 1] = False  # node 2 = (0, 0, 2): +y edge joins node 6
    edge_ok[9, 2] = False  # node 9 = (0, 2, 1): +z edge joins node 21
    assert MC.cut_edge_nodes(edge_ok, grid).tolist() == [2, 5, 6, 9, 21]
    assert MC.cut_edge_nodes(np.ones((n, 3), dtype=bool), grid).size == 0


def test_hex8_corners_are_unit_steps_in_the_documented_order():
    grid = (3, 4, 5)
    el = MC.hex8_elements(grid)
    assert el.shape == (2 * 3 * 4, 8)
    nz, ny, nx = grid
    iz, rem = np.divmod(el, ny * nx)
    iy, ix = np.divmod(rem, nx)
    steps = np.stack([ix - ix[:, :1], iy - iy[:, :1], iz - iz[:, :1]], -1)  # (E, 8, [x, y, z])
    order = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    assert all((steps[e] == np.array(order)).all() for e in range(el.shape[0]))
    # the low corners of each axis' four edges are the corners whose partner is one step up that axis
    for axis, starts in enumerate(MC.EDGE_STARTS):
        for s in starts:
            partner = list(order[s])
            partner[axis] += 1
            assert tuple(partner) in order


def test_wide_element_boxes_match_connected_components():
    """The flood fill for element boxes wider than the LDS kernel's 64 voxels (torch, here on CPU
    tensors) against scipy's 6-connected labelling of each box: walls, a holed sheet, scattered
    masked voxels, a first corner on masked material, boxes clamped at the volume's faces."""
    import torch
    from scipy import ndimage

    rng = np.random.default_rng(3)
    shape = (60, 150, 170)
    mask = np.ones(shape, np.uint8)
    mask[:, :, 90:93] = 0  # a wall normal to x
    mask[30:32, :, :] = 0  # a sheet normal to z with a hole in it
    mask[30:32, 60:70, 40:50] = 1
    mask &= (rng.random(shape) > 0.03).astype(np.uint8)
    ax = [np.array([-10.0, 60.0, 130.0, 175.0]), np.array([5.0, 75.0, 145.0]), np.array([2.0, 40.0, 66.0])]  # x, y, z
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    coords = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])
    elements = MC.hex8_elements((ax[2].size, ax[1].size, ax[0].size))
    corners = np.round(coords[elements]).astype(np.int64)
    corners[0, 0] = [91, 10, 10]  # a first corner inside the wall
    got = MC._bridging_wide(torch.from_numpy(mask), corners, budget=1 << 22)
    six = ndimage.generate_binary_structure(3, 1)
    top = np.array(shape[::-1]) - 1
    ref = np.zeros(len(corners), dtype=bool)
    for e, c in enumerate(corners):
        lo, hi = np.clip(c.min(0), 0, top), np.clip(c.max(0), 0, top)
        lab, _ = ndimage.label(mask[lo[2] : hi[2] + 1, lo[1] : hi[1] + 1, lo[0] : hi[0] + 1] > 0, structure=six)
        rel = np.clip(c - lo, 0, hi - lo)
        l0 = lab[rel[0, 2], rel[0, 1], rel[0, 0]]
        ref[e] = l0 == 0 or any(lab[r[2], r[1], r[0]] != l0 for r in rel[1:])
    assert ref.any() and not ref.all() and ref[0]
    assert np.array_equal(got, ref)
