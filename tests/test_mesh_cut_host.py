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


def test_element_boxes_wider_than_64_voxels_are_refused_before_device_work():
    mask = np.ones((40, 200, 200), np.uint8)
    grid = (2, 2, 2)
    ax = np.array([10.0, 80.0])  # node step 70 along x and y: boxes of 71 voxels
    z, y, x = np.meshgrid(np.array([5.0, 30.0]), ax, ax, indexing="ij")
    coords = np.column_stack([x.ravel(), y.ravel(), z.ravel()])
    with pytest.raises(ValueError, match="at most 64 voxels"):
        MC.bridging_elements(mask, coords, MC.hex8_elements(grid))
