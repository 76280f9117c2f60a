"""Masking and cutting of the node grid on the GPU (pyaldvc_amd/mesh_cut.py) against a brute-force
connected-component restatement and the reference's own cases (tests/test_mesh.py:82-97,
tests/test_subset_split.py:250-275)."""

from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from pyaldvc_amd import mesh_cut as MC

pytestmark = pytest.mark.gpu
SIX = ndimage.generate_binary_structure(3, 1)


def _lattice(shape, w, step):
    axes = [w // 2 + step * np.arange((n - w) // step + 1) for n in shape[::-1]]  # x, y, z
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    coords = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()]).astype(np.float64)
    return coords, (axes[2].size, axes[1].size, axes[0].size)


def _brute_bridging(mask, coords, elements):
    """mesh_cut.py:44-95 by labelling: the component of the first corner within the corners' box."""
    nz, ny, nx = mask.shape
    out = np.zeros(len(elements), dtype=bool)
    for e, el in enumerate(elements):
        c = np.round(coords[el]).astype(int)
        lo = np.maximum(c.min(0), 0)
        hi = np.minimum(c.max(0), [nx - 1, ny - 1, nz - 1])
        box = mask[lo[2] : hi[2] + 1, lo[1] : hi[1] + 1, lo[0] : hi[0] + 1] > 0
        lab, _ = ndimage.label(box, structure=SIX)
        rel = np.clip(c - lo, 0, hi - lo)
        l0 = lab[rel[0, 2], rel[0, 1], rel[0, 0]]
        out[e] = l0 == 0 or any(lab[r[2], r[1], r[0]] != l0 for r in rel[1:])
    return out


@pytest.mark.parametrize("step,seed", [(6, 0), (8, 1), (13, 2)])
def test_bridging_elements_match_connected_components(step, seed):
    shape = (40, 46, 52)
    rng = np.random.default_rng(seed)
    mask = np.ones(shape, np.uint8)
    mask[:, :, 20:22] = 0  # a wall normal to x
    mask[10:12, :, :] = 0  # a sheet normal to z with a hole in it
    mask[10:12, 20:26, 30:36] = 1
    mask &= (rng.random(shape) > 0.02).astype(np.uint8)  # scattered masked voxels
    coords, grid = _lattice(shape, 8, step)
    elements = MC.hex8_elements(grid)
    got = MC.bridging_elements(mask, coords, elements)
    ref = _brute_bridging(mask, coords, elements)
    assert ref.any() and not ref.all()
    assert np.array_equal(got, ref)


def test_wide_and_narrow_element_boxes_together():
    """Node steps of 70 voxels along x and y (boxes wider than the LDS kernel's 64) with 12 along z:
    every element through the device flood fill, against the brute-force labelling."""
    shape = (50, 160, 180)
    rng = np.random.default_rng(9)
    mask = np.ones(shape, np.uint8)
    mask[:, :, 100:102] = 0
    mask[24:26, :, :] = 0
    mask[24:26, 50:60, 20:30] = 1
    mask &= (rng.random(shape) > 0.02).astype(np.uint8)
    ax = [np.array([8.0, 78.0, 148.0, 171.0]), np.array([6.0, 76.0, 146.0]), np.array([6.0, 18.0, 30.0, 42.0])]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    coords = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])
    elements = MC.hex8_elements((ax[2].size, ax[1].size, ax[0].size))
    got = MC.bridging_elements(mask, coords, elements)
    ref = _brute_bridging(mask, coords, elements)
    assert ref.any() and not ref.all()
    assert np.array_equal(got, ref)


def test_the_mesh_is_cut_at_a_wall():
    """tests/test_subset_split.py:250-275: a wall three voxels wide through the volume; the dropped
    elements are exactly the ones that cross it, and the cut edges follow from them."""
    shape = (48, 64, 96)
    mask = np.ones(shape, np.uint8)
    mask[:, :, 46:49] = 0
    coords, grid = _lattice(shape, 16, 6)
    elements = MC.hex8_elements(grid)
    out, edge_ok, dropped = MC.cut_mesh(elements, coords, mask, coords.shape[0])
    assert dropped > 0 and not edge_ok.all()
    live = out[out[:, 0] >= 0]
    x = coords[:, 0]
    assert not np.any((x[live[:, 0]] < 45) & (x[live[:, 1]] > 49))
    crossing = (x[elements[:, 0]] < 46) & (x[elements[:, 1]] > 48)
    assert np.array_equal(out[:, 0] < 0, crossing)
    # an x edge is cut exactly when it crosses the wall; no y or z edge is
    e = edge_ok.reshape(grid + (3,))
    xs = np.unique(x)
    cut_x = (xs[:-1] < 46) & (xs[1:] > 48)
    assert np.array_equal(~e[:, :, :-1, 0], np.broadcast_to(cut_x, e[:, :, :-1, 0].shape))
    assert e[..., 1].all() and e[..., 2].all()


def test_apply_mask_to_mesh_reference_case_and_definition():
    """tests/test_mesh.py:82-97 (right half masked): nodes valid by the definition (centre in the
    mask and at least half the subset), elements with an invalid corner dropped, boundary nodes
    include the nodes next to invalid ones."""
    x0 = np.array([8.0, 16.0, 24.0, 32.0])
    Z, Y, X = np.meshgrid(x0, x0, x0, indexing="ij")
    coords = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])
    mask = np.ones((40, 40, 40), dtype=np.uint8)
    mask[:, :, 20:] = 0
    m2 = MC.apply_mask_to_mesh(coords, (4, 4, 4), mask, (8, 8, 8), 0.5)
    x = coords[:, 0]
    assert m2.node_valid[x == 8].all() and not m2.node_valid[x == 24].any()
    act = m2.elements[m2.elements[:, 0] >= 0]
    assert 0 < act.shape[0] < MC.hex8_elements((4, 4, 4)).shape[0]
    assert m2.node_valid[act].all()
    # the definition, node by node
    for i, (cx, cy, cz) in enumerate(coords.astype(int)):
        sub = mask[max(cz - 4, 0) : cz + 5, max(cy - 4, 0) : cy + 5, max(cx - 4, 0) : cx + 5]
        want = mask[cz, cy, cx] > 0 and np.count_nonzero(sub) / 729.0 >= 0.5
        assert m2.node_valid[i] == want
    v = m2.node_valid.reshape(4, 4, 4)
    inner_edge = v & ~np.roll(v, -1, axis=2)  # valid with an invalid +x neighbour
    assert set(np.flatnonzero(inner_edge.ravel())) <= set(m2.boundary_nodes.tolist())
    # with cut_bridging on a wall the cut edges join the boundary set
    mask_w = np.ones((40, 40, 40), dtype=np.uint8)
    mask_w[:, :, 19:21] = 0
    m3 = MC.apply_mask_to_mesh(coords, (4, 4, 4), mask_w, (8, 8, 8), 0.5, cut_bridging=True)
    assert m3.edge_ok.shape == (64, 3) and not m3.edge_ok[:, 0].all()
    assert set(MC.cut_edge_nodes(m3.edge_ok, (4, 4, 4)).tolist()) <= set(m3.boundary_nodes.tolist())
    assert m3.node_mesh().is_cut
