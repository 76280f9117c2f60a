"""Masking and cutting of the node grid on the GPU (mesh/grid_mesh.py:191-255, mesh/mesh_cut.py).

``apply_mask_to_mesh`` marks a node valid when its centre voxel is in the mask and at least
``min_valid_ratio`` of its subset is (the fraction from the summed-volume table kernel,
``aldvc_subset_valid_fraction``), drops the elements with an invalid corner and, with
``cut_bridging``, the elements that span a masked boundary: ``aldvc_bridging_elements`` grows,
per element, the 6-connected in-mask component of its first corner inside the corners' bounding
box (one wave per element, the box as 64-bit rows in LDS) and flags the element when a corner lies
outside it. An edge of the grid is cut when every element containing it is bridging
(``edge_ok``). Only node- and element-sized arrays touch the host.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import kernels as K
from ._lib import check, lib

MAX_ELEMENT_EXTENT = 64  # voxels per axis of an element box the LDS kernel takes (one 64-bit word per row)


def hex8_elements(grid_shape) -> np.ndarray:
    """``(E, 8)`` node indices of the grid's elements, corners ordered (x, y, z) = 000, 100, 110,
    010, 001, 101, 111, 011 with node ``iz*ny*nx + iy*nx + ix`` (grid_mesh.py:76-100)."""
    nz, ny, nx = (int(v) for v in grid_shape)
    iz, iy, ix = (a.ravel() for a in np.meshgrid(np.arange(nz - 1), np.arange(ny - 1), np.arange(nx - 1), indexing="ij"))

    def nid(a, b, c):
        return c * ny * nx + b * nx + a

    return np.column_stack([nid(ix, iy, iz), nid(ix + 1, iy, iz), nid(ix + 1, iy + 1, iz), nid(ix, iy + 1, iz),
                            nid(ix, iy, iz + 1), nid(ix + 1, iy, iz + 1), nid(ix + 1, iy + 1, iz + 1),
                            nid(ix, iy + 1, iz + 1)]).astype(np.int64)  # fmt: skip


def bridging_elements(mask, coordinates, elements) -> np.ndarray:
    """``(E,)`` True for every element (rows of 8 node indices) that spans a masked boundary
    (mesh_cut.py:98-106): its first corner is masked out, or some corner lies outside the first
    corner's 6-connected in-mask component within the element's box (clamped to the volume).
    Boxes of at most 64 voxels per axis run in the LDS kernel (one 64-bit word per box row); wider
    ones (node steps of 64 and more) in :func:`_bridging_wide`."""
    elements = np.asarray(elements, dtype=np.int64)
    E = elements.shape[0]
    if E == 0:
        return np.zeros(0, dtype=bool)
    shape = tuple(int(v) for v in np.shape(mask))
    corners = np.round(np.asarray(coordinates, dtype=np.float64)[elements]).astype(np.int64)  # (E, 8, 3)
    top = np.array(shape[::-1], dtype=np.int64) - 1  # x, y, z
    extent = (np.clip(corners.max(axis=1), 0, top) - np.clip(corners.min(axis=1), 0, top)).max(axis=1) + 1
    dev = K.device()
    m = K._Args.vol(mask, np.uint8)
    nz, ny, nx = (int(v) for v in m.shape)
    res = np.zeros(E, dtype=bool)
    narrow = np.flatnonzero(extent <= MAX_ELEMENT_EXTENT)
    if narrow.size:
        cd = torch.from_numpy(np.ascontiguousarray(corners[narrow])).to(dev)
        out = torch.zeros(narrow.size, dtype=torch.uint8, device=dev)
        max_ext = int(extent[narrow].max())
        check(lib().aldvc_bridging_elements(K._ptr(m), nz, ny, nx, K._ptr(cd), int(narrow.size), max_ext, K._ptr(out),
                                            K.stream_handle()), "aldvc_bridging_elements")  # fmt: skip
        got = out.cpu().numpy()
        if (got == 2).any():
            raise RuntimeError("an element box exceeded the declared extent")
        res[narrow] = got == 1
    wide = np.flatnonzero(extent > MAX_ELEMENT_EXTENT)
    if wide.size:
        res[wide] = _bridging_wide(m, corners[wide])
    return res


def _bridging_wide(m: torch.Tensor, corners: np.ndarray, budget: int = 1 << 28) -> np.ndarray:
    """The bridging test of :func:`bridging_elements` for element boxes of any size, on the device:
    the boxes of a batch are cut out of the mask into one zero-padded boolean tensor, and the first
    corner's component grows by one face step per round (the union of the six shifted copies, kept
    inside the mask) until a round adds nothing. ``budget``: voxels per batch."""
    nz, ny, nx = (int(v) for v in m.shape)
    top = np.array([nx - 1, ny - 1, nz - 1], dtype=np.int64)
    lo = np.clip(corners.min(axis=1), 0, top)
    hi = np.clip(corners.max(axis=1), 0, top)
    local = np.clip(corners - lo[:, None, :], 0, (hi - lo)[:, None, :])  # (E, 8, 3) x, y, z in the box
    Sx, Sy, Sz = (int(v) for v in (hi - lo).max(axis=0) + 1)
    per = max(1, budget // (Sx * Sy * Sz))
    out = np.zeros(corners.shape[0], dtype=bool)
    for b0 in range(0, corners.shape[0], per):
        idx = range(b0, min(corners.shape[0], b0 + per))
        sub = torch.zeros((len(idx), Sz, Sy, Sx), dtype=torch.bool, device=m.device)
        for j, e in enumerate(idx):
            (x0, y0, z0), (x1, y1, z1) = lo[e], hi[e]
            sub[j, : z1 - z0 + 1, : y1 - y0 + 1, : x1 - x0 + 1] = m[z0 : z1 + 1, y0 : y1 + 1, x0 : x1 + 1] != 0
        lz, ly, lx = (torch.from_numpy(local[list(idx), :, a]).to(m.device) for a in (2, 1, 0))  # (B, 8)
        rows = torch.arange(len(idx), device=m.device)
        comp = torch.zeros_like(sub)
        comp[rows, lz[:, 0], ly[:, 0], lx[:, 0]] = sub[rows, lz[:, 0], ly[:, 0], lx[:, 0]]
        while True:
            grown = comp.clone()
            for d in (1, 2, 3):  # z, y, x
                n = comp.shape[d]
                if n > 1:
                    grown.narrow(d, 1, n - 1).logical_or_(comp.narrow(d, 0, n - 1))
                    grown.narrow(d, 0, n - 1).logical_or_(comp.narrow(d, 1, n - 1))
            grown.logical_and_(sub)
            if torch.equal(grown, comp):
                break
            comp = grown
        reached = comp[rows[:, None], lz, ly, lx]  # (B, 8)
        out[list(idx)] = (~reached.all(dim=1)).cpu().numpy()
    return out


# For each axis (x, y, z): the corners at the low end of the element's four edges along that axis,
# in the corner order of hex8_elements (their partners are one node step up the axis).
EDGE_STARTS = ((0, 3, 4, 7), (0, 1, 4, 5), (0, 1, 2, 3))


def edge_ok_from_elements(elements, bridging, n_nodes: int) -> np.ndarray:
    """``(N, 3)`` flags of the +x, +y, +z edge of every node: False when the edge belongs to at least
    one of ``elements`` and every element it belongs to is bridging (mesh_cut.py:109-119). Counted
    on the device with integer scatter-adds."""
    dev = K.device()
    el = torch.as_tensor(np.asarray(elements, dtype=np.int64), device=dev).reshape(-1, 8)
    br = torch.as_tensor(np.asarray(bridging, dtype=bool), device=dev).to(torch.int32)
    members = torch.zeros(n_nodes * 3, dtype=torch.int32, device=dev)
    bridged = torch.zeros_like(members)
    for axis, starts in enumerate(EDGE_STARTS):
        slot = (el[:, list(starts)] * 3 + axis).reshape(-1)
        members.index_add_(0, slot, torch.ones_like(slot, dtype=torch.int32))
        bridged.index_add_(0, slot, br.repeat_interleave(len(starts)))
    return ((members == 0) | (bridged < members)).reshape(n_nodes, 3).cpu().numpy()


def cut_mesh(elements, coordinates, mask, n_nodes: int):
    """``(elements, edge_ok, n_dropped)``: the live elements (first corner >= 0) that span a masked
    boundary are dropped (rows set to -1) and the edges only they held are cut
    (mesh_cut.py:122-138). Rows dropped before stay dropped and do not count against an edge."""
    out = np.array(elements, dtype=np.int64, copy=True).reshape(-1, 8)
    live = out[:, 0] >= 0
    if not live.any():
        return out, np.ones((n_nodes, 3), dtype=bool), 0
    drop = np.zeros(out.shape[0], dtype=bool)
    drop[live] = bridging_elements(mask, coordinates, out[live])
    edge_ok = edge_ok_from_elements(out[live], drop[live], n_nodes)
    out[drop] = -1
    return out, edge_ok, int(drop.sum())


def grid_surface_nodes(grid_shape) -> np.ndarray:
    """Nodes on the outer surface of the node grid (grid_mesh.py:125-132)."""
    on = np.zeros(tuple(int(v) for v in grid_shape), dtype=bool)
    on[0], on[-1] = True, True
    on[:, 0], on[:, -1] = True, True
    on[:, :, 0], on[:, :, -1] = True, True
    return np.flatnonzero(on.ravel()).astype(np.int64)


def cut_edge_nodes(edge_ok, grid_shape) -> np.ndarray:
    """Sorted indices of the nodes at either end of a cut edge (mesh_cut.py:154-165)."""
    grid = tuple(int(v) for v in grid_shape)
    cut = ~np.asarray(edge_ok, dtype=bool).reshape(grid + (3,))
    ends = np.zeros(grid, dtype=bool)
    for comp, axis in ((0, 2), (1, 1), (2, 0)):  # edge component -> array axis of x, y, z
        c = cut[..., comp]
        ends |= c
        n = grid[axis]
        if n > 1:  # the node one step up the axis closes the edge
            up = [slice(None)] * 3
            down = [slice(None)] * 3
            up[axis], down[axis] = slice(1, n), slice(0, n - 1)
            ends[tuple(up)] |= c[tuple(down)]
    return np.flatnonzero(ends.ravel()).astype(np.int64)


@dataclass
class MaskedMesh:
    """What ``apply_mask_to_mesh`` returns for the local step: the reference's ``DVCMesh`` fields
    this path uses (coordinates, grid_shape, node_valid, edge_ok) plus its elements and boundary
    nodes."""

    coordinates: np.ndarray
    grid_shape: tuple
    node_valid: np.ndarray
    edge_ok: np.ndarray
    elements: np.ndarray
    boundary_nodes: np.ndarray

    @property
    def n_nodes(self) -> int:
        return int(self.coordinates.shape[0])

    def node_mesh(self):
        from .local import NodeMesh

        return NodeMesh(self.coordinates, tuple(self.grid_shape), self.node_valid, self.edge_ok)


def apply_mask_to_mesh(coordinates, grid_shape, mask, winsize, min_valid_ratio: float, cut_bridging: bool = False,
                       elements=None) -> MaskedMesh:  # fmt: skip
    """Mark nodes invalid (centre outside the mask, or less than ``min_valid_ratio`` of the subset
    inside) and drop elements with an invalid corner; with ``cut_bridging`` drop the elements that
    span a masked boundary and record the cut edges (grid_mesh.py:191-255)."""
    coordinates = np.asarray(coordinates, dtype=np.float64)
    grid_shape = tuple(int(v) for v in grid_shape)
    n = coordinates.shape[0]
    elements = hex8_elements(grid_shape) if elements is None else np.array(elements, dtype=np.int64, copy=True)
    if mask is None:
        node_valid = np.ones(n, dtype=bool)
    else:
        m = K._Args.vol(mask, np.uint8)
        nz, ny, nx = (int(v) for v in m.shape)
        c = np.round(coordinates).astype(np.int64)
        cx, cy, cz = (np.clip(c[:, a], 0, lim - 1) for a, lim in ((0, nx), (1, ny), (2, nz)))
        idx = torch.from_numpy((cz * ny + cy) * nx + cx).to(m.device)
        centre_ok = (m.reshape(-1)[idx] > 0).cpu().numpy()
        frac = np.asarray(K.subset_valid_fraction(m, torch.from_numpy(c).to(m.device), tuple(int(w) for w in winsize)).cpu())
        node_valid = centre_ok & (frac >= min_valid_ratio)
        if elements.size:
            elements[~node_valid[elements].all(axis=1)] = -1
    edge_ok = np.zeros((0, 3), dtype=bool)
    if mask is not None and cut_bridging and elements.size:
        elements, edge_ok, _n = cut_mesh(elements, coordinates, mask, n)
    boundary = set(grid_surface_nodes(grid_shape).tolist())
    if not node_valid.all():
        v = node_valid.reshape(grid_shape)
        pad = np.pad(v, 1, mode="constant", constant_values=True)
        near = np.zeros_like(v)
        for sh in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            near |= ~pad[tuple(slice(1 + s, 1 + s + d) for s, d in zip(sh, grid_shape))]
        boundary |= set(np.flatnonzero((near & v).ravel()).tolist())
    if edge_ok.size:
        boundary |= set(cut_edge_nodes(edge_ok, grid_shape).tolist())
    return MaskedMesh(coordinates, grid_shape, node_valid, edge_ok, elements, np.array(sorted(boundary), dtype=np.int64))
