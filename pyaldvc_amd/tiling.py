"""Local steps over sub-boxes of the volume (the reference's solver/tiling.py, al_dvc 1.2.1).

Every kernel of the local step addresses the volumes relative to a node centre and none reduces
across nodes, so a group of nodes can be solved against a crop of the volumes with the crop's origin
subtracted from the node coordinates: the results are bit-identical to the whole-volume solve as
long as every window the kernels read lies inside the crop.

Planning here is vectorised: each node of the lattice gets a block index per axis by integer
division of its lattice index (``count`` near-equal blocks per axis), the nodes are sorted by the
combined block id, and each block's box is the min/max of its node centres widened by the halo
(segment reductions over the sorted nodes), clamped to the volume. Two reasons to tile:

* ``para.tile_local`` asks for boxes of at most that many voxels per edge;
* the IC-GN kernels address a volume with 32-bit element offsets, so volumes beyond
  ``KERNEL_MAX_VOXELS`` are always cut into boxes below it (the block counts grow, longest box edge
  first, until the largest box fits).

The halo (:func:`halo`) is the reference's rule (tiling.py:141-175): half-width + the gradient
stencil's 3 voxels (+ 4 sigma of a prefilter) on the reference side; on the deformed side also the
largest displacement, a node step of margin, a quarter of the half-width for stretch, the search
radius and the 2 voxels a cubic sample reaches. The crops are device-to-device copies.
"""

from __future__ import annotations

import logging
import math
from dataclasses import dataclass

import numpy as np
import torch

logger = logging.getLogger(__name__)

STENCIL_REACH = 3  # voxels the 7-point gradient stencil reads beyond a window
CUBIC_REACH = 2  # voxels above its floor that a cubic sample reads
STRETCH_SHARE = 0.25  # share of the half-width allowed for stretch of the deformed subset
KERNEL_MAX_VOXELS = (1 << 31) - 1  # the IC-GN kernels' 32-bit element offsets


@dataclass(frozen=True)
class Box:
    """Voxels ``lo <= (z, y, x) < hi`` of a volume; ``cut`` flags the faces (-z +z -y +y -x +x) that a
    tile boundary made rather than the volume's own faces."""

    lo: tuple
    hi: tuple
    cut: tuple = (False,) * 6

    @staticmethod
    def of_volume(shape) -> "Box":
        return Box((0, 0, 0), tuple(int(s) for s in shape))

    @property
    def origin(self) -> tuple:
        return tuple(self.lo)

    @property
    def shape(self) -> tuple:
        return tuple(int(b) - int(a) for a, b in zip(self.lo, self.hi))

    @property
    def voxels(self) -> int:
        return int(np.prod(self.shape))

    @property
    def whole(self) -> bool:
        return tuple(self.lo) == (0, 0, 0) and not any(self.cut)

    @property
    def open(self) -> bool:
        """A face of the box lies inside the volume: a sample can leave the box without leaving the volume."""
        return any(self.cut)

    def crop(self, vol):
        """The box's voxels of the device volume ``vol``, contiguous; 1x1x1 placeholders, ``None`` and a
        volume already of the box's shape pass through."""
        if vol is None or tuple(vol.shape) in ((1, 1, 1), self.shape):
            return vol
        return vol[tuple(slice(a, b) for a, b in zip(self.lo, self.hi))].contiguous()

    def local(self, coords_xyz) -> np.ndarray:
        """Node centres ``[x, y, z]`` in the box's frame."""
        return np.asarray(coords_xyz, dtype=np.int64) - np.asarray(self.lo[::-1], dtype=np.int64)


@dataclass(frozen=True)
class Tiling:
    """``parts``: ``(node indices, Box)`` pairs; the halos they were planned with; ``note`` says why a
    requested tiling was not applied."""

    parts: list
    ref_halo: tuple
    def_halo: tuple
    note: str = ""

    def __len__(self) -> int:
        return len(self.parts)

    def __iter__(self):
        return iter(self.parts)

    @property
    def whole(self) -> bool:
        return len(self.parts) == 1 and self.parts[0][1].whole

    @property
    def largest(self) -> int:
        return max((box.voxels for _, box in self.parts), default=0)


def halo(para, half, disp=None, deformed: bool = False) -> tuple:
    """Voxels ``(x, y, z)`` a box keeps beyond a node centre (tiling.py:141-175)."""
    sigma = float(getattr(para, "prefilter_sigma", 0.0) or 0.0)
    ref = np.asarray(half, dtype=np.int64) + STENCIL_REACH + (int(4.0 * sigma + 0.5) if sigma > 0 else 0)
    if not deformed:
        return tuple(int(v) for v in ref)
    moved = np.zeros(3, dtype=np.int64)
    if disp is not None and np.size(disp):
        d = np.abs(np.asarray(disp, dtype=np.float64)).reshape(-1, 3)
        d = d[np.isfinite(d).all(1)]
        if d.size:
            moved = np.ceil(d.max(0)).astype(np.int64)
    step_margin = int(getattr(para, "tile_disp_margin", 0) or 0)
    if step_margin <= 0:
        step_margin = int(np.max(np.broadcast_to(np.asarray(getattr(para, "winstepsize", 0), dtype=np.int64), (3,))))
    share = float(getattr(para, "tile_strain_margin", STRETCH_SHARE) or 0.0)
    stretch = np.array([math.ceil(share * int(h)) for h in half], dtype=np.int64)
    radius = np.broadcast_to(np.asarray(getattr(para, "search_radius", 0), dtype=np.int64), (3,))
    return tuple(int(v) for v in ref + moved + step_margin + stretch + radius + CUBIC_REACH)


def _group(coords, nodes, grid_shape, counts, reach, volume_shape) -> list:
    """Blocks of ``nodes``: ``counts`` (z, y, x) near-equal slabs of the lattice per axis, each
    block's box = its nodes' centres +- ``reach`` (x, y, z) clamped to the volume."""
    gz, gy, gx = (int(s) for s in grid_shape)
    iz, rem = np.divmod(nodes, gy * gx)
    iy, ix = np.divmod(rem, gx)
    block = ((iz * counts[0]) // gz * counts[1] + (iy * counts[1]) // gy) * counts[2] + (ix * counts[2]) // gx
    order = np.argsort(block, kind="stable")
    sorted_nodes, sorted_block = nodes[order], block[order]
    starts = np.flatnonzero(np.r_[True, sorted_block[1:] != sorted_block[:-1]])
    c = coords[sorted_nodes][:, ::-1]  # (z, y, x)
    lo = np.maximum(np.minimum.reduceat(c, starts, axis=0) - np.asarray(reach[::-1]), 0)
    hi = np.minimum(np.maximum.reduceat(c, starts, axis=0) + np.asarray(reach[::-1]) + 1, np.asarray(volume_shape))
    vol = np.asarray(volume_shape)
    parts = []
    for k, a in enumerate(starts):
        b = starts[k + 1] if k + 1 < starts.size else sorted_nodes.size
        cut = tuple(bool(x) for pair in zip(lo[k] > 0, hi[k] < vol) for x in pair)
        parts.append((np.sort(sorted_nodes[a:b]), Box(tuple(int(v) for v in lo[k]), tuple(int(v) for v in hi[k]), cut)))
    return parts


def counts_within(grid_shape, coords, nodes, volume_shape, reach, limit: int = KERNEL_MAX_VOXELS) -> list:
    """Block counts (z, y, x) whose boxes all hold at most ``limit`` voxels: starting from one block,
    the count of the axis along which the largest box is longest is doubled until every box fits."""
    grid = [int(s) for s in grid_shape]
    counts = [1, 1, 1]
    while True:
        largest = max((box for _, box in _group(coords, nodes, grid, counts, reach, volume_shape)), key=lambda b: b.voxels)
        if largest.voxels <= limit:
            return counts
        growable = [ax for ax in np.argsort(largest.shape)[::-1] if counts[ax] < grid[ax]]
        if not growable:
            raise ValueError("a single node's subset and halo already exceed the kernels' volume limit")
        counts[growable[0]] = min(grid[growable[0]], 2 * counts[growable[0]])


def plan_tiles(grid_shape, coords_xyz, volume_shape, para, half, disp=None, nodes=None, edge=None,
               limit: int = KERNEL_MAX_VOXELS) -> Tiling:  # fmt: skip
    """Blocks of the node lattice (all nodes, or ``nodes``) with their boxes: boxes of at most
    ``edge`` (default ``para.tile_local``) voxels per edge when that is asked, and always below
    ``limit`` voxels (tiling.py:204-289)."""
    volume_shape = tuple(int(s) for s in volume_shape)
    coords = np.asarray(coords_xyz, dtype=np.int64)
    idx = np.arange(coords.shape[0], dtype=np.int64) if nodes is None else np.asarray(nodes, dtype=np.int64)
    rh, dh = halo(para, half), halo(para, half, disp=disp, deformed=True)
    reach = tuple(max(a, b) for a, b in zip(rh, dh))
    single = Tiling([(idx, Box.of_volume(volume_shape))], rh, dh)
    oversize = int(np.prod(volume_shape)) > limit
    want = int(edge if edge is not None else getattr(para, "tile_local", 0) or 0)
    if idx.size == 0 or (want <= 0 and not oversize):
        return single
    if int(np.prod(grid_shape)) != coords.shape[0]:
        if oversize:
            raise ValueError("a volume beyond the kernels' limit can only be tiled over a full node lattice")
        return Tiling(single.parts, rh, dh, note="tiling skipped: the nodes do not form a full lattice")
    parts = None
    if want > 0:
        if max(2 * r + 1 for r in reach) > want and not oversize:
            return Tiling(single.parts, rh, dh,
                          note=f"tile_local={want} cannot hold a subset with its {max(reach)}-voxel halo; "
                               "solving on the whole volume")  # fmt: skip
        spans = coords.max(0) - coords.min(0) + 1  # (x, y, z)
        counts = [max(1, -(-int(spans[2 - a]) // max(want - 2 * reach[2 - a], 1))) for a in range(3)]
        parts = _group(coords, idx, grid_shape, counts, reach, volume_shape)
        if oversize and max(box.voxels for _, box in parts) > limit:
            parts = None
    if parts is None:
        parts = _group(coords, idx, grid_shape, counts_within(grid_shape, coords, idx, volume_shape, reach, limit),
                       reach, volume_shape)  # fmt: skip
    if not np.array_equal(np.sort(np.concatenate([p for p, _ in parts])), np.sort(idx)):
        raise RuntimeError("tile planning lost or repeated a node")
    out = Tiling(parts, rh, dh)
    logger.info("Local steps in %d boxes (largest %d voxels; halo %s reference, %s deformed)", len(parts), out.largest,
                rh, dh)  # fmt: skip
    return out


def local_split(split_index: torch.Tensor | None, split_keep: torch.Tensor | None, nodes: torch.Tensor) -> dict:
    """The split rows of ``nodes``, renumbered from zero in node order, as kernel keyword arguments
    (tiling.py:345-361); empty when nothing is split."""
    if split_index is None or split_keep is None:
        return {}
    rows = split_index[nodes]
    has = rows >= 0
    renum = torch.cumsum(has.to(torch.int64), 0) - 1
    local = torch.where(has, renum, torch.full_like(rows, -1))
    keep = split_keep[rows[has]].contiguous()
    if keep.shape[0] == 0:
        keep = torch.zeros((1, split_keep.shape[1]), dtype=torch.uint8, device=split_keep.device)
    return {"split_index": local, "split_keep": keep}
