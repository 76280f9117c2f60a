"""Local steps over sub-boxes of the volume (solver/tiling.py, al_dvc 1.2.1).

Every kernel of the local step addresses the volumes relative to a node centre and none reduces
across nodes, so a block of nodes can be solved against a crop of the volumes with the crop's
origin subtracted from the node coordinates: the results are bit-identical to the whole-volume
solve as long as every window the kernels read lies inside the crop. This module plans such
blocks and boxes, on the reference's rules (tiling.py:141-289):

* ``tile_pad``: the voxels a box must hold beyond a node centre -- the half-width and the gradient
  stencil for the reference side, plus the measured displacement, a step of margin, the subset's
  stretch allowance, the search radius and the interpolation reach for the deformed side;
* ``plan_tiles``: the node lattice split along each axis until every box edge fits
  ``para.tile_local`` voxels.

On the GPU a second reason to tile exists: the IC-GN kernels address a volume with 32-bit element
offsets, so volumes beyond ``KERNEL_MAX_VOXELS`` are solved in boxes below that size even when
``tile_local`` is 0 (``auto_tile_edge``). The crops are device-to-device copies; the volumes stay
in HBM.
"""

from __future__ import annotations

import logging
from dataclasses import dataclass

import numpy as np
import torch

logger = logging.getLogger(__name__)

REF_STENCIL_PAD = 3  # tiling.py:34: the 7-point gradient stencil reaches 3 voxels beyond a window
SAMPLE_PAD = 2  # tiling.py:37: the highest sample a cubic kernel touches
MIN_TILE_NODES = 1  # tiling.py:42
KERNEL_MAX_VOXELS = (1 << 31) - 1  # the IC-GN kernels' 32-bit element offsets


@dataclass(frozen=True)
class TileBox:
    """A sub-box ``origin (z0, y0, x0)`` + ``shape (nz, ny, nx)``; ``open_faces`` (-z +z -y +y -x +x)
    marks the faces cut by the tile rather than by the volume (tiling.py:48-110)."""

    origin: tuple
    shape: tuple
    open_faces: tuple = (False,) * 6

    @property
    def slices(self) -> tuple:
        return tuple(slice(o, o + n) for o, n in zip(self.origin, self.shape))

    @property
    def is_whole(self) -> bool:
        return tuple(self.origin) == (0, 0, 0) and not any(self.open_faces)

    @property
    def has_open_face(self) -> bool:
        return any(self.open_faces)

    @property
    def n_voxels(self) -> int:
        return int(np.prod(self.shape))

    def crop(self, vol):
        """The box's voxels of ``vol`` (a device tensor), contiguous; 1x1x1 placeholders and a
        whole-volume box pass through."""
        if vol is None or tuple(vol.shape) == (1, 1, 1) or tuple(vol.shape) == tuple(self.shape):
            return vol
        return vol[self.slices].contiguous()

    def shift(self, coords_xyz: np.ndarray) -> np.ndarray:
        """``[x, y, z] - [x0, y0, z0]``."""
        z0, y0, x0 = self.origin
        return np.asarray(coords_xyz, dtype=np.int64) - np.array([x0, y0, z0], dtype=np.int64)[None, :]


@dataclass(frozen=True)
class TilePlan:
    """``(node indices, box)`` pairs and the pads they were cut with (tiling.py:113-133)."""

    tiles: list
    ref_pad: tuple
    def_pad: tuple
    fallback: str = ""

    def __len__(self) -> int:
        return len(self.tiles)

    def __iter__(self):
        return iter(self.tiles)

    @property
    def is_whole(self) -> bool:
        return len(self.tiles) == 1 and self.tiles[0][1].is_whole

    @property
    def max_voxels(self) -> int:
        return max((box.n_voxels for _n, box in self.tiles), default=0)


def whole_box_tile(shape) -> TileBox:
    return TileBox(origin=(0, 0, 0), shape=tuple(int(s) for s in shape))


def tile_pad(para, half, disp=None, deformed: bool = False) -> tuple:
    """Voxels a box holds beyond a node centre, ``(px, py, pz)`` (tiling.py:141-175)."""
    pad = [int(h) + REF_STENCIL_PAD for h in half]
    sigma = float(getattr(para, "prefilter_sigma", 0.0) or 0.0)
    if sigma > 0:
        pad = [p + int(4.0 * sigma + 0.5) for p in pad]
    if not deformed:
        return tuple(pad)
    strain = float(getattr(para, "tile_strain_margin", 0.25) or 0.0)
    radius = np.broadcast_to(np.asarray(getattr(para, "search_radius", 0), dtype=np.int64), (3,))
    margin = int(getattr(para, "tile_disp_margin", 0) or 0)
    if margin <= 0:
        margin = int(np.max(np.broadcast_to(np.asarray(getattr(para, "winstepsize", 0), dtype=np.int64), (3,))))
    moved = [0, 0, 0]
    if disp is not None and np.size(disp):
        d = np.abs(np.asarray(disp, dtype=np.float64)).reshape(-1, 3)
        d = d[np.isfinite(d).all(axis=1)]
        if d.size:
            moved = [int(np.ceil(d[:, k].max())) for k in range(3)]
    return tuple(int(p + moved[k] + margin + int(np.ceil(strain * h)) + int(radius[k]) + SAMPLE_PAD)
                 for k, (h, p) in enumerate(zip(half, pad)))  # fmt: skip


def _box_for(coords: np.ndarray, pad, shape) -> TileBox:
    """Smallest box holding every node's centre +- pad, clamped to the volume (tiling.py:184-201)."""
    lo, hi, faces = [], [], []
    for axis, size, p in ((2, shape[0], pad[2]), (1, shape[1], pad[1]), (0, shape[2], pad[0])):
        a = max(0, int(coords[:, axis].min()) - p)
        b = min(int(size), int(coords[:, axis].max()) + p + 1)
        lo.append(a)
        hi.append(b)
        faces += [a > 0, b < int(size)]
    return TileBox(tuple(lo), tuple(h - l for h, l in zip(hi, lo)), tuple(faces))


def _blocks(n: int, count: int) -> list:
    count = max(1, min(int(count), int(n)))
    return [b for b in np.array_split(np.arange(n, dtype=np.int64), count) if b.size]


def _split_lattice(grid_shape, coords, index, pad, volume_shape, counts) -> list:
    gz, gy, gx = (int(s) for s in grid_shape)
    selected = np.zeros(gz * gy * gx, dtype=bool)
    selected[index] = True
    lattice = np.arange(gz * gy * gx, dtype=np.int64).reshape(gz, gy, gx)
    tiles = []
    for bz in _blocks(gz, counts[0]):
        for by in _blocks(gy, counts[1]):
            for bx in _blocks(gx, counts[2]):
                block = lattice[np.ix_(bz, by, bx)].ravel()
                block = block[selected[block]]
                if block.size >= MIN_TILE_NODES:
                    tiles.append((block, _box_for(coords[block], pad, volume_shape)))
    return tiles


def auto_tile_edge(grid_shape, coords, volume_shape, pad, limit: int = KERNEL_MAX_VOXELS) -> list:
    """Block counts per lattice axis (z, y, x) whose boxes all hold at most ``limit`` voxels: the
    longest box edge is halved until they fit."""
    coords = np.asarray(coords, dtype=np.int64)
    index = np.arange(coords.shape[0], dtype=np.int64)
    counts = [1, 1, 1]
    grid = [int(s) for s in grid_shape]
    while True:
        tiles = _split_lattice(grid_shape, coords, index, pad, volume_shape, counts)
        worst = max(tiles, key=lambda t: t[1].n_voxels)[1]
        if worst.n_voxels <= limit:
            return counts
        order = np.argsort(worst.shape)[::-1]  # split the axis with the longest box edge first
        for ax in order:
            if counts[ax] < grid[ax]:
                counts[ax] = min(grid[ax], counts[ax] * 2)
                break
        else:
            raise ValueError("one node's window and halo exceed the kernels' volume limit")


def plan_tiles(grid_shape, coords_xyz, volume_shape, para, half, disp=None, nodes=None, edge=None,
               limit: int = KERNEL_MAX_VOXELS) -> TilePlan:  # fmt: skip
    """Blocks of the node lattice whose boxes fit ``para.tile_local`` voxels per edge
    (tiling.py:204-289); volumes beyond ``limit`` voxels are always split so that every box fits it."""
    volume_shape = tuple(int(s) for s in volume_shape)
    coords_xyz = np.asarray(coords_xyz, dtype=np.int64)
    n_total = coords_xyz.shape[0]
    index = np.arange(n_total, dtype=np.int64) if nodes is None else np.asarray(nodes, dtype=np.int64)
    ref_pad = tile_pad(para, half)
    def_pad = tile_pad(para, half, disp=disp, deformed=True)
    pad = tuple(max(r, d) for r, d in zip(ref_pad, def_pad))
    target = int(edge if edge is not None else getattr(para, "tile_local", 0) or 0)
    whole = TilePlan([(index, whole_box_tile(volume_shape))], ref_pad, def_pad)
    big = int(np.prod(volume_shape)) > limit
    if index.size == 0 or (target <= 0 and not big):
        return whole
    gz, gy, gx = (int(s) for s in grid_shape)
    if gz * gy * gx != n_total:
        if big:
            raise ValueError("volumes beyond the kernels' limit need a full node lattice to be tiled")
        return TilePlan(whole.tiles, ref_pad, def_pad, fallback="the node grid is not a full lattice")
    if target > 0:
        if max(2 * p + 1 for p in pad) > target and not big:
            return TilePlan(whole.tiles, ref_pad, def_pad,
                            fallback=f"tile_local={target} is smaller than one subset plus its halo; "
                                     "the local steps run on the whole volume")  # fmt: skip
        counts = []
        for axis in range(3):  # z, y, x
            p = pad[2 - axis]
            span = int(coords_xyz[:, 2 - axis].max() - coords_xyz[:, 2 - axis].min()) + 1
            counts.append(max(1, int(np.ceil(span / max(target - 2 * p, 1)))))
        tiles = _split_lattice(grid_shape, coords_xyz, index, pad, volume_shape, counts)
        if big and max(t[1].n_voxels for t in tiles) > limit:
            tiles = None
    else:
        tiles = None
    if tiles is None:
        counts = auto_tile_edge(grid_shape, coords_xyz, volume_shape, pad, limit)
        tiles = _split_lattice(grid_shape, coords_xyz, index, pad, volume_shape, counts)
    covered = np.sort(np.concatenate([t[0] for t in tiles])) if tiles else np.zeros(0, np.int64)
    if not np.array_equal(covered, np.sort(index)):
        raise RuntimeError("the tile plan did not cover every node exactly once")
    plan = TilePlan(tiles, ref_pad, def_pad)
    logger.info("Local steps tiled: %d boxes, largest %d voxels, halo %s / %s", len(tiles), plan.max_voxels,
                ref_pad, def_pad)  # fmt: skip
    return plan


def local_split(split_index: torch.Tensor | None, split_keep: torch.Tensor | None, nodes: torch.Tensor) -> dict:
    """The split rows of ``nodes`` renumbered from zero, as kernel keyword arguments
    (tiling.py:345-361); empty when nothing is split."""
    if split_index is None or split_keep is None:
        return {}
    idx = split_index[nodes]
    rows = torch.nonzero(idx >= 0).flatten()
    local = torch.full_like(idx, -1)
    if rows.numel() == 0:
        return {"split_index": local, "split_keep": torch.zeros((1, split_keep.shape[1]), dtype=torch.uint8,
                                                                device=split_keep.device)}  # fmt: skip
    local[rows] = torch.arange(rows.numel(), dtype=torch.int64, device=idx.device)
    return {"split_index": local, "split_keep": split_keep[idx[rows]].contiguous()}
