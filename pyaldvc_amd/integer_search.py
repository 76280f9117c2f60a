"""Integer NCC search of the local DVC step, resident on the GPU.

Host contract of the reference (solver/integer_search.py, al_dvc 1.2.1):

* :func:`ncc_search`            integer_search.py:329-483
* :func:`ncc_search_expanding`  integer_search.py:486-524

Everything per node stays on the device: the search windows (clamped to the volume), the ZNCC maps
(``aldvc_zncc_map``, FP32 or BF16), the peak statistics (``aldvc_ncc_peaks``: first-index argmax,
PCE, 27-point quadratic sub-voxel offset) and the result fields. :class:`Found` holds those fields
as device tensors; the reference-signature functions copy them to the host once at the end. The
only host reads inside a search are the two counts that decide whether clipped peaks are searched
again with a doubled radius.

Data contract kept from the reference: the window of node c starts at ``c + shift - h - r``,
clamped into the volume; a peak is "clipped" when it lies on a face of the window that the clamp
did not move; clipped peaks are searched again with twice the radius while they are more than 5 %
of the solved nodes, at most ``max_expand`` times, and a node keeps a re-search only where it
solved.
"""

from __future__ import annotations

import logging
from dataclasses import dataclass

import numpy as np
import torch

from . import kernels as K

logger = logging.getLogger(__name__)

CLIPPED_EXPAND_FRACTION = 0.05  # integer_search.py:38
MAP_BYTES_PER_CHUNK = 1 << 30  # device memory of the ZNCC maps of one chunk of nodes

# 27-point quadratic design (integer_search.py:210-227): columns 1, dx, dy, dz, dx dy, dx dz,
# dy dz, dx^2, dy^2, dz^2 over the 3x3x3 neighbourhood in (dz, dy, dx) C order
_dz, _dy, _dx = (a.ravel().astype(float) for a in np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij"))
DESIGN = np.column_stack([np.ones(27), _dx, _dy, _dz, _dx * _dy, _dx * _dz, _dy * _dz, _dx**2, _dy**2, _dz**2])
PINV = np.ascontiguousarray(np.linalg.pinv(DESIGN))  # (10, 27)

FIELDS = ("disp", "cc", "pce", "clipped", "ok", "peak")


@dataclass
class Found:
    """Search results of n nodes, device tensors: ``disp`` (n, 3) float64 (NaN where not solved),
    ``peak`` (n, 3) int64 integer displacement of the map peak, ``cc``, ``pce`` (n,) float64,
    ``clipped``, ``ok`` (n,) bool; ``radius`` the last radius searched, ``expansions`` how many
    times clipped peaks were searched again."""

    disp: torch.Tensor
    peak: torch.Tensor
    cc: torch.Tensor
    pce: torch.Tensor
    clipped: torch.Tensor
    ok: torch.Tensor
    radius: tuple = (0, 0, 0)
    expansions: int = 0

    @staticmethod
    def unsolved(n: int, dev) -> "Found":
        return Found(disp=torch.full((n, 3), float("nan"), dtype=torch.float64, device=dev),
                     peak=torch.zeros((n, 3), dtype=torch.int64, device=dev),
                     cc=torch.full((n,), float("nan"), dtype=torch.float64, device=dev),
                     pce=torch.full((n,), float("nan"), dtype=torch.float64, device=dev),
                     clipped=torch.zeros(n, dtype=torch.bool, device=dev),
                     ok=torch.zeros(n, dtype=torch.bool, device=dev))  # fmt: skip

    def host(self) -> dict:
        """The reference's result dictionary (plus ``peak``), host arrays."""
        out = {k: getattr(self, k).cpu().numpy() for k in FIELDS}
        out["radius"] = tuple(int(r) for r in self.radius)
        out["expansions"] = int(self.expansions)
        return out

    def scatter_into(self, full: "Found", rows: torch.Tensor) -> None:
        """Write these results into rows ``rows`` of ``full``."""
        for k in FIELDS:
            getattr(full, k)[rows] = getattr(self, k)
        full.radius, full.expansions = self.radius, self.expansions


def _half(subset):
    return tuple(int(w) // 2 for w in subset)


def search(fd: torch.Tensor, gd: torch.Tensor, coords: torch.Tensor, half, radius, shift=None,
           precision: str = "fp32") -> Found:  # fmt: skip
    """One ZNCC search of every node of ``coords`` (n, 3) int64 on the device, templates of
    half-widths ``half`` around the nodes, windows ``2 r + 1`` offsets wide around ``c + shift``."""
    dev = fd.device
    n = int(coords.shape[0])
    res = Found.unsolved(n, dev)
    res.radius = tuple(int(r) for r in radius)
    nz, ny, nx = (int(s) for s in fd.shape)
    extent = [2 * int(q) + 1 + 2 * int(p) for q, p in zip(half, res.radius)]  # window extent (x, y, z)
    if n == 0 or any(e > d for e, d in zip(extent, (nx, ny, nz))):
        return res
    vx, vy, vz = (2 * int(q) + 1 for q in res.radius)
    # the per-axis constants (x, y, z) in one upload
    consts = [int(q) for q in half] + list(res.radius) + extent + [nx, ny, nz, vx - 1, vy - 1, vz - 1]
    h, r, span, dims, edge_hi = torch.tensor(consts, dtype=torch.int64, device=dev).reshape(5, 3)
    start = coords - h - r + (0 if shift is None else shift)
    win0 = torch.minimum(torch.clamp(start, min=0), dims - span)
    # faces the clamp left alone: a peak there may lie beyond the window
    free = torch.cat([start >= 0, start <= dims - span], 1)
    t_ok = ((coords - h >= 0) & (coords + h < dims)).all(1).to(torch.uint8)
    rows = max(1, min(n, MAP_BYTES_PER_CHUNK // (4 * vx * vy * vz)))
    for a in range(0, n, rows):
        b = min(n, a + rows)
        c = coords[a:b].contiguous()
        w = win0[a:b]
        wx0, wy0, wz0 = (w[:, k].contiguous() for k in range(3))
        ok_t = t_ok[a:b].contiguous()
        maps = K.zncc_direct(fd, gd, c, *half, *res.radius, wx0, wy0, wz0, ok_t, precision=precision)
        pk, pv, pce, sub = K.ncc_peaks(maps, *res.radius, PINV)
        at = torch.stack([pk % vx, (pk // vx) % vy, pk // (vx * vy)], 1)  # peak offset (x, y, z) in the map
        peak = w + at + h - c
        on_face = ((at == 0) & free[a:b, :3]) | ((at == edge_hi) & free[a:b, 3:])
        solved = torch.isfinite(pv) & (maps[:, 0, 0, 0] > -1.5) & (ok_t > 0)  # a flat template's map is all -2
        res.peak[a:b] = peak
        res.disp[a:b] = torch.where(solved[:, None], peak.double() + sub, torch.full_like(sub, float("nan")))
        res.cc[a:b] = pv
        res.pce[a:b] = pce
        if vx > 1 or vy > 1 or vz > 1:
            res.clipped[a:b] = on_face.any(1) & (ok_t > 0)  # nodes whose template leaves the volume are not searched
        res.ok[a:b] = solved
    return res


def search_expanding(fd: torch.Tensor, gd: torch.Tensor, coords: torch.Tensor, half, radius, shift=None,
                     auto_expand: bool = True, max_expand: int = 3, precision: str = "fp32") -> Found:  # fmt: skip
    """:func:`search`, then clipped peaks searched again with a doubled radius while they are more
    than ``CLIPPED_EXPAND_FRACTION`` of the solved nodes (at most ``max_expand`` times)."""
    res = search(fd, gd, coords, half, radius, shift, precision)
    rounds = max_expand if auto_expand else 0
    while res.expansions < rounds:
        again = res.clipped & res.ok
        n_again, n_ok = (int(v) for v in torch.stack([again.sum(), res.ok.sum()]).tolist())
        if n_ok == 0 or n_again <= CLIPPED_EXPAND_FRACTION * n_ok:
            break
        wider = tuple(2 * int(q) for q in res.radius)
        logger.info("NCC: searching %d clipped peaks of %d again with radius %s", n_again, n_ok, wider)
        rows = torch.nonzero(again).squeeze(1)
        redo = search(fd, gd, coords[rows], half, wider, None if shift is None else shift[rows], precision)
        keep = rows[redo.ok]
        if keep.numel() == 0:
            break  # the wider window no longer fits the volume: the previous peaks stand
        for k in FIELDS:
            getattr(res, k)[keep] = getattr(redo, k)[redo.ok]
        res.radius = wider
        res.expansions += 1
    return res


def _device_inputs(f, g, coords, shift0):
    fd = K._Args.vol(f)
    gd = K._Args.vol(g)
    dev = fd.device
    c = torch.as_tensor(np.asarray(coords, dtype=np.int64) if not isinstance(coords, torch.Tensor) else coords,
                        dtype=torch.int64, device=dev).reshape(-1, 3).contiguous()  # fmt: skip
    s = None
    if shift0 is not None:
        s0 = shift0 if isinstance(shift0, torch.Tensor) else torch.from_numpy(np.asarray(shift0, dtype=np.float64))
        s = torch.round(s0.to(device=dev, dtype=torch.float64)).to(torch.int64).reshape(-1, 3)
    return fd, gd, c, s


def ncc_search(f, g, coords, subset, radius, shift0=None, chunk: int = 64, backend: str = "hip", precision: str = "fp32"):
    """Per-node ZNCC search with the reference's result dictionary: ``disp`` (N, 3) (NaN where not
    solvable), ``cc``, ``pce``, ``clipped``, ``ok`` (host arrays), plus ``peak``. ``chunk`` and
    ``backend`` are accepted for signature parity; chunks are sized by map memory."""
    fd, gd, c, s = _device_inputs(f, g, coords, shift0)
    out = search(fd, gd, c, _half(subset), radius, s, precision).host()
    del out["radius"], out["expansions"]
    return out


def ncc_search_expanding(f, g, coords, subset, radius, shift0=None, auto_expand=True, max_expand=3, chunk=64,
                         backend="hip", precision="fp32"):  # fmt: skip
    """:func:`ncc_search` with radius doubling for clipped peaks; adds ``expansions`` and
    ``radius`` (the last radius searched) to the dictionary."""
    fd, gd, c, s = _device_inputs(f, g, coords, shift0)
    return search_expanding(fd, gd, c, _half(subset), radius, s, auto_expand, max_expand, precision).host()
