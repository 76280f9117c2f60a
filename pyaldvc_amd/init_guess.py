"""Initial guess of the local step on the GPU: block-mean pyramid, phase-correlation shift,
coarse-to-fine NCC search, cleaning (median test + spring inpainting), and the coarse-lattice start.

The reference runs this stage on the host around its NCC search (solver/integer_search.py:235-694,
solver/init_disp.py:29-189, solver/coarse_init.py:33-128). Here the node fields never leave the
device between the stages: a :class:`Pyramid` keeps its running estimate ``(N, 3)`` in HBM, each
level downsamples the volumes with ``aldvc_block_downsample``, searches with
``integer_search.search_expanding`` (device maps, peaks and result assembly), and cleans the level's
field with the device median test and spring fill (``aldvc_subpb1_cleanup``, ``aldvc_spring_fill``).
Host reads inside a run are a node count per level, the expansion decisions and the conjugate-
gradient stop flags; the fields come back once, at the end.

Contract kept from the reference (its rules, not its code):

* level factor ``2**l``; template ``max(8, even(w // 2**l))``; coarsest radius ``ceil(r / 2**l) + 1``
  (at least 2), finer levels ``fine_radius``; the estimate is scaled into each level with
  round-half-to-even, and a level whose search solves nothing keeps the previous estimate;
* the automatic level count: consecutive levels from 1 while the level's template + window + a
  4-voxel margin fits every axis and the level keeps 35 % of the full-resolution standard deviation;
* cleaning between levels: median test (threshold, eps 0.1) then spring inpainting, uncut grid;
* the final cleaning: not solved, non-finite, invalid node, PCE below ``init_min_pce``, median-test
  outlier (cut mesh: reachable neighbours), and split nodes below 90 % of their subset with at least
  two usable same-side neighbours are all replaced by spring inpainting (no spring across a cut).
"""

from __future__ import annotations

import logging
import math
import time

import numpy as np
import torch

from . import integer_search as IS
from . import kernels as K
from ._lib import check, lib

logger = logging.getLogger(__name__)

PYRAMID_FINE_RADIUS = 2
COARSE_RADIUS_MARGIN = 1
MIN_COARSE_RADIUS = 2
MIN_PYRAMID_SUBSET = 8
DEFAULT_INIT_SUBSET = 16
MIN_COARSE_STD_RATIO = 0.35
LEVEL_FIT_MARGIN = 4
SPLIT_GUESS_MAX_KEEP = 0.9
SPLIT_GUESS_MIN_NEIGHBOURS = 2
MIN_COARSE_NODES_PER_AXIS = 2


def _device_f32(vol) -> torch.Tensor:
    if isinstance(vol, torch.Tensor):
        return vol.to(device=K.device(), dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(vol, dtype=np.float32)).to(K.device())


def _device(a, dtype) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(device=K.device(), dtype=dtype)
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=K.device(), dtype=dtype)


# ---------------------------------------------------------------------------- volumes
def block_downsample(vol, factor: int) -> torch.Tensor:
    """Block-mean downsampling by an integer ``factor``, trailing voxels dropped
    (integer_search.py:235-243); a device float32 volume. Each mean is summed in float64 and
    rounded once, so it agrees with NumPy's float32 mean to a few ulps."""
    v = _device_f32(vol)
    if factor <= 1:
        return v
    nz, ny, nx = (int(s) for s in v.shape)
    out = torch.empty((nz // factor, ny // factor, nx // factor), dtype=torch.float32, device=v.device)
    check(lib().aldvc_block_downsample(K._ptr(v), nz, ny, nx, int(factor), K._ptr(out), K.stream_handle()),
          "aldvc_block_downsample")  # fmt: skip
    return out


def phase_correlation_shift(f, g, voi=None, max_dim: int = 128) -> np.ndarray:
    """Global integer shift ``(dx, dy, dz)`` of ``g`` relative to ``f`` (integer_search.py:246-294):
    the VOI block-downsampled so that no axis exceeds ``max_dim``, centred, Hann-windowed, whitened
    cross power spectrum, peak of its inverse transform, scaled back. On the device."""
    fd, gd = _device_f32(f), _device_f32(g)
    if voi is not None:
        sl = voi.clamp(tuple(fd.shape)).slices
        fd, gd = fd[sl].contiguous(), gd[sl].contiguous()
    factor = max(1, int(math.ceil(max(fd.shape) / max_dim)))
    fs = block_downsample(fd, factor).double()
    gs = block_downsample(gd, factor).double()
    win = torch.ones_like(fs)
    for ax, n in enumerate(fs.shape):
        shape = [1, 1, 1]
        shape[ax] = n
        win = win * torch.from_numpy(np.hanning(n)).to(fs).reshape(shape)
    cross = torch.fft.fftn((gs - gs.mean()) * win) * torch.conj(torch.fft.fftn((fs - fs.mean()) * win))
    mag = torch.abs(cross)
    corr = torch.real(torch.fft.ifftn(cross / torch.where(mag > 1e-12, mag, torch.ones_like(mag))))
    at = np.array(np.unravel_index(int(torch.argmax(corr).item()), tuple(corr.shape)), dtype=np.float64)
    n = np.array(corr.shape, dtype=np.float64)
    at = np.where(at > n // 2, at - n, at) * factor  # wrap to signed: g(x) = f(x - shift)
    return at[::-1].copy()  # (z, y, x) -> (x, y, z)


def _voi_std(vol: torch.Tensor) -> float:
    from .volumes import voi_mean_std

    return float(voi_mean_std(vol)[1])


# ---------------------------------------------------------------------------- level rules
def coarse_radius(radius, fac: int) -> tuple[int, int, int]:
    """Search radius at the coarsest level: the radius in that level's voxels, rounded up, plus one
    (integer_search.py:566-572)."""
    return tuple(max(MIN_COARSE_RADIUS, math.ceil(int(r) / fac) + COARSE_RADIUS_MARGIN) for r in radius)


def level_template(subset, fac: int) -> tuple[int, int, int]:
    """Template size at a level: the size in that level's voxels, made even, at least 8."""
    return tuple(max(MIN_PYRAMID_SUBSET, 2 * ((int(w) // fac) // 2)) for w in subset)


def _level_fits(shape, subset, radius, fac: int) -> bool:
    """Whether the level's template, search window and margin fit every axis of the level's volume."""
    tpl = level_template(subset, fac)
    rad = coarse_radius(radius, fac)
    return all(int(n) // fac >= t + 1 + 2 * max(r, 1) + LEVEL_FIT_MARGIN for n, t, r in zip(tuple(shape)[::-1], tpl, rad))


def auto_pyramid_levels(shape, subset, radius, max_levels: int = 3, f=None, min_std_ratio: float = MIN_COARSE_STD_RATIO) -> int:
    """Number of coarse levels (integer_search.py:575-614): consecutive levels from 1 while a level
    fits (:func:`_level_fits`) and, when ``f`` is given, its block means keep ``min_std_ratio`` of the
    full-resolution standard deviation (a finer speckle than the block averages away)."""
    fd = _device_f32(f) if f is not None else None
    std_full = _voi_std(fd) if fd is not None else 0.0
    count = 0
    while count < max_levels:
        fac = 2 ** (count + 1)
        if not _level_fits(shape, subset, radius, fac):
            break
        if std_full > 0 and _voi_std(block_downsample(fd, fac)) < min_std_ratio * std_full:
            break
        count += 1
    return count


# ---------------------------------------------------------------------------- cleaning on the device
def _median_outliers(d: torch.Tensor, usable: torch.Tensor, grid, threshold: float, edge_ok=None) -> torch.Tensor:
    """Nodes the universal median test flags among the ``usable`` ones of the field ``d`` (N, 3), the
    others spring-filled first; nothing when the test is off or 27 or fewer nodes are usable."""
    from .admm import subpb1_cleanup

    if threshold <= 0:
        return torch.zeros_like(usable)
    status = torch.where(usable, 0, 1).to(torch.int8)
    vals = torch.nan_to_num(d)
    _, flagged = subpb1_cleanup(vals, vals, status, grid, float(threshold), edge_ok=edge_ok)
    return flagged & usable


def _inpaint(d: torch.Tensor, grid, edge_ok=None) -> torch.Tensor:
    from .admm import spring_fill

    return spring_fill(d.contiguous(), grid, edge_ok=edge_ok)


def _clean_level(d: torch.Tensor, grid, threshold: float) -> torch.Tensor:
    """A pyramid level's field with its median-test outliers removed and every hole inpainted
    (integer_search.py:545-556: uncut grid)."""
    usable = torch.isfinite(d).all(1)
    out = d.clone()
    out[_median_outliers(d, usable, grid, threshold)] = float("nan")
    return _inpaint(out, grid)


def _same_side_count(good: torch.Tensor, edge_ok: torch.Tensor, grid) -> torch.Tensor:
    """Per node, how many of its six face neighbours are ``good`` and joined to it by an intact edge
    (init_disp.py:29-41). ``edge_ok`` (N, 3): the +x, +y, +z edge of every node."""
    nz, ny, nx = grid
    gd = good.reshape(nz, ny, nx).to(torch.int32)
    eo = edge_ok.reshape(nz, ny, nx, 3).to(torch.int32)
    count = torch.zeros((nz, ny, nx), dtype=torch.int32, device=good.device)
    for axis, comp in ((2, 0), (1, 1), (0, 2)):  # array axis of x, y, z and its edge component
        n = gd.shape[axis]
        if n < 2:
            continue
        lo, hi = gd.narrow(axis, 0, n - 1), gd.narrow(axis, 1, n - 1)
        link = eo[..., comp].narrow(axis, 0, n - 1)  # the edge between node i and i + 1
        count.narrow(axis, 0, n - 1).add_(hi * link)  # neighbour ahead
        count.narrow(axis, 1, n - 1).add_(lo * link)  # neighbour behind
    return count.reshape(-1)


# ---------------------------------------------------------------------------- the pyramid
class Pyramid:
    """Coarse-to-fine NCC search over ``levels + 1`` levels, state on the device.

    ``coords`` (N, 3) node centres, ``grid_shape`` (nz, ny, nx) of the node lattice; ``active`` the
    nodes correlated (their reference subset is usable), the rest are inpainted at every level."""

    def __init__(self, f, g, coords, grid_shape, subset, radius, levels: int, global_shift=None, active=None,
                 fine_radius: int = PYRAMID_FINE_RADIUS, precision: str = "fp32"):  # fmt: skip
        self.f, self.g = _device_f32(f), _device_f32(g)
        dev = self.f.device
        self.coords = _device(np.asarray(coords, dtype=np.int64), torch.int64).reshape(-1, 3)
        self.N = int(self.coords.shape[0])
        self.grid = tuple(int(v) for v in grid_shape)
        self.subset, self.radius = tuple(int(w) for w in subset), tuple(int(r) for r in radius)
        self.levels, self.fine_radius, self.precision = int(levels), int(fine_radius), precision
        rows = None if active is None else torch.nonzero(_device(active, torch.bool)).squeeze(1)
        self.rows = torch.arange(self.N, device=dev) if rows is None or rows.numel() == 0 else rows
        self.estimate = torch.zeros((self.N, 3), dtype=torch.float64, device=dev)
        if global_shift is not None:
            self.estimate += torch.as_tensor(np.asarray(global_shift, dtype=np.float64), device=dev)
        self.found = IS.Found.unsolved(self.N, dev)  # the finest level that solved anything
        self.report: list = []

    def _level(self, lv: int, threshold: float, auto_expand: bool, max_expand: int) -> int:
        fac = 2**lv
        fl = block_downsample(self.f, fac)
        gl = block_downsample(self.g, fac)
        tpl = level_template(self.subset, fac)
        rad = coarse_radius(self.radius, fac) if lv == self.levels else (self.fine_radius,) * 3
        c = torch.round(self.coords[self.rows].double() / fac).long()
        s = torch.round(self.estimate[self.rows] / fac).long()
        res = IS.search_expanding(fl, gl, c, tuple(t // 2 for t in tpl), rad, s, auto_expand, max_expand, self.precision)
        n_ok = int(res.ok.sum())
        entry = {"level": lv, "n_ok": n_ok, "radius": res.radius, "subset": tpl, "expansions": res.expansions}
        self.report.append(entry)
        if n_ok == 0:
            logger.warning("Pyramid level %d: no NCC peak found; the previous estimate stands.", lv)
            return res.expansions
        field = torch.full((self.N, 3), float("nan"), dtype=torch.float64, device=self.f.device)
        field[self.rows] = res.disp * fac  # NaN where the search did not solve
        self.estimate = _clean_level(field, self.grid, threshold)
        self.found = IS.Found.unsolved(self.N, self.f.device)
        res.scatter_into(self.found, self.rows)
        entry["_stats"] = (res.cc, res.clipped, res.ok)
        return res.expansions

    def run(self, threshold: float = 2.0, auto_expand: bool = True, max_expand: int = 3) -> dict:
        expansions = sum(self._level(lv, threshold, auto_expand, max_expand) for lv in range(self.levels, -1, -1))
        for entry in self.report:
            stats = entry.pop("_stats", None)
            if stats is not None:
                cc, clipped, ok = stats
                entry["median_cc"] = float(torch.nanmedian(cc)) if cc.numel() else float("nan")
                entry["clipped_frac"] = float(clipped[ok].double().mean()) if bool(ok.any()) else 0.0
        host = self.found.host()
        return {"levels": self.levels, "per_level": self.report, "expansions": expansions, "n_searched": int(self.rows.numel()),
                "disp": self.estimate.cpu().numpy(), "cc": host["cc"], "pce": host["pce"], "ok": host["ok"],
                "clipped": host["clipped"], "estimate_device": self.estimate}  # fmt: skip


def pyramid_search(f, g, coords, grid_shape, subset, radius, levels: int = 0, global_shift=None,
                   outlier_threshold: float = 2.0, auto_expand: bool = True, max_expand: int = 3,
                   fine_radius: int = PYRAMID_FINE_RADIUS, backend: str = "hip", valid=None,
                   precision: str = "fp32") -> dict:  # fmt: skip
    """Coarse-to-fine NCC search (integer_search.py:617-694): ``disp`` (N, 3) in full-resolution
    voxels plus the last solving level's ``cc``, ``pce``, ``ok``, ``clipped`` and a per-level report.
    ``valid`` restricts the correlation to those nodes; ``levels`` 0 picks the count automatically."""
    if backend != "hip":
        raise ValueError("pyramid_search here runs on the HIP backend")
    fd = _device_f32(f)
    if levels <= 0:
        levels = auto_pyramid_levels(tuple(fd.shape), subset, radius, f=fd)
        logger.info("Pyramid: %d coarse level(s) chosen", levels)
    pyr = Pyramid(fd, g, coords, grid_shape, subset, radius, levels, global_shift, valid, fine_radius, precision)
    return pyr.run(outlier_threshold, auto_expand, max_expand)


# ---------------------------------------------------------------------------- the guess
def clean_initial_guess(disp, ok, pce, mesh, para, split_fraction=None):
    """``(U0 (N, 3), bad (N,))``: the nodes that did not solve, are not finite, are not valid, fall
    below ``para.init_min_pce`` or fail the median test are inpainted; split nodes below 90 % of their
    subset with enough usable same-side neighbours are inpainted from their own side
    (init_disp.py:144-189). On the device; host arrays out."""
    N = mesh.n_nodes
    grid = tuple(int(v) for v in mesh.grid_shape)
    d = _device(disp, torch.float64).reshape(N, 3).clone()
    bad = ~_device(ok, torch.bool).reshape(N) | ~torch.isfinite(d).all(1)
    if np.size(mesh.node_valid) == N:
        bad |= ~_device(mesh.node_valid, torch.bool)
    if para.init_min_pce > 0 and pce is not None:
        p = _device(pce, torch.float64)
        bad |= torch.isfinite(p) & (p < para.init_min_pce)
    d[bad] = float("nan")
    edge_grid = mesh.edges_ok_grid()
    bad |= _median_outliers(d, ~bad, grid, float(para.init_outlier_threshold), edge_ok=edge_grid)
    d[bad] = float("nan")
    if split_fraction is not None and edge_grid is not None:
        sf = _device(split_fraction, torch.float64)
        cut = torch.isfinite(sf) & (sf < SPLIT_GUESS_MAX_KEEP) & ~bad
        if bool(cut.any()):
            sides = _same_side_count(~bad, _device(mesh.edges_ok(), torch.bool), grid)
            own = cut & (sides >= SPLIT_GUESS_MIN_NEIGHBOURS)
            bad |= own
            d[own] = float("nan")
            logger.info("Initial guess: %d split nodes filled from their own side", int(own.sum()))
    n_bad = int(bad.sum())
    if n_bad == N:
        logger.warning("Initial guess: no node solved; U0 = 0.")
        return np.zeros((N, 3)), bad.cpu().numpy()
    out = _inpaint(d, grid, edge_ok=edge_grid)
    logger.info("Initial guess: %d of %d nodes inpainted", n_bad, N)
    return out.cpu().numpy(), bad.cpu().numpy()


def _template_size(para) -> tuple[int, int, int]:
    """``para.init_subset``, or the subset capped at 16 voxels (init_disp.py:76-81)."""
    if para.init_subset is not None:
        return tuple(int(w) for w in para.init_subset)
    return tuple(min(int(w), DEFAULT_INIT_SUBSET) for w in para.winsize)


def _search_direct(f, g, coords, para, shift, active):
    """The single-level search with radius doubling (init_guess_method "ncc")."""
    N = coords.shape[0]
    fd, gd = _device_f32(f), _device_f32(g)
    rows = torch.arange(N, device=fd.device) if active is None else torch.nonzero(_device(active, torch.bool)).squeeze(1)
    if rows.numel() == 0:
        rows = torch.arange(N, device=fd.device)
    c = _device(coords, torch.int64)[rows]
    s = None if shift is None else torch.round(torch.as_tensor(shift, device=fd.device)).long().expand(rows.numel(), 3)
    tpl = _template_size(para)
    res = IS.search_expanding(fd, gd, c, tuple(w // 2 for w in tpl), tuple(int(r) for r in para.search_radius),
                              s, bool(para.ncc_auto_expand), int(para.ncc_max_expand))  # fmt: skip
    full = IS.Found.unsolved(N, fd.device)
    res.scatter_into(full, rows)
    host = full.host()
    return host["disp"], host["ok"], host["pce"], {"method": "ncc", "radius": host["radius"], "expansions": host["expansions"],
                                                   "cc": host["cc"], "pce": host["pce"], "ok": host["ok"],
                                                   "clipped": host["clipped"]}  # fmt: skip


def compute_initial_guess(f, g, mesh, para, previous=None, split_fraction=None):
    """``U0`` (N, 3) and an info dict by ``para.init_guess_method`` (init_disp.py:53-141): "zero";
    "previous" (the previous frame's field when it is whole, else "pyramid"); "pyramid"; "ncc" (the
    single-level search with radius doubling). The searched guess is then cleaned
    (:func:`clean_initial_guess`)."""
    N = mesh.n_nodes
    method = para.init_guess_method
    if method == "zero":
        return np.zeros((N, 3)), {"method": "zero"}
    if method == "previous":
        if previous is not None and np.shape(previous) == (N, 3) and np.isfinite(previous).all():
            return np.array(previous, dtype=np.float64), {"method": "previous"}
        method = "pyramid"
    coords = np.round(np.asarray(mesh.coordinates)).astype(np.int64)
    active = mesh.node_valid if np.size(mesh.node_valid) == N else None
    shift = phase_correlation_shift(f, g, para.voi) if para.global_shift else None
    if shift is not None:
        logger.info("Global shift (dx, dy, dz) = %s voxels", np.array2string(shift, precision=1))
    if method == "pyramid":
        info = pyramid_search(f, g, coords, mesh.grid_shape, _template_size(para), para.search_radius, para.pyramid_levels,
                              global_shift=shift, outlier_threshold=para.init_outlier_threshold,
                              auto_expand=para.ncc_auto_expand, max_expand=para.ncc_max_expand,
                              fine_radius=int(getattr(para, "pyramid_fine_radius", PYRAMID_FINE_RADIUS)), valid=active)  # fmt: skip
        info.pop("estimate_device", None)
        disp, ok, pce = info["disp"], info["ok"], info["pce"]
        info["method"] = "pyramid"
    else:
        disp, ok, pce, info = _search_direct(f, g, coords, para, shift, active)
    U0, bad = clean_initial_guess(disp, ok, pce, mesh, para, split_fraction)
    info["n_bad"] = int(bad.sum())
    info["global_shift"] = shift
    return U0, info


# ---------------------------------------------------------------------------- coarse-lattice start
# The contract (coarse_init.py:33-128): solve the initial guess and the 12-DOF local step on every
# k-th node along each axis of the node grid only, then carry U and F to every node by linear
# interpolation in node position -- extended linearly past the last lattice node, non-finite lattice
# values first replaced by the nearest finite node's -- so that the fine solve starts from both; with
# fewer than two lattice nodes along an axis the plain initial guess is used and F0 is None.
#
# Here the lift is separable: the coarse table is interpolated along x, then y, then z, each a
# two-tap linear map per fine grid line on the device (the fine nodes form a full grid, so this is
# the trilinear interpolant at every node to rounding), and U and F travel as one 12-column table.


def coarse_lattice(mesh, factor: int):
    """``(coarse mesh, fine rows)``: the nodes whose grid index is a multiple of ``factor`` along
    every axis, as a mesh of their own (their validity carried over), and their rows in ``mesh``;
    ``None`` for ``factor`` 1 or when an axis would keep fewer than two nodes."""
    from .local import NodeMesh

    k = int(factor)
    grid = tuple(int(v) for v in mesh.grid_shape)
    picked = np.arange(mesh.n_nodes).reshape(grid)[::k, ::k, ::k] if k > 1 else None
    if picked is None or min(picked.shape) < MIN_COARSE_NODES_PER_AXIS:
        return None
    rows = picked.ravel()
    valid = np.asarray(mesh.node_valid, dtype=bool)
    sub_valid = valid[rows].copy() if valid.size == mesh.n_nodes else np.ones(rows.size, dtype=bool)
    return NodeMesh(np.asarray(mesh.coordinates, dtype=np.float64)[rows], picked.shape, sub_valid), rows


def _grid_axes(mesh):
    """Node positions along z, y and x of a full node grid."""
    nz, ny, nx = (int(v) for v in mesh.grid_shape)
    c = np.asarray(mesh.coordinates, dtype=np.float64).reshape(nz, ny, nx, 3)
    return c[:, 0, 0, 2].copy(), c[0, :, 0, 1].copy(), c[0, 0, :, 0].copy()


def _fill_holes_from_nearest(table: torch.Tensor) -> torch.Tensor:
    """``table`` (gz, gy, gx, C) with every row holding a non-finite entry replaced by the row of the
    nearest grid node whose row is finite (Euclidean distance in grid steps; all zeros when no row
    is finite). The reference takes the node from scipy's Euclidean distance transform
    (coarse_init.py:70-78); this is the exact transform as three separable passes -- along z, then
    y, then x, each position taking the position q along the line that minimises (p - q)^2 plus the
    squared distance found so far at q, the lower q on a tie -- which also picks the transform's
    node where several are equally near."""
    shape = table.shape
    rows = table.reshape(-1, shape[-1])
    good = torch.isfinite(rows).all(1)
    if bool(good.all()):
        return table
    if not bool(good.any()):
        return torch.zeros_like(table)
    dev = table.device
    d2 = torch.where(good, 0.0, float("inf")).to(torch.float64).reshape(shape[:3])
    donor = torch.arange(rows.shape[0], device=dev).reshape(shape[:3])
    for axis in range(3):
        n = shape[axis]
        pos = torch.arange(n, device=dev, dtype=torch.float64)
        dm, om = d2.movedim(axis, -1), donor.movedim(axis, -1)  # (..., n)
        cost = dm.unsqueeze(-2) + (pos[:, None] - pos[None, :]) ** 2  # (..., p, q)
        q = torch.argmin(cost, -1)  # the first (lowest) q of equal cost
        d2 = torch.gather(cost, -1, q.unsqueeze(-1)).squeeze(-1).movedim(-1, axis)
        donor = torch.gather(om, -1, q).movedim(-1, axis)
    return rows[donor.reshape(-1)].reshape(shape)


def _axis_lift(coarse_pos: torch.Tensor, fine_pos: torch.Tensor):
    """``(lo, w)`` per fine position: the lattice interval [lo, lo + 1] it is interpolated on (the
    first or last one beyond the lattice's ends) and its weight on lo + 1 (outside [0, 1] beyond the
    ends: linear extrapolation)."""
    lo = torch.clamp(torch.searchsorted(coarse_pos, fine_pos, right=True) - 1, 0, coarse_pos.numel() - 2)
    return lo, (fine_pos - coarse_pos[lo]) / (coarse_pos[lo + 1] - coarse_pos[lo])


def interpolate_to_mesh(coarse, values, mesh) -> np.ndarray:
    """Per-node values ``(Nc, ...)`` on the lattice ``coarse`` carried to every node of ``mesh`` (the
    contract above; float64 on the device)."""
    v = np.asarray(values, dtype=np.float64)
    dev = K.device()
    table = torch.from_numpy(np.ascontiguousarray(v.reshape(tuple(int(n) for n in coarse.grid_shape) + (-1,)))).to(dev)
    table = _fill_holes_from_nearest(table)
    for dim, (cpos, fpos) in zip((2, 1, 0), zip(_grid_axes(coarse)[::-1], _grid_axes(mesh)[::-1])):
        lo, w = _axis_lift(torch.from_numpy(cpos).to(dev), torch.from_numpy(fpos).to(dev))
        shape = [1] * table.dim()
        shape[dim] = -1
        w = w.reshape(shape)
        table = table.index_select(dim, lo) * (1.0 - w) + table.index_select(dim, lo + 1) * w
    return table.reshape((mesh.n_nodes,) + v.shape[1:]).cpu().numpy()


def coarse_initial_guess(ref, g_norm, g_prep, mesh, para):
    """``(U0, F0, info)`` from the coarse lattice of every ``para.init_coarse_factor``-th node
    (coarse_init.py:81-128; the contract above). ``info`` holds the initial guess's own entries plus
    ``method``, ``coarse_factor``, ``coarse_nodes``, ``coarse_converged``, ``coarse_bad`` and the
    ``coarse_ncc_time``, ``coarse_icgn_time`` and ``coarse_total_time`` the reference reports."""
    from .local import local_icgn, precompute_local_context

    clock = time.perf_counter
    t0 = clock()
    k = int(para.init_coarse_factor)
    lattice = coarse_lattice(mesh, k)
    if lattice is None:
        U0, info = compute_initial_guess(ref.f, g_norm, mesh, para)
        return U0, None, {**info, "coarse_factor": 1}
    sub = lattice[0]
    U_sub, info = compute_initial_guess(ref.f, g_norm, sub, para)
    t_guess = clock() - t0
    U, F, solve, bad = local_icgn(precompute_local_context(sub, ref, para), ref, g_prep, U_sub, para, sub)
    lifted = interpolate_to_mesh(sub, np.concatenate([np.asarray(U), np.asarray(F).reshape(sub.n_nodes, 9)], 1), mesh)
    converged = float(np.mean(solve.status == 0))
    info.update(method=f"coarse x{k} + {info.get('method', 'ncc')}", coarse_factor=k, coarse_nodes=int(sub.n_nodes),
                coarse_converged=converged, coarse_bad=int(bad.sum()), coarse_ncc_time=t_guess,
                coarse_icgn_time=float(solve.solve_time), coarse_total_time=clock() - t0)  # fmt: skip
    logger.info("Coarse start x%d: %d lattice nodes for %d, %.1f %% converged, %d filled; guess %.2f s, IC-GN %.2f s",
                k, sub.n_nodes, mesh.n_nodes, 100.0 * converged, int(bad.sum()), t_guess, solve.solve_time)  # fmt: skip
    return lifted[:, :3].copy(), lifted[:, 3:].reshape(mesh.n_nodes, 3, 3), info
