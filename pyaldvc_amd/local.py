"""The local subset DVC step on the GPU: per-reference context, 12-DOF IC-GN over all nodes, and
the cleanup of failed / outlying nodes (solver/local_icgn.py:40-522).

Same entry points, arguments and results as the reference driver:

* :func:`precompute_local_context` ``(mesh, ref, para) -> LocalContext`` (local_icgn.py:223-347)
* :func:`local_icgn` ``(ctx, ref, g, U0, para, mesh, F0=None) -> (U, F, info, bad)``
  (local_icgn.py:398-489)
* :func:`fill_bad_nodes` ``(U, F, bad, mesh) -> (U, F)`` (local_icgn.py:492-522)
* :func:`subpb1_solver` ``(ctx, ref, g, U_hat, F_hat, vdual, mu, para, mesh) -> (U, info, bad)``
  (subpb1_solver.py:77-164), the ADMM local step

``mesh``, ``ref`` and ``para`` are duck-typed, so the reference's own ``DVCMesh``,
``ReferenceBundle`` / ``ReferenceSource`` and ``DVCPara`` objects work unchanged:

* ``mesh``: ``coordinates (N, 3)``, ``grid_shape``, ``node_valid (N,)``, ``edges_ok_grid()``;
* ``ref``: ``f`` and ``mask`` (a volume, a 1x1x1 placeholder, or ``None``), and either stored
  gradients ``gx, gy, gz`` (1x1x1 placeholders: evaluated on the fly) or a ``gradient_mode``;
* ``para``: ``winsize``, ``subset_stride``, ``min_valid_ratio``, ``hessian_cond_max``,
  ``subset_split``, ``interp_method``, ``icgn_tol``, ``icgn_dp_tol``, ``icgn_max_iter``,
  ``icgn_patience``, ``icgn_predictive_stop``, ``icgn_noise_hessian``,
  ``local_outlier_threshold``.

Everything per voxel runs in the HIP library; the context keeps its per-node arrays on the GPU
(the host copies the reference exposes are made on first access). The whole volume is one tile:
at 21 bytes per voxel a 1024x1024x306 run needs ~7 GB of the 288 GB of HBM, so the reference's
block tiling (``tiling.py``, written for 24 GB cards) and its escaped-node retry have nothing to
do here. The per-node cleanup (median test, inpainting) is host NumPy/SciPy, as in the reference.
"""

from __future__ import annotations

import logging
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import kernels as K
from .noise import STENCIL_NOISE_GAIN, noise_hessian_pattern
from .tiling import KERNEL_MAX_VOXELS, Box, local_split, plan_tiles

logger = logging.getLogger(__name__)

STATUS_CONVERGED = 0
STATUS_OUT_OF_BOUNDS = 2
STATUS_INVALID_SUBSET = 3
STATUS_SKIPPED = 6
INTERP_MODE_BY_NAME = {"linear": 0, "cubic": 1, "bspline": 2}  # interp_kernels.py:27
ICGN12_VARIANT = "auto"  # the fused kernel with the autotuned launch shape
MAX_SPLIT_BYTES = 512 * 1024 * 1024  # packed keep rows beyond this: no splitting (local_icgn.py:37)
GRADIENT_AUTO_BYTES = 8 * 1024**3  # "auto" gradients are stored below this (volume_ops.py:247)


@dataclass
class LocalSolveInfo:
    """Diagnostics of one local IC-GN pass (data_structures.py:344-351)."""

    n_iter: np.ndarray
    status: np.ndarray
    zncc: np.ndarray
    solve_time: float
    n_bad: int


@dataclass
class NodeMesh:
    """The node-grid view of ``DVCMesh`` this path uses (data_structures.py:208-280): coordinates
    ``(N, 3)`` ``[x, y, z]`` with node ``n = iz*ny*nx + iy*nx + ix``, ``grid_shape (nz, ny, nx)``,
    ``node_valid (N,)`` and ``edge_ok (N, 3)`` (the +x, +y, +z edge of every node; empty when the
    mesh was not cut)."""

    coordinates: np.ndarray
    grid_shape: tuple
    node_valid: np.ndarray
    edge_ok: np.ndarray = None

    @property
    def n_nodes(self) -> int:
        return int(self.coordinates.shape[0])

    @property
    def is_cut(self) -> bool:
        e = self.edge_ok
        return e is not None and e.shape == (self.n_nodes, 3) and not bool(e.all())

    def edges_ok(self):
        """``(N, 3)`` edge flags, all True for a mesh that was not cut (data_structures.py:276-280)."""
        e = self.edge_ok
        if e is not None and np.shape(e) == (self.n_nodes, 3):
            return np.asarray(e, dtype=bool)
        return np.ones((self.n_nodes, 3), dtype=bool)

    def edges_ok_grid(self):
        """``(nz, ny, nx, 3)`` edge flags, or ``None`` when nothing was cut."""
        return np.asarray(self.edge_ok).reshape(tuple(self.grid_shape) + (3,)) if self.is_cut else None


class LocalContext:
    """Per-(reference, mesh) data of the local solvers (local_icgn.py:40-108).

    The per-node arrays live on the GPU (``device``); the attributes the reference exposes
    (``H_all``, ``L_all``, ``meanf``, ``bottomf``, ``n_valid``, ``valid``, ``split_index``,
    ``split_keep``) are host copies made on first access.
    """

    def __init__(self, coords_int, half, device, precompute_time, stride, noise_pattern, split_fraction, volumes):
        self.coords_int = coords_int
        self.half = half
        self.device = device  # name -> torch tensor on the GPU
        self.precompute_time = precompute_time
        self.stride = stride
        self.noise_pattern = noise_pattern
        self.split_fraction = split_fraction
        self._volumes = volumes  # device copies of f, gx, gy, gz, mask for the solves
        self._host: dict = {}
        # separable passes for the whole reference and every box of it (None: each call decides)
        self.separable = None

    def _h(self, name):
        if name not in self._host:
            t = self.device.get(name)
            self._host[name] = None if t is None else t.cpu().numpy()
        return self._host[name]

    H_all = property(lambda self: self._h("H"))
    L_all = property(lambda self: self._h("L"))
    meanf = property(lambda self: self._h("meanf"))
    bottomf = property(lambda self: self._h("bottomf"))
    n_valid = property(lambda self: self._h("n_valid"))
    split_index = property(lambda self: self._h("split_index"))
    split_keep = property(lambda self: self._h("split_keep"))

    @property
    def valid(self) -> np.ndarray:
        return self._h("valid").astype(bool)

    @property
    def n_nodes(self) -> int:
        return int(self.coords_int.shape[0])

    @property
    def n_split(self) -> int:
        si = self.device.get("split_index")
        return 0 if si is None else int((si >= 0).sum())

    def noise_args(self, para):
        """``(pattern, gain)`` for the kernels; gain 0 keeps the stored Hessian (local_icgn.py:60-66)."""
        gain = float(STENCIL_NOISE_GAIN) if getattr(para, "icgn_noise_hessian", False) else 0.0
        return np.ascontiguousarray(self.noise_pattern, dtype=np.float64), gain

    def split_args(self) -> dict:
        si = self.device.get("split_index")
        return {} if si is None else {"split_index": si, "split_keep": self.device["split_keep"]}


# ---------------------------------------------------------------------------- inputs
def _is_volume(x, shape) -> bool:
    return x is not None and tuple(x.shape) == tuple(shape)


def _reference_volumes(ref, whole_gradients: bool = True) -> dict:
    """Device copies of the reference volumes: f, stored gradients (or None: on the fly), mask
    (or None), and whether gradients are stored. With ``whole_gradients`` False, stored gradients
    that the reference does not carry are left to be computed per tile box."""
    fd = K._Args.vol(ref.f)
    shape = tuple(fd.shape)
    mask = getattr(ref, "mask", None)
    md = K._Args.vol(mask, np.uint8) if _is_volume(mask, shape) else None
    gx = getattr(ref, "gx", None)
    if gx is not None:
        stored = _is_volume(gx, shape)
        grads = (K._Args.vol(gx), K._Args.vol(ref.gy), K._Args.vol(ref.gz)) if stored else (None, None, None)
    else:  # a source that builds its bundle: the reference's gradient_mode rule (volume_ops.py:247-265)
        mode = getattr(ref, "gradient_mode", "stored")
        if mode == "auto":
            mode = "on_the_fly" if 12 * int(np.prod(shape)) > GRADIENT_AUTO_BYTES else "stored"
        stored = mode == "stored"
        grads = K.compute_gradients(fd) if stored and whole_gradients else (None, None, None)
    return {"f": fd, "gx": grads[0], "gy": grads[1], "gz": grads[2], "mask": md, "stored": stored}


def _tile_volumes(vols, box):
    """``(f, (gx, gy, gz), mask)`` of one box: crops of the device volumes; stored gradients the
    reference does not carry are computed from the crop (the box holds every window plus the
    stencil's 3 voxels, so the values the kernels read equal the whole-volume ones)."""
    f = box.crop(vols["f"])
    mask = box.crop(vols["mask"]) if vols["mask"] is not None else _nomask()
    if not vols["stored"]:
        grads = (_otf(), _otf(), _otf())
    elif vols["gx"] is not None:
        grads = tuple(box.crop(vols[k]) for k in ("gx", "gy", "gz"))
    else:
        grads = K.compute_gradients(f)
    return f, grads, mask


def _volume_shape(vols) -> tuple:
    return tuple(int(v) for v in vols["f"].shape)


def _otf():
    return torch.zeros((1, 1, 1), dtype=torch.float32, device=K.device())


def _nomask():
    return torch.ones((1, 1, 1), dtype=torch.uint8, device=K.device())


def split_keep_bytes(n_candidates: int, half, stride: int) -> int:
    """Bytes of the packed keep rows of ``n_candidates`` subsets (local_icgn.py:140-143)."""
    return int(n_candidates) * ((K.subset_count(*half, stride) + 7) // 8)


def _split_rows(vols, para, coords_d, node_valid_d, half, stride, limit=MAX_SPLIT_BYTES):
    """``(split_index, split_keep, n_keep, n_inmask, cand)`` on the device, or ``None`` when
    splitting is off or over budget (local_icgn.py:146-220)."""
    if not bool(getattr(para, "subset_split", False)) or vols["mask"] is None:
        return None
    frac = K.subset_valid_fraction(vols["mask"], coords_d, tuple(int(w) for w in para.winsize))
    cand = torch.nonzero(node_valid_d & (frac < 1.0)).flatten()
    need = split_keep_bytes(int(cand.numel()), half, stride)
    if need > limit:
        logger.warning(
            "Subset splitting off for this reference: %d subsets touch a boundary and their keep rows would need "
            "%.1f GB (limit %.1f GB).", int(cand.numel()), need / 1e9, limit / 1e9,
        )  # fmt: skip
        return None
    N = int(coords_d.shape[0])
    split_index = torch.full((N,), -1, dtype=torch.int64, device=coords_d.device)
    if cand.numel() == 0:
        return split_index, torch.zeros((1, 1), dtype=torch.uint8, device=coords_d.device), None, None, cand
    margin = 0 if vols["stored"] else 3  # on-the-fly gradients read 3 voxels around every kept voxel
    rows, n_keep, n_inmask = K.build_split_rows(coords_d, cand, *half, stride, vols["mask"], margin)
    split_index[cand] = torch.arange(cand.numel(), dtype=torch.int64, device=coords_d.device)
    return split_index, rows, n_keep, n_inmask, cand


# ---------------------------------------------------------------------------- precompute
def precompute_local_context(mesh, ref, para) -> LocalContext:
    """Hessians, Cholesky factors, normalisation and validity of every node (local_icgn.py:223-347)."""
    t0 = time.perf_counter()
    dev = K.device()
    coords_int = np.round(np.asarray(mesh.coordinates, dtype=np.float64)).astype(np.int64)
    hx, hy, hz = (int(w) // 2 for w in para.winsize)
    half = (hx, hy, hz)
    stride = int(getattr(para, "subset_stride", 1))
    shape = tuple(int(v) for v in np.shape(ref.f))
    plan = plan_tiles(mesh.grid_shape, coords_int, shape, para, half)
    if plan.note:
        logger.warning("%s", plan.note)
    vols = _reference_volumes(ref, whole_gradients=plan.whole)
    coords_d = torch.from_numpy(coords_int).to(dev)
    node_valid = np.asarray(mesh.node_valid, dtype=bool)
    node_valid_d = torch.from_numpy(node_valid).to(dev)
    N = coords_int.shape[0]
    # one choice of pass for the whole reference, boxes included, so that they give the same bits
    sep = K.separable_wanted(N, half, stride, shape, vols["stored"])
    if not plan.whole:
        return _precompute_tiled(plan, vols, para, coords_int, coords_d, node_valid_d, half, stride, t0, sep)

    split = _split_rows(vols, para, coords_d, node_valid_d, half, stride)
    split_kw = {} if split is None else {"split_index": split[0], "split_keep": split[1]}
    gx, gy, gz = (vols[k] if vols["stored"] else _otf() for k in ("gx", "gy", "gz"))
    mask = vols["mask"] if vols["mask"] is not None else _nomask()
    H, L, meanf, bottomf, n_valid, valid = K.precompute_nodes(
        coords_d, hx, hy, hz, vols["f"], gx, gy, gz, mask, float(para.min_valid_ratio), float(para.hessian_cond_max),
        stride, **split_kw, separable=sep,
    )  # fmt: skip
    valid = valid & node_valid_d
    device = {"coords": coords_d, "H": H, "L": L, "meanf": meanf, "bottomf": bottomf, "n_valid": n_valid,
              "valid": valid.to(torch.uint8)}  # fmt: skip
    split_fraction = None
    if split is not None:
        split_index, rows, n_keep, n_inmask, cand = split
        device["split_index"] = split_index
        device["split_keep"] = rows
        sf = torch.ones(N, dtype=torch.float32, device=dev)
        if cand.numel():
            sf[cand] = (n_keep.double() / torch.clamp(n_inmask, min=1).double()).float()
        sf[~valid] = float("nan")
        split_fraction = sf.cpu().numpy()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    logger.info("Local precompute: %d nodes, %d valid, %.2fs", N, int(valid.sum()), dt)
    ctx = LocalContext(coords_int, half, device, dt, stride, noise_hessian_pattern(half, stride), split_fraction, vols)
    ctx.separable = sep
    return ctx


def _precompute_tiled(plan, vols, para, coords_int, coords_d, node_valid_d, half, stride, t0, sep) -> LocalContext:
    """precompute_local_context over the boxes of ``plan`` (local_icgn.py:250-300): per box the
    crops of f, mask and gradients, the split rows of its nodes, the precompute kernel on the
    shifted coordinates; the per-node results and split rows scattered back into whole arrays."""
    dev = coords_d.device
    N = coords_int.shape[0]
    hx, hy, hz = half
    H = torch.zeros((N, 12, 12), dtype=torch.float64, device=dev)
    L = torch.zeros_like(H)
    meanf = torch.zeros(N, dtype=torch.float64, device=dev)
    bottomf = torch.ones(N, dtype=torch.float64, device=dev)
    n_valid = torch.zeros(N, dtype=torch.int64, device=dev)
    valid = torch.zeros(N, dtype=torch.bool, device=dev)
    winsize = tuple(int(w) for w in para.winsize)
    split_on = bool(getattr(para, "subset_split", False)) and vols["mask"] is not None
    tiles = []
    for nodes, box in plan:
        nodes_d = torch.from_numpy(nodes).to(dev)
        loc = torch.from_numpy(box.local(coords_int[nodes])).to(dev)
        cand = None
        if split_on:  # one decision for the whole reference: count the candidates of every box first
            frac = K.subset_valid_fraction(box.crop(vols["mask"]), loc, winsize)
            cand = torch.nonzero(node_valid_d[nodes_d] & (frac < 1.0)).flatten()
        tiles.append((nodes_d, loc, box, cand))
    if split_on:
        n_cand = sum(int(t[3].numel()) for t in tiles)
        need = split_keep_bytes(n_cand, half, stride)
        if need > MAX_SPLIT_BYTES:
            logger.warning("Subset splitting off for this reference: %d subsets touch a boundary and their keep rows "
                           "would need %.1f GB (limit %.1f GB).", n_cand, need / 1e9, MAX_SPLIT_BYTES / 1e9)  # fmt: skip
            split_on = False
    margin = 0 if vols["stored"] else 3
    parts = []
    for nodes_d, loc, box, cand in tiles:
        f_t, (gx, gy, gz), m_t = _tile_volumes(vols, box)
        kw = {}
        if split_on:
            si = torch.full((loc.shape[0],), -1, dtype=torch.int64, device=dev)
            rows = torch.zeros((1, 1), dtype=torch.uint8, device=dev)
            n_keep = n_inmask = None
            if cand.numel():
                rows, n_keep, n_inmask = K.build_split_rows(loc, cand, hx, hy, hz, stride, m_t, margin)
                si[cand] = torch.arange(cand.numel(), dtype=torch.int64, device=dev)
            kw = {"split_index": si, "split_keep": rows}
            parts.append((nodes_d, cand, rows, n_keep, n_inmask))
        out = K.precompute_nodes(loc, hx, hy, hz, f_t, gx, gy, gz, m_t, float(para.min_valid_ratio),
                                 float(para.hessian_cond_max), stride, **kw, separable=sep)  # fmt: skip
        for dst, src in zip((H, L, meanf, bottomf, n_valid), out[:5]):
            dst[nodes_d] = src
        valid[nodes_d] = out[5].bool()
        del f_t, gx, gy, gz, m_t
    valid &= node_valid_d
    device = {"coords": coords_d, "H": H, "L": L, "meanf": meanf, "bottomf": bottomf, "n_valid": n_valid,
              "valid": valid.to(torch.uint8)}  # fmt: skip
    split_fraction = None
    if split_on:
        split_index = torch.full((N,), -1, dtype=torch.int64, device=dev)
        sf = torch.ones(N, dtype=torch.float32, device=dev)
        rows_all, offset = [], 0
        for nodes_d, cand, rows, n_keep, n_inmask in parts:
            if cand.numel() == 0:
                continue
            split_index[nodes_d[cand]] = offset + torch.arange(cand.numel(), dtype=torch.int64, device=dev)
            rows_all.append(rows)
            offset += int(cand.numel())
            sf[nodes_d[cand]] = (n_keep.double() / torch.clamp(n_inmask, min=1).double()).float()
        width = (K.subset_count(hx, hy, hz, stride) + 7) // 8
        device["split_index"] = split_index
        device["split_keep"] = torch.cat(rows_all) if rows_all else torch.zeros((1, width), dtype=torch.uint8, device=dev)
        sf[~valid] = float("nan")
        split_fraction = sf.cpu().numpy()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    logger.info("Local precompute: %d nodes, %d valid, %d tiles, %.2fs", N, int(valid.sum()), len(plan), dt)
    ctx = LocalContext(coords_int, half, device, dt, stride, noise_hessian_pattern(half, stride), split_fraction, vols)
    ctx.separable = sep
    ctx.n_tiles = len(plan)
    return ctx


def _solve_tiled(ctx, mesh, para, disp, run, n_out):
    """Run ``run(nodes_d, box) -> n_out per-node tensors`` over the boxes planned around ``disp``
    and scatter the results; nodes that left their box through a face the tile cut (status out of
    bounds) are solved again on the whole volume, or, beyond the kernels' size limit, in boxes with
    twice the deformed halo (local_icgn.py:427-454, subpb1_solver.py:110-135)."""
    dev = ctx.device["coords"].device
    N = ctx.n_nodes
    shape = _volume_shape(ctx._volumes)
    plan = plan_tiles(mesh.grid_shape, ctx.coords_int, shape, para, ctx.half, disp=disp)
    outs = None
    open_face = np.zeros(N, dtype=bool)

    def solve(plan_):
        nonlocal outs
        for nodes, box in plan_:
            nodes_d = torch.from_numpy(np.asarray(nodes, dtype=np.int64)).to(dev)
            res = run(nodes_d, box)
            if outs is None:
                outs = [torch.zeros((N,) + tuple(r.shape[1:]), dtype=r.dtype, device=dev) for r in res]
            for dst, src in zip(outs, res):
                dst[nodes_d] = src
            open_face[nodes] = box.open

    solve(plan)
    status = outs[n_out - 2].cpu().numpy()
    escaped = np.flatnonzero((status == STATUS_OUT_OF_BOUNDS) & open_face)
    if escaped.size:
        logger.warning("%d node(s) left their tile instead of the volume (the halo was short); solving them again",
                       int(escaped.size))  # fmt: skip
        open_face[:] = False
        if int(np.prod(shape)) <= KERNEL_MAX_VOXELS:
            solve([(escaped, Box.of_volume(shape))])
        else:
            d = np.abs(np.asarray(disp, dtype=np.float64)).reshape(N, 3)
            wider = np.where(np.isfinite(d), 2.0 * d, 0.0) + np.asarray(ctx.half, dtype=np.float64)[None, :]
            solve(plan_tiles(mesh.grid_shape, ctx.coords_int, shape, para, ctx.half, disp=wider, nodes=escaped))
    return outs


# ---------------------------------------------------------------------------- solve + cleanup
def _tiled(ctx, para) -> bool:
    """Whether the solves run in boxes: the context was tiled, ``para.tile_local`` asks for boxes,
    or the volume exceeds the kernels' 32-bit offsets."""
    if getattr(ctx, "n_tiles", 1) > 1 or int(getattr(para, "tile_local", 0) or 0) > 0:
        return True
    return int(np.prod(_volume_shape(ctx._volumes))) > KERNEL_MAX_VOXELS


def _P_from_UF(U, F):
    n = U.shape[0]
    P = np.empty((n, 12))
    P[:, :9] = F.reshape(n, 9)
    P[:, 9:] = U
    return P


def _outliers(U, status, mesh, para) -> np.ndarray:
    """bad = not converged, or flagged by the universal median test of U when the test is on and
    more than 27 nodes converged (local_icgn.py:470-480, subpb1_solver.py:140-150); the test runs
    on the device, the unknown nodes spring-filled there first."""
    from .admm import subpb1_cleanup

    _, bad = subpb1_cleanup(U, U, status, tuple(mesh.grid_shape), float(para.local_outlier_threshold),
                            edge_ok=mesh.edges_ok_grid())  # fmt: skip
    return bad


def local_icgn(ctx: LocalContext, ref, g, U0, para, mesh, F0=None):
    """12-DOF IC-GN over all nodes from ``U0`` (and ``F0``), then the median test and inpainting
    of the failed / outlying nodes: ``(U (N,3), F (N,3,3), info, bad (N,))`` (local_icgn.py:398-489).

    ``g`` is the deformed volume prepared for ``para.interp_method`` (NaN where masked)."""
    N = ctx.n_nodes
    U0 = np.asarray(U0, dtype=np.float64).reshape(N, 3)
    F0 = np.zeros((N, 3, 3)) if F0 is None else np.asarray(F0, dtype=np.float64).reshape(N, 3, 3)
    mode = INTERP_MODE_BY_NAME[para.interp_method]
    hx, hy, hz = ctx.half
    pattern, gain = ctx.noise_args(para)
    vols = ctx._volumes
    d = ctx.device
    t0 = time.perf_counter()
    gd = K._Args.vol(g)
    gx, gy, gz = (vols[k] if vols["stored"] and vols[k] is not None else _otf() for k in ("gx", "gy", "gz"))
    mask = vols["mask"] if vols["mask"] is not None else _nomask()
    P0 = torch.from_numpy(_P_from_UF(U0, F0)).to(gd.device)
    if _tiled(ctx, para):

        def run(nodes_d, box):
            f_t, (gx_t, gy_t, gz_t), m_t = _tile_volumes(vols, box)
            loc = torch.from_numpy(box.local(ctx.coords_int[nodes_d.cpu().numpy()])).to(gd.device)
            return K.icgn_12dof_parallel(
                loc, P0[nodes_d], hx, hy, hz, f_t, gx_t, gy_t, gz_t, m_t, box.crop(gd), mode, d["L"][nodes_d],
                d["meanf"][nodes_d], d["bottomf"][nodes_d], d["valid"][nodes_d], float(para.icgn_tol),
                float(para.icgn_dp_tol), int(para.icgn_max_iter), int(para.icgn_patience), ctx.stride, d["H"][nodes_d],
                pattern, gain, bool(para.icgn_predictive_stop),
                **local_split(d.get("split_index"), d.get("split_keep"), nodes_d), variant=ICGN12_VARIANT,
                separable=ctx.separable,
            )  # fmt: skip

        P, n_iter, status, zncc = _solve_tiled(ctx, mesh, para, U0, run, 4)
    else:
        P, n_iter, status, zncc = K.icgn_12dof_parallel(
            d["coords"], P0, hx, hy, hz, vols["f"], gx, gy, gz, mask, gd, mode, d["L"], d["meanf"], d["bottomf"],
            d["valid"], float(para.icgn_tol), float(para.icgn_dp_tol), int(para.icgn_max_iter), int(para.icgn_patience),
            ctx.stride, d["H"], pattern, gain, bool(para.icgn_predictive_stop), **ctx.split_args(), variant=ICGN12_VARIANT,
            separable=ctx.separable,
        )  # fmt: skip
    P = P.cpu().numpy()
    n_iter = n_iter.cpu().numpy().astype(np.int32)
    status = status.cpu().numpy().astype(np.int8)
    zncc = zncc.cpu().numpy().astype(np.float64)
    solve_time = time.perf_counter() - t0

    U = P[:, 9:12].copy()
    F = P[:, :9].reshape(N, 3, 3).copy()
    bad = _outliers(U, status, mesh, para)
    n_bad = int(np.sum(bad & (status != STATUS_INVALID_SUBSET) & (status != STATUS_SKIPPED)))
    logger.info("Local IC-GN: %d/%d converged, %d bad, %.2fs", int(np.sum(status == 0)), N, int(bad.sum()), solve_time)
    U, F = fill_bad_nodes(U, F, bad, mesh)
    info = LocalSolveInfo(n_iter=n_iter, status=status, zncc=zncc, solve_time=solve_time, n_bad=n_bad)
    return U, F, info, bad


def subpb1_solver(ctx: LocalContext, ref, g, U_hat, F_hat, vdual, mu, para, mesh):
    """ADMM local step: per node, the 3 translations minimising
    ``ZNSSD(u; F_hat) + mu/2 |u - (U_hat + vdual)|^2`` from ``U_hat``; median test on the result,
    and the nodes that fail keep ``U_hat`` (subpb1_solver.py:77-164). ``(U, info, bad)``."""
    N = ctx.n_nodes
    U_hat = np.ascontiguousarray(U_hat, dtype=np.float64).reshape(N, 3)
    F_hat = np.ascontiguousarray(F_hat, dtype=np.float64).reshape(N, 3, 3)
    vdual = np.ascontiguousarray(vdual, dtype=np.float64).reshape(N, 3)
    mode = INTERP_MODE_BY_NAME[para.interp_method]
    hx, hy, hz = ctx.half
    pattern, gain = ctx.noise_args(para)
    vols = ctx._volumes
    d = ctx.device
    t0 = time.perf_counter()
    gd = K._Args.vol(g)
    dev = gd.device
    gx, gy, gz = (vols[k] if vols["stored"] and vols[k] is not None else _otf() for k in ("gx", "gy", "gz"))
    mask = vols["mask"] if vols["mask"] is not None else _nomask()
    Uh_d, Fh_d, vd_d = (torch.from_numpy(a).to(dev) for a in (U_hat, F_hat, vdual))
    if _tiled(ctx, para):

        def run(nodes_d, box):
            f_t, (gx_t, gy_t, gz_t), m_t = _tile_volumes(vols, box)
            loc = torch.from_numpy(box.local(ctx.coords_int[nodes_d.cpu().numpy()])).to(dev)
            return K.icgn_3dof_parallel(
                loc, Uh_d[nodes_d], Fh_d[nodes_d], vd_d[nodes_d], hx, hy, hz, f_t, gx_t, gy_t, gz_t, m_t, box.crop(gd),
                mode, d["H"][nodes_d], d["meanf"][nodes_d], d["bottomf"][nodes_d], d["valid"][nodes_d], float(mu),
                float(para.icgn_tol), float(para.icgn_dp_tol), int(para.icgn_max_iter), int(para.icgn_patience),
                ctx.stride, float(pattern[9, 9]), gain, bool(para.icgn_predictive_stop),
                **local_split(d.get("split_index"), d.get("split_keep"), nodes_d), separable=ctx.separable,
            )  # fmt: skip

        U, n_iter, status, zncc = _solve_tiled(ctx, mesh, para, U_hat + vdual, run, 4)
    else:
        U, n_iter, status, zncc = K.icgn_3dof_parallel(
            d["coords"], Uh_d, Fh_d, vd_d, hx, hy, hz, vols["f"], gx, gy, gz, mask, gd, mode, d["H"], d["meanf"],
            d["bottomf"], d["valid"], float(mu), float(para.icgn_tol), float(para.icgn_dp_tol), int(para.icgn_max_iter),
            int(para.icgn_patience), ctx.stride, float(pattern[9, 9]), gain, bool(para.icgn_predictive_stop),
            **ctx.split_args(), separable=ctx.separable,
        )  # fmt: skip
    # the cleanup on the device: median test (unknown nodes spring-filled first) and
    # U_filled = bad ? U_hat : U (nodes the local solver could not handle keep the global estimate)
    from .admm import subpb1_cleanup

    U_filled, bad = subpb1_cleanup(U, Uh_d, status, tuple(mesh.grid_shape), float(para.local_outlier_threshold),
                                   edge_ok=mesh.edges_ok_grid())  # fmt: skip
    U_filled = U_filled.cpu().numpy().reshape(N, 3)
    bad = bad.cpu().numpy()
    n_iter = n_iter.cpu().numpy().astype(np.int32)
    status = status.cpu().numpy().astype(np.int8)
    zncc = zncc.cpu().numpy().astype(np.float64)
    solve_time = time.perf_counter() - t0
    n_bad = int(np.sum(bad & (status != STATUS_INVALID_SUBSET) & (status != STATUS_SKIPPED)))
    logger.info("Subpb1 (3-DOF): %d/%d converged, %d bad, %.2fs", int(np.sum(status == 0)), N, int(bad.sum()), solve_time)
    return U_filled, LocalSolveInfo(n_iter=n_iter, status=status, zncc=zncc, solve_time=solve_time, n_bad=n_bad), bad


def fill_bad_nodes(U, F, bad, mesh):
    """The ``bad`` nodes of U (and F) replaced by the spring model on the node grid, no spring across
    a cut edge (local_icgn.py:492-522). U and the nine F columns are filled on the device, three
    columns per solve."""
    from .admm import spring_fill

    if not bad.any():
        return U, F
    if bad.all():
        logger.warning("Every node is bad: U and F set to zero.")
        return np.zeros_like(U), (None if F is None else np.zeros_like(F))
    shape = tuple(mesh.grid_shape)
    edge_ok = mesh.edges_ok_grid()
    holes = np.asarray(bad, dtype=bool).reshape(-1)
    cols = np.asarray(U, dtype=np.float64).reshape(-1, 3)
    if F is not None:
        cols = np.concatenate([cols, np.asarray(F, dtype=np.float64).reshape(-1, 9)], 1)
    cols = np.array(cols, copy=True)
    cols[holes] = np.nan
    filled = np.concatenate([spring_fill(cols[:, k : k + 3], shape, edge_ok=edge_ok) for k in range(0, cols.shape[1], 3)], 1)
    U_out = filled[:, :3].reshape(np.shape(U))
    F_out = None if F is None else filled[:, 3:].reshape(np.shape(F))
    return U_out, F_out
