"""Initial-guess orchestration on the GPU: block-mean pyramid, phase-correlation shift, coarse-to-fine
NCC search, and the cleaning of the guess (median test + spring inpainting).

The reference runs these on the host around its GPU NCC search (solver/integer_search.py:235-694,
solver/init_disp.py:53-189). Here the volumes stay on the device: the pyramid levels come from
``aldvc_block_downsample``, the search is ``integer_search.ncc_search_expanding`` (HIP maps and the
fused peak kernel), the global shift is a phase correlation with the FFT on the device, and the
cleaning uses the device median test and spring fill of the local step (``admm.subpb1_cleanup``,
``admm.spring_fill``) up to ``SPRING_MAX_UNKNOWN`` nodes, the host restatement beyond (as the
local step does). Only node-sized arrays move to the host.
"""

from __future__ import annotations

import logging
import math
import time

import numpy as np
import torch

from . import kernels as K
from ._lib import check, lib
from .cleanup import SPRING_MAX_UNKNOWN, fill_nan_grid, universal_median_test
from .integer_search import ncc_search_expanding

logger = logging.getLogger(__name__)

# integer_search.py:33-38, init_disp.py:25-26
PYRAMID_FINE_RADIUS = 2
COARSE_RADIUS_MARGIN = 1
MIN_COARSE_RADIUS = 2
MIN_PYRAMID_SUBSET = 8
DEFAULT_INIT_SUBSET = 16
MIN_COARSE_STD_RATIO = 0.35
SPLIT_GUESS_MAX_KEEP = 0.9
SPLIT_GUESS_MIN_NEIGHBOURS = 2
MIN_COARSE_NODES_PER_AXIS = 2  # coarse_init.py:30


def _device_f32(vol) -> torch.Tensor:
    if isinstance(vol, torch.Tensor):
        return vol.to(device=K.device(), dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(vol, dtype=np.float32)).to(K.device())


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
    fs = fs - fs.mean()
    gs = gs - gs.mean()
    win = torch.ones_like(fs)
    for ax, n in enumerate(fs.shape):
        w1 = torch.from_numpy(np.hanning(n)).to(fs)  # the reference's window, float64 here
        shape = [1, 1, 1]
        shape[ax] = n
        win = win * w1.reshape(shape)
    Ff = torch.fft.fftn(fs * win)
    Fg = torch.fft.fftn(gs * win)
    R = Fg * torch.conj(Ff)
    mag = torch.abs(R)
    R = R / torch.where(mag > 1e-12, mag, torch.ones_like(mag))
    r = torch.real(torch.fft.ifftn(R))
    peak = np.unravel_index(int(torch.argmax(r).item()), tuple(r.shape))
    shift = np.array(peak, dtype=np.float64)
    for ax, n in enumerate(r.shape):
        if shift[ax] > n // 2:
            shift[ax] -= n
    dz, dy, dx = shift * factor  # r peaks at the shift that maps f onto g: g(x) = f(x - shift)
    return np.array([dx, dy, dz], dtype=np.float64)


def coarse_radius(radius, fac: int) -> tuple[int, int, int]:
    """Search radius at a coarse level (integer_search.py:566-572)."""
    return tuple(max(MIN_COARSE_RADIUS, -(-int(r) // fac) + COARSE_RADIUS_MARGIN) for r in radius)


def _std(vol: torch.Tensor) -> float:
    from .volumes import voi_mean_std

    return float(voi_mean_std(vol)[1])


def auto_pyramid_levels(shape, subset, radius, max_levels: int = 3, f=None, min_std_ratio: float = MIN_COARSE_STD_RATIO) -> int:
    """Number of coarse levels that fit the search window and keep texture (integer_search.py:575-614):
    a level needs the coarse template + search window + margin inside the volume and, when ``f`` is
    given, at least ``min_std_ratio`` of the full-resolution standard deviation."""
    levels = 0
    std0 = _std(_device_f32(f)) if f is not None else None
    for lv in range(1, max_levels + 1):
        fac = 2**lv
        fits = True
        for n, w, r in zip(tuple(shape)[::-1], subset, coarse_radius(radius, fac)):
            w_l = max(MIN_PYRAMID_SUBSET, (w // fac) // 2 * 2)
            need = w_l + 1 + 2 * max(r, 1) + 4
            if n // fac < need:
                fits = False
        if fits and std0 is not None and std0 > 0:
            if _std(block_downsample(f, fac)) < min_std_ratio * std0:
                fits = False
        if fits:
            levels = lv
        else:
            break
    return levels


def _scatter_search(sub: dict, active: np.ndarray, n: int) -> dict:
    """Search results of the ``active`` nodes as full-size arrays (integer_search.py:527-542)."""
    if active.size == n:
        return sub
    res = {"disp": np.full((n, 3), np.nan), "cc": np.full(n, np.nan), "pce": np.full(n, np.nan),
           "clipped": np.zeros(n, dtype=bool), "ok": np.zeros(n, dtype=bool),
           "expansions": sub.get("expansions", 0), "radius": sub.get("radius")}  # fmt: skip
    for key in ("disp", "cc", "pce", "clipped", "ok"):
        res[key][active] = sub[key]
    return res


def _median_flags(d: np.ndarray, valid: np.ndarray, grid_shape, threshold: float, edge_ok=None) -> np.ndarray:
    """Universal median test flags of the ``valid`` nodes of ``d`` (N, 3) on the node grid, the
    invalid nodes spring-filled first (outlier_detection.py:135-160); on the device up to
    SPRING_MAX_UNKNOWN nodes."""
    N = d.shape[0]
    if N <= SPRING_MAX_UNKNOWN:
        from .admm import subpb1_cleanup

        status = np.where(valid, 0, 1).astype(np.int8)
        _, bad = subpb1_cleanup(np.nan_to_num(d), np.nan_to_num(d), status, grid_shape, float(threshold), edge_ok=edge_ok)
        return bad & valid
    flag = universal_median_test(d.reshape(tuple(grid_shape) + (3,)), valid.reshape(grid_shape), threshold, edge_ok=edge_ok)
    return flag.ravel() & valid


def _fill(d: np.ndarray, grid_shape, edge_ok=None) -> np.ndarray:
    """Spring inpainting of the NaN entries of ``d`` (N, 3), column by column (inpaint.py:35-121):
    on the device up to SPRING_MAX_UNKNOWN unknowns, the host restatement beyond."""
    if int(np.isnan(d).sum(axis=0).max()) <= SPRING_MAX_UNKNOWN:
        from .admm import spring_fill

        return spring_fill(d, grid_shape, edge_ok=edge_ok)
    return np.stack([fill_nan_grid(d[:, c].reshape(grid_shape), edge_ok=edge_ok).ravel() for c in range(3)], 1)


def _clean_field(disp: np.ndarray, grid_shape, threshold: float) -> np.ndarray:
    """Median-test outlier removal + spring inpainting on the node grid (integer_search.py:545-556)."""
    d = np.array(disp, dtype=np.float64).reshape(-1, 3)
    valid = np.all(np.isfinite(d), axis=1)
    if threshold > 0 and valid.sum() > 27:
        flag = _median_flags(d, valid, grid_shape, threshold)
        d[flag] = np.nan
    return _fill(d, grid_shape)


def pyramid_search(f, g, coords, grid_shape, subset, radius, levels: int = 0, global_shift=None,
                   outlier_threshold: float = 2.0, auto_expand: bool = True, max_expand: int = 3,
                   fine_radius: int = PYRAMID_FINE_RADIUS, backend: str = "hip", valid=None,
                   precision: str = "fp32") -> dict:  # fmt: skip
    """Coarse-to-fine NCC search (integer_search.py:617-694): ``disp`` (N, 3) in full-resolution
    voxels + info. The coarsest level searches ``radius`` scaled to its voxels, the finer levels
    refine with ``fine_radius`` around the up-scaled estimate, and the estimate is median-cleaned
    and inpainted between levels. ``valid`` restricts the correlation to those nodes."""
    if backend != "hip":
        raise ValueError("pyramid_search here runs on the HIP backend")
    fd, gd = _device_f32(f), _device_f32(g)
    coords = np.asarray(coords, dtype=np.int64)
    grid_shape = tuple(int(v) for v in grid_shape)
    if levels <= 0:
        levels = auto_pyramid_levels(tuple(fd.shape), subset, radius, f=fd)
        logger.info("Pyramid: %d coarse level(s) selected automatically", levels)
    N = coords.shape[0]
    active = np.arange(N) if valid is None else np.flatnonzero(np.asarray(valid, dtype=bool))
    if active.size == 0:
        active = np.arange(N)
    disp_full = np.zeros((N, 3), dtype=np.float64)
    if global_shift is not None:
        disp_full += np.asarray(global_shift, dtype=np.float64)[None, :]
    info: dict = {"levels": levels, "per_level": [], "expansions": 0, "n_searched": int(active.size)}
    for lv in range(levels, -1, -1):
        fac = 2**lv
        fl = block_downsample(fd, fac) if fac > 1 else fd
        gl = block_downsample(gd, fac) if fac > 1 else gd
        subset_l = tuple(max(MIN_PYRAMID_SUBSET, (w // fac) // 2 * 2) for w in subset)
        rad_l = coarse_radius(radius, fac) if lv == levels else (int(fine_radius),) * 3
        coords_l = np.round(coords[active] / fac).astype(np.int64)
        shift_l = np.round(disp_full[active] / fac).astype(np.int64)
        sub = ncc_search_expanding(fl, gl, coords_l, subset_l, rad_l, shift_l, auto_expand, max_expand, precision=precision)
        res = _scatter_search(sub, active, N)
        rad_l = res["radius"]
        info["expansions"] += res["expansions"]
        d = res["disp"] * fac
        d[~res["ok"]] = np.nan
        n_ok = int(res["ok"].sum())
        if n_ok == 0:
            logger.warning("Pyramid level %d produced no valid NCC peaks; keeping previous estimate.", lv)
            info["per_level"].append({"level": lv, "n_ok": 0, "radius": rad_l, "subset": subset_l})
            continue
        disp_full = _clean_field(d, grid_shape, outlier_threshold)
        info["per_level"].append({"level": lv, "n_ok": n_ok, "radius": rad_l, "subset": subset_l,
                                  "median_cc": float(np.nanmedian(res["cc"])),
                                  "clipped_frac": float(np.mean(res["clipped"][res["ok"]]))})  # fmt: skip
        info["cc"], info["pce"], info["ok"], info["clipped"] = res["cc"], res["pce"], res["ok"], res["clipped"]
    info["disp"] = disp_full
    return info


def same_side_neighbours(good, edge_ok, grid_shape) -> np.ndarray:
    """``(N,)`` count of the six direct neighbours that are ``good`` and not across a cut edge
    (init_disp.py:29-41)."""
    g = np.asarray(good, dtype=bool).reshape(grid_shape)
    eo = np.asarray(edge_ok, dtype=bool).reshape(tuple(grid_shape) + (3,))
    cnt = np.zeros(grid_shape, dtype=np.int64)
    cnt[:, :, :-1] += g[:, :, 1:] & eo[:, :, :-1, 0]
    cnt[:, :, 1:] += g[:, :, :-1] & eo[:, :, :-1, 0]
    cnt[:, :-1, :] += g[:, 1:, :] & eo[:, :-1, :, 1]
    cnt[:, 1:, :] += g[:, :-1, :] & eo[:, :-1, :, 1]
    cnt[:-1, :, :] += g[1:, :, :] & eo[:-1, :, :, 2]
    cnt[1:, :, :] += g[:-1, :, :] & eo[:-1, :, :, 2]
    return cnt.ravel()


def clean_initial_guess(disp, ok, pce, mesh, para, split_fraction=None):
    """Reject failed / low-quality / outlying nodes and inpaint them (init_disp.py:144-189):
    ``(U0 (N, 3), bad (N,))``. Split nodes that lost more than ``1 - SPLIT_GUESS_MAX_KEEP`` of their
    subset take their guess from their own side when enough same-side neighbours are usable."""
    N = mesh.n_nodes
    d = np.array(disp, dtype=np.float64).reshape(N, 3)
    bad = ~np.asarray(ok, dtype=bool) | ~np.all(np.isfinite(d), axis=1)
    if mesh.node_valid.size == N:
        bad |= ~mesh.node_valid
    if para.init_min_pce > 0 and pce is not None:
        p = np.asarray(pce, dtype=np.float64)
        bad |= np.isfinite(p) & (p < para.init_min_pce)
    d[bad] = np.nan
    grid = tuple(mesh.grid_shape)
    edge_ok = mesh.edges_ok_grid()
    if para.init_outlier_threshold > 0 and (~bad).sum() > 27:
        bad |= _median_flags(d, ~bad, grid, para.init_outlier_threshold, edge_ok=edge_ok)
        d[bad] = np.nan
    if split_fraction is not None and edge_ok is not None:
        sf = np.asarray(split_fraction, dtype=np.float64)
        cut = np.isfinite(sf) & (sf < SPLIT_GUESS_MAX_KEEP) & ~bad
        if cut.any():
            n_side = same_side_neighbours(~bad, mesh.edges_ok(), grid)
            take = cut & (n_side >= SPLIT_GUESS_MIN_NEIGHBOURS)
            bad |= take
            d[take] = np.nan
            logger.info("Initial guess: %d split nodes take their guess from their own side", int(take.sum()))
    if bad.all():
        logger.warning("Initial guess failed at every node; starting from zero displacement.")
        return np.zeros((N, 3)), bad
    out = _fill(d, grid, edge_ok=edge_ok)
    logger.info("Initial guess: %d/%d nodes replaced by inpainting (%.1f%%)", int(bad.sum()), N, 100 * bad.mean())
    return out, bad


def compute_initial_guess(f, g, mesh, para, previous=None, split_fraction=None):
    """``U0`` (N, 3) and an info dict according to ``para.init_guess_method`` (init_disp.py:53-141):
    "zero", "previous" (falls back to "pyramid"), "pyramid", or the direct expanding NCC search."""
    N = mesh.n_nodes
    method = para.init_guess_method
    if method == "zero":
        return np.zeros((N, 3)), {"method": "zero"}
    if method == "previous":
        if previous is not None and previous.shape == (N, 3) and np.all(np.isfinite(previous)):
            return previous.copy(), {"method": "previous"}
        method = "pyramid"
    coords = np.round(mesh.coordinates).astype(np.int64)
    subset = para.init_subset if para.init_subset is not None else tuple(min(int(w), DEFAULT_INIT_SUBSET) for w in para.winsize)
    valid = mesh.node_valid if mesh.node_valid.size == N else None
    shift = None
    if para.global_shift:
        shift = phase_correlation_shift(f, g, para.voi)
        logger.info("Global phase-correlation shift (dx, dy, dz) = %s voxels", np.array2string(shift, precision=1))
    if method == "pyramid":
        info = pyramid_search(f, g, coords, mesh.grid_shape, subset, para.search_radius, para.pyramid_levels,
                              global_shift=shift, outlier_threshold=para.init_outlier_threshold,
                              auto_expand=para.ncc_auto_expand, max_expand=para.ncc_max_expand,
                              fine_radius=int(getattr(para, "pyramid_fine_radius", PYRAMID_FINE_RADIUS)), valid=valid)  # fmt: skip
        disp = info["disp"]
        ok = info.get("ok", np.ones(N, dtype=bool))
        pce = info.get("pce", np.full(N, np.nan))
        info["method"] = "pyramid"
    else:
        shift0 = None if shift is None else np.tile(np.round(shift).astype(np.int64), (N, 1))
        active = np.arange(N) if valid is None or not valid.any() else np.flatnonzero(valid)
        res = ncc_search_expanding(f, g, coords[active], subset, tuple(int(r) for r in para.search_radius),
                                   None if shift0 is None else shift0[active], para.ncc_auto_expand,
                                   para.ncc_max_expand)  # fmt: skip
        res = _scatter_search(res, active, N)
        disp, ok, pce = res["disp"], res["ok"], res["pce"]
        info = {"method": "ncc", "radius": res["radius"], "expansions": res["expansions"], "cc": res["cc"],
                "pce": pce, "ok": ok, "clipped": res["clipped"]}  # fmt: skip
    U0, bad = clean_initial_guess(disp, ok, pce, mesh, para, split_fraction)
    info["n_bad"] = int(bad.sum())
    info["global_shift"] = shift
    return U0, info


# ---------------------------------------------------------------------------- coarse-lattice start
def _axes(mesh):
    """The lattice axes ``(z0, y0, x0)`` of a full node grid."""
    nz, ny, nx = (int(v) for v in mesh.grid_shape)
    c = np.asarray(mesh.coordinates, dtype=np.float64).reshape(nz, ny, nx, 3)
    return c[:, 0, 0, 2].copy(), c[0, :, 0, 1].copy(), c[0, 0, :, 0].copy()


def coarse_lattice(mesh, factor: int):
    """Every ``factor``-th node per axis as a mesh of its own, and the fine index of each of its
    nodes (coarse_init.py:33-48); ``None`` when fewer than two nodes per axis would remain."""
    from .local import NodeMesh

    k = int(factor)
    if k <= 1:
        return None
    sel = [np.arange(0, int(n), k) for n in mesh.grid_shape]
    if min(a.size for a in sel) < MIN_COARSE_NODES_PER_AXIS:
        return None
    nz, ny, nx = (int(v) for v in mesh.grid_shape)
    iz, iy, ix = np.meshgrid(*sel, indexing="ij")
    fine_index = (iz * ny * nx + iy * nx + ix).ravel()
    coords = np.asarray(mesh.coordinates, dtype=np.float64)[fine_index]
    valid = np.asarray(mesh.node_valid, dtype=bool)
    node_valid = valid[fine_index].copy() if valid.size == mesh.n_nodes else np.ones(fine_index.size, dtype=bool)
    return NodeMesh(coords, (sel[0].size, sel[1].size, sel[2].size), node_valid), fine_index


def _nearest_fill(flat: np.ndarray) -> np.ndarray:
    """Non-finite coarse nodes take the values of the nearest finite one (coarse_init.py:70-78)."""
    from scipy.ndimage import distance_transform_edt

    bad = ~np.all(np.isfinite(flat), axis=-1)
    if bad.all():
        return np.nan_to_num(flat)
    _, idx = distance_transform_edt(bad, return_distances=True, return_indices=True)
    return flat[tuple(idx)]


def interpolate_to_mesh(coarse, values, mesh) -> np.ndarray:
    """Trilinear interpolation of per-node values ``(Nc, ...)`` from the coarse lattice to every node
    of ``mesh``, linear extrapolation beyond the lattice, non-finite coarse values filled from the
    nearest finite node first (coarse_init.py:51-67). On the device in float64."""
    v = np.asarray(values, dtype=np.float64)
    gz, gy, gx = (int(n) for n in coarse.grid_shape)
    flat = v.reshape(gz, gy, gx, -1)
    if not np.all(np.isfinite(flat)):
        flat = _nearest_fill(flat)
    dev = K.device()
    table = torch.from_numpy(np.ascontiguousarray(flat)).to(dev)
    pts = torch.from_numpy(np.asarray(mesh.coordinates, dtype=np.float64)).to(dev)
    idx, frac = [], []
    for axis, col in zip(_axes(coarse), (2, 1, 0)):
        a = torch.from_numpy(axis).to(dev)
        x = pts[:, col].contiguous()
        i = torch.clamp(torch.searchsorted(a, x, right=True) - 1, 0, a.numel() - 2)
        idx.append(i)
        frac.append((x - a[i]) / (a[i + 1] - a[i]))  # outside [0, 1] beyond the lattice: extrapolation
    (iz, iy, ix), (tz, ty, tx) = idx, frac
    out = torch.zeros((pts.shape[0], flat.shape[-1]), dtype=torch.float64, device=dev)
    for dz in (0, 1):
        wz = tz if dz else 1.0 - tz
        for dy in (0, 1):
            wy = ty if dy else 1.0 - ty
            for dx in (0, 1):
                wx = tx if dx else 1.0 - tx
                out += (wz * wy * wx)[:, None] * table[iz + dz, iy + dy, ix + dx]
    return out.cpu().numpy().reshape((mesh.n_nodes,) + v.shape[1:])


def coarse_initial_guess(ref, g_norm, g_prep, mesh, para):
    """``(U0, F0, info)`` for every node from the solve on every ``para.init_coarse_factor``-th node:
    the initial guess and the 12-DOF local step on the coarse lattice, U and F interpolated to all
    nodes (coarse_init.py:81-128). Falls back to :func:`compute_initial_guess` (``F0 = None``)
    when the coarse lattice would have fewer than two nodes per axis."""
    from .local import local_icgn, precompute_local_context

    t0 = time.perf_counter()
    lattice = coarse_lattice(mesh, int(para.init_coarse_factor))
    if lattice is None:
        U0, info = compute_initial_guess(ref.f, g_norm, mesh, para)
        info["coarse_factor"] = 1
        return U0, None, info
    coarse, _fine_index = lattice
    U0c, info = compute_initial_guess(ref.f, g_norm, coarse, para)
    t_ncc = time.perf_counter() - t0
    ctx_c = precompute_local_context(coarse, ref, para)
    Uc, Fc, info_c, bad_c = local_icgn(ctx_c, ref, g_prep, U0c, para, coarse)
    U0 = interpolate_to_mesh(coarse, Uc, mesh)
    F0 = interpolate_to_mesh(coarse, np.asarray(Fc).reshape(coarse.n_nodes, 9), mesh).reshape(mesh.n_nodes, 3, 3)
    dt = time.perf_counter() - t0
    info.update({"method": f"coarse x{int(para.init_coarse_factor)} + {info.get('method', 'ncc')}",
                 "coarse_factor": int(para.init_coarse_factor), "coarse_nodes": int(coarse.n_nodes),
                 "coarse_converged": float(np.mean(info_c.status == 0)), "coarse_bad": int(bad_c.sum()),
                 "coarse_ncc_time": t_ncc, "coarse_icgn_time": float(info_c.solve_time), "coarse_total_time": dt})  # fmt: skip
    logger.info("Coarse initial guess: %d of %d nodes solved (%.1f%% converged, %d filled), %.2fs NCC + %.2fs IC-GN",
                coarse.n_nodes, mesh.n_nodes, 100.0 * info["coarse_converged"], int(bad_c.sum()), t_ncc,
                info_c.solve_time)  # fmt: skip
    return U0, F0, info
