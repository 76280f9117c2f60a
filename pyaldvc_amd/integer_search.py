"""Integer NCC search of the local DVC step on the GPU.

Keeps the reference's host contract (solver/integer_search.py, al_dvc 1.2.1):

* :func:`ncc_search`            integer_search.py:329-483
* :func:`ncc_search_expanding`  integer_search.py:486-524

The ZNCC maps come from ``aldvc_zncc_map`` and the per-node peak statistics (first-index argmax,
PCE, 27-point quadratic sub-voxel offset) from ``aldvc_ncc_peaks``; only the window clamping and
the assembly of the result dictionary run on the host, as they do in the reference.
"""

from __future__ import annotations

import logging

import numpy as np
import torch

from . import kernels as K

logger = logging.getLogger(__name__)

CLIPPED_EXPAND_FRACTION = 0.05  # integer_search.py:38
DEFAULT_CHUNK_BYTES = 1 << 30  # map memory per chunk of nodes

# 27-point quadratic design (integer_search.py:210-227): columns 1, dx, dy, dz, dx dy, dx dz,
# dy dz, dx^2, dy^2, dz^2 over the 3x3x3 neighbourhood in (dz, dy, dx) C order
_dz, _dy, _dx = (a.ravel().astype(float) for a in np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij"))
DESIGN = np.column_stack([np.ones(27), _dx, _dy, _dz, _dx * _dy, _dx * _dz, _dy * _dz, _dx**2, _dy**2, _dz**2])
PINV = np.ascontiguousarray(np.linalg.pinv(DESIGN))  # (10, 27)


def _empty(N):
    return {
        "disp": np.full((N, 3), np.nan),
        "cc": np.full(N, np.nan),
        "pce": np.full(N, np.nan),
        "clipped": np.zeros(N, dtype=bool),
        "ok": np.zeros(N, dtype=bool),
        "peak": np.zeros((N, 3), dtype=np.int64),  # integer displacement of the map peak (extra key)
    }


def ncc_search(f, g, coords, subset, radius, shift0=None, chunk: int = 64, backend: str = "hip", precision: str = "fp32"):
    """Per-node ZNCC search; returns ``disp`` (N,3) (NaN where not solvable), ``cc``, ``pce``,
    ``clipped`` and ``ok`` exactly as the reference does. ``chunk`` and ``backend`` are accepted
    for signature parity; chunks are sized by map memory."""
    coords = np.asarray(coords, dtype=np.int64)
    nz, ny, nx = (int(s) for s in f.shape)
    N = coords.shape[0]
    hx, hy, hz = (int(w) // 2 for w in subset)
    rx, ry, rz = (int(r) for r in radius)
    sx, sy, sz = 2 * hx + 1, 2 * hy + 1, 2 * hz + 1
    wx, wy, wz = sx + 2 * rx, sy + 2 * ry, sz + 2 * rz
    if shift0 is None:
        shift0 = np.zeros((N, 3), dtype=np.int64)
    shift0 = np.asarray(np.round(shift0), dtype=np.int64)
    res = _empty(N)
    if wx > nx or wy > ny or wz > nz or N == 0:
        return res
    cx, cy, cz = coords[:, 0], coords[:, 1], coords[:, 2]
    t_ok = (cx - hx >= 0) & (cx + hx < nx) & (cy - hy >= 0) & (cy + hy < ny) & (cz - hz >= 0) & (cz + hz < nz)
    dx0 = cx + shift0[:, 0] - hx - rx
    dy0 = cy + shift0[:, 1] - hy - ry
    dz0 = cz + shift0[:, 2] - hz - rz
    wx0 = np.clip(dx0, 0, nx - wx)
    wy0 = np.clip(dy0, 0, ny - wy)
    wz0 = np.clip(dz0, 0, nz - wz)
    free_lo = np.column_stack([dx0 >= 0, dy0 >= 0, dz0 >= 0])
    free_hi = np.column_stack([dx0 <= nx - wx, dy0 <= ny - wy, dz0 <= nz - wz])
    idx_all = np.flatnonzero(t_ok)
    vz, vy, vx = 2 * rz + 1, 2 * ry + 1, 2 * rx + 1
    n_off = vz * vy * vx
    step = max(1, min(idx_all.size, DEFAULT_CHUNK_BYTES // (4 * n_off)))
    fd = K._Args.vol(f)
    gd = K._Args.vol(g)
    for start in range(0, idx_all.size, step):
        ids = idx_all[start : start + step]
        n = ids.size
        maps = K.zncc_direct(fd, gd, torch.from_numpy(coords[ids]), hx, hy, hz, rx, ry, rz,
                             torch.from_numpy(wx0[ids]), torch.from_numpy(wy0[ids]), torch.from_numpy(wz0[ids]),
                             torch.ones(n, dtype=torch.uint8), precision=precision)  # fmt: skip
        flat_tpl = (maps[:, 0, 0, 0] > -1.5).cpu().numpy()  # a flat template's map is all -2
        pk, pv, pce, sub = (t.cpu().numpy() for t in K.ncc_peaks(maps, rx, ry, rz, PINV))
        pz, rem = np.divmod(pk, vy * vx)
        py, px = np.divmod(rem, vx)
        lo, hi = free_lo[ids], free_hi[ids]
        if rx > 0 or ry > 0 or rz > 0:
            on_edge = (
                ((px == 0) & lo[:, 0])
                | ((px == 2 * rx) & hi[:, 0])
                | ((py == 0) & lo[:, 1])
                | ((py == 2 * ry) & hi[:, 1])
                | ((pz == 0) & lo[:, 2])
                | ((pz == 2 * rz) & hi[:, 2])
            )
        else:
            on_edge = np.zeros(n, dtype=bool)
        res["peak"][ids] = np.column_stack([wx0[ids] + px + hx - cx[ids], wy0[ids] + py + hy - cy[ids],
                                            wz0[ids] + pz + hz - cz[ids]])  # fmt: skip
        res["disp"][ids, 0] = wx0[ids] + px + hx - cx[ids] + sub[:, 0]
        res["disp"][ids, 1] = wy0[ids] + py + hy - cy[ids] + sub[:, 1]
        res["disp"][ids, 2] = wz0[ids] + pz + hz - cz[ids] + sub[:, 2]
        res["cc"][ids] = pv
        res["pce"][ids] = pce
        res["clipped"][ids] = on_edge
        res["ok"][ids] = np.isfinite(pv) & flat_tpl
    res["disp"][~res["ok"]] = np.nan
    return res


def ncc_search_expanding(f, g, coords, subset, radius, shift0=None, auto_expand=True, max_expand=3, chunk=64, backend="hip", precision="fp32"):
    """:func:`ncc_search` plus radius doubling for clipped peaks (integer_search.py:486-524)."""
    res = ncc_search(f, g, coords, subset, radius, shift0, chunk=chunk, precision=precision)
    res["expansions"] = 0
    res["radius"] = tuple(int(r) for r in radius)
    rad = res["radius"]
    coords = np.asarray(coords, dtype=np.int64)
    for _ in range(max_expand if auto_expand else 0):
        sel = res["clipped"] & res["ok"]
        n_ok = int(res["ok"].sum())
        if n_ok == 0 or sel.sum() <= CLIPPED_EXPAND_FRACTION * n_ok:
            break
        rad = tuple(2 * r for r in rad)
        logger.info("NCC: %d/%d peaks clipped; re-searching them with radius %s", int(sel.sum()), n_ok, rad)
        sub_shift = None if shift0 is None else np.asarray(shift0)[sel]
        sub = ncc_search(f, g, coords[sel], subset, rad, sub_shift, chunk=chunk, precision=precision)
        if not sub["ok"].any():
            break
        idx = np.flatnonzero(sel)[sub["ok"]]
        for key in ("disp", "cc", "pce", "clipped", "ok", "peak"):
            res[key][idx] = sub[key][sub["ok"]]
        res["expansions"] += 1
        res["radius"] = rad
    return res
