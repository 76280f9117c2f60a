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
    dev = fd.device
    half = torch.tensor([hx, hy, hz], dtype=torch.int64, device=dev)
    top = torch.tensor([2 * rx, 2 * ry, 2 * rz], dtype=torch.int64, device=dev)
    win0 = np.column_stack([wx0, wy0, wz0])
    free = np.concatenate([free_lo, free_hi], axis=1)
    for start in range(0, idx_all.size, step):
        ids = idx_all[start : start + step]
        n = ids.size
        # contiguous node ranges (every template inside the volume) are written through slices
        sel = slice(int(ids[0]), int(ids[0]) + n) if int(ids[-1]) - int(ids[0]) + 1 == n else ids
        cd = torch.from_numpy(coords[sel]).to(dev)
        w0 = torch.from_numpy(win0[sel]).to(dev)
        fr = torch.from_numpy(free[sel]).to(dev)
        maps = K.zncc_direct(fd, gd, cd, hx, hy, hz, rx, ry, rz, w0[:, 0].contiguous(), w0[:, 1].contiguous(),
                             w0[:, 2].contiguous(), torch.ones(n, dtype=torch.uint8, device=dev), precision=precision)  # fmt: skip
        flat_ok = maps[:, 0, 0, 0] > -1.5  # a flat template's map is all -2
        pk, pv, pce, sub = K.ncc_peaks(maps, rx, ry, rz, PINV)
        # the result assembly on the device: map peak (x, y, z) -> integer and sub-voxel displacement,
        # clipped flag (peak on a face of the window that the volume did not clamp), ok flag
        pz, rem = pk // (vy * vx), pk % (vy * vx)
        pidx = torch.stack([rem % vx, rem // vx, pz], 1)
        peak = w0 + pidx + half - cd
        if rx > 0 or ry > 0 or rz > 0:
            on_edge = ((pidx == 0) & fr[:, :3]).any(1) | ((pidx == top) & fr[:, 3:]).any(1)
        else:
            on_edge = torch.zeros(n, dtype=torch.bool, device=dev)
        ok = torch.isfinite(pv) & flat_ok
        packed = torch.cat([peak.double() + sub, peak.double(), pv[:, None], pce[:, None], on_edge.double()[:, None],
                            ok.double()[:, None]], 1).cpu().numpy()  # fmt: skip
        res["disp"][sel] = packed[:, 0:3]
        res["peak"][sel] = packed[:, 3:6].astype(np.int64)
        res["cc"][sel] = packed[:, 6]
        res["pce"][sel] = packed[:, 7]
        res["clipped"][sel] = packed[:, 8] > 0
        res["ok"][sel] = packed[:, 9] > 0
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
