"""ORACLE -- test infrastructure only, never product code.

NumPy restatement of the reference's integer NCC search (solver/integer_search.py) on top of
the float64 direct ZNCC map of ``cpu_ref.zncc_direct``:

* ``DESIGN`` / ``PINV``   the 27-point quadratic design matrix and its pseudo-inverse
                          (integer_search.py:210-227)
* ``subvoxel_peak``       integer_search.py:297-321
* ``peak_stats``          the per-node peak logic of ncc_search (integer_search.py:444-477):
                          first-index argmax in C order, PCE, clipped flag, sub-voxel offset
* ``ncc_search``          integer_search.py:329-483 with the direct engine (the path the
                          reference takes whenever offsets x template <= 3e7, :395)
* ``ncc_search_expanding`` integer_search.py:486-524
* ``block_downsample``     integer_search.py:235-243
* ``phase_correlation_shift`` integer_search.py:246-294 (numpy FFT)
* ``coarse_radius`` / ``auto_pyramid_levels`` integer_search.py:566-614
* ``pyramid_search``       integer_search.py:617-694; the between-level cleaning (median test +
                          spring inpainting, integer_search.py:545-556) is passed in as ``clean``
                          so that a test can supply a separately pinned implementation

Pinned by closed forms in tests/test_oracle_init_guess.py: block means of known arrays, an exact
cyclic shift recovered by the phase correlation, the coarse-radius and level rules on hand-worked
shapes, and a rigid integer translation recovered exactly by the pyramid.
"""

from __future__ import annotations

import numpy as np

from . import cpu_ref

CLIPPED_EXPAND_FRACTION = 0.05  # integer_search.py:38

_dz, _dy, _dx = np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij")
_dx, _dy, _dz = _dx.ravel().astype(float), _dy.ravel().astype(float), _dz.ravel().astype(float)
DESIGN = np.column_stack([np.ones(27), _dx, _dy, _dz, _dx * _dy, _dx * _dz, _dy * _dz, _dx**2, _dy**2, _dz**2])
PINV = np.linalg.pinv(DESIGN)  # (10, 27)
NB_DZ, NB_DY, NB_DX = (a.ravel()[None, :] for a in np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij"))


def subvoxel_peak(neigh):
    """Quadratic sub-voxel refinement of ``(n, 27)`` neighbourhoods (integer_search.py:297-321)."""
    n = neigh.shape[0]
    coef = neigh @ PINV.T
    a = coef[:, 1:4]
    H = np.empty((n, 3, 3))
    H[:, 0, 0] = 2 * coef[:, 7]
    H[:, 1, 1] = 2 * coef[:, 8]
    H[:, 2, 2] = 2 * coef[:, 9]
    H[:, 0, 1] = H[:, 1, 0] = coef[:, 4]
    H[:, 0, 2] = H[:, 2, 0] = coef[:, 5]
    H[:, 1, 2] = H[:, 2, 1] = coef[:, 6]
    out = np.zeros((n, 3))
    det = np.linalg.det(H)
    ok = np.abs(det) > 1e-12
    if ok.any():
        try:
            sol = np.linalg.solve(H[ok], -a[ok][:, :, None])[:, :, 0]
            good = np.all(np.abs(sol) <= 1.0, axis=1) & np.all(np.isfinite(sol), axis=1)
            tmp = np.zeros((int(ok.sum()), 3))
            tmp[good] = sol[good]
            out[ok] = tmp
        except np.linalg.LinAlgError:
            pass
    return out


def peak_stats(ncc, rx, ry, rz, free_lo, free_hi):
    """``(pk, px, py, pz, peak_val, pce, on_edge, sub)`` for ``ncc`` (n, vz, vy, vx) (integer_search.py:444-473)."""
    n = ncc.shape[0]
    flat = ncc.reshape(n, -1)
    pk = np.argmax(flat, axis=1)
    pz, py, px = np.unravel_index(pk, ncc.shape[1:])
    peak_val = flat[np.arange(n), pk]
    cmin = flat - flat.min(axis=1, keepdims=True)
    energy = np.mean(cmin * cmin, axis=1)
    pk_min = cmin[np.arange(n), pk]
    pce = np.divide(pk_min * pk_min, energy, out=np.zeros(n), where=energy > 1e-30)
    lo, hi = free_lo, free_hi
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
    sub = np.zeros((n, 3))
    at_border = (px == 0) | (px == 2 * rx) | (py == 0) | (py == 2 * ry) | (pz == 0) | (pz == 2 * rz)
    interior = ~at_border & (rx > 0) & (ry > 0) & (rz > 0)
    if interior.any():
        ii = np.flatnonzero(interior)
        neigh = ncc[ii[:, None], pz[ii][:, None] + NB_DZ, py[ii][:, None] + NB_DY, px[ii][:, None] + NB_DX]
        sub[ii] = subvoxel_peak(neigh)
    return pk, px, py, pz, peak_val, pce, on_edge, sub


def window_starts(coords, shape, half, radius, shift0):
    """Clamped window starts and the free-side flags (integer_search.py:374-386)."""
    nz, ny, nx = shape
    hx, hy, hz = half
    rx, ry, rz = radius
    wx, wy, wz = 2 * hx + 1 + 2 * rx, 2 * hy + 1 + 2 * ry, 2 * hz + 1 + 2 * rz
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
    return t_ok, wx0, wy0, wz0, free_lo, free_hi


def ncc_search(f, g, coords, subset, radius, shift0=None, chunk=4096, return_maps=False):
    """Direct-engine ``ncc_search`` in float64 (integer_search.py:329-483)."""
    f = np.ascontiguousarray(f, dtype=np.float32)
    g = np.ascontiguousarray(g, dtype=np.float32)
    coords = np.asarray(coords, dtype=np.int64)
    nz, ny, nx = f.shape
    N = coords.shape[0]
    hx, hy, hz = (int(w) // 2 for w in subset)
    rx, ry, rz = (int(r) for r in radius)
    sx, sy, sz = 2 * hx + 1, 2 * hy + 1, 2 * hz + 1
    wx, wy, wz = sx + 2 * rx, sy + 2 * ry, sz + 2 * rz
    if shift0 is None:
        shift0 = np.zeros((N, 3), dtype=np.int64)
    shift0 = np.asarray(np.round(shift0), dtype=np.int64)
    disp = np.full((N, 3), np.nan)
    cc = np.full(N, np.nan)
    pce = np.full(N, np.nan)
    clipped = np.zeros(N, dtype=bool)
    ok = np.zeros(N, dtype=bool)
    vz, vy, vx = 2 * rz + 1, 2 * ry + 1, 2 * rx + 1
    maps = np.full((N, vz, vy, vx), np.nan) if return_maps else None
    res = {"disp": disp, "cc": cc, "pce": pce, "clipped": clipped, "ok": ok}
    if wx > nx or wy > ny or wz > nz:
        if return_maps:
            res["maps"] = maps
        return res
    t_ok, wx0, wy0, wz0, free_lo, free_hi = window_starts(coords, f.shape, (hx, hy, hz), (rx, ry, rz), shift0)
    cx, cy, cz = coords[:, 0], coords[:, 1], coords[:, 2]
    idx_all = np.flatnonzero(t_ok)
    for start in range(0, idx_all.size, chunk):
        ids = idx_all[start : start + chunk]
        n = ids.size
        ncc = np.empty((n, vz, vy, vx))
        cpu_ref.zncc_direct(f, g, coords[ids], hx, hy, hz, rx, ry, rz, wx0[ids], wy0[ids], wz0[ids], np.ones(n, dtype=bool), ncc)
        t_norm = np.where(ncc[:, 0, 0, 0] > -1.5, 1.0, 0.0)
        ncc[ncc < -1.0] = 0.0
        if return_maps:
            maps[ids] = ncc
        pk, px, py, pz, peak_val, pce_k, on_edge, sub = peak_stats(ncc, rx, ry, rz, free_lo[ids], free_hi[ids])
        disp[ids, 0] = wx0[ids] + px + hx - cx[ids] + sub[:, 0]
        disp[ids, 1] = wy0[ids] + py + hy - cy[ids] + sub[:, 1]
        disp[ids, 2] = wz0[ids] + pz + hz - cz[ids] + sub[:, 2]
        cc[ids] = peak_val
        pce[ids] = pce_k
        clipped[ids] = on_edge
        ok[ids] = np.isfinite(peak_val) & (t_norm > 1e-10)
    disp[~ok] = np.nan
    if return_maps:
        res["maps"] = maps
    return res


def ncc_search_expanding(f, g, coords, subset, radius, shift0=None, auto_expand=True, max_expand=3):
    """integer_search.py:486-524."""
    res = ncc_search(f, g, coords, subset, radius, shift0)
    res["expansions"] = 0
    res["radius"] = tuple(int(r) for r in radius)
    rad = res["radius"]
    for _ in range(max_expand if auto_expand else 0):
        sel = res["clipped"] & res["ok"]
        n_ok = int(res["ok"].sum())
        if n_ok == 0 or sel.sum() <= CLIPPED_EXPAND_FRACTION * n_ok:
            break
        rad = tuple(2 * r for r in rad)
        sub_shift = None if shift0 is None else np.asarray(shift0)[sel]
        sub = ncc_search(f, g, np.asarray(coords)[sel], subset, rad, sub_shift)
        if not sub["ok"].any():
            break
        idx = np.flatnonzero(sel)[sub["ok"]]
        for key in ("disp", "cc", "pce", "clipped", "ok"):
            res[key][idx] = sub[key][sub["ok"]]
        res["expansions"] += 1
        res["radius"] = rad
    return res


# ---------------------------------------------------------------------------- coarse-to-fine search
PYRAMID_FINE_RADIUS = 2  # integer_search.py:33-36
COARSE_RADIUS_MARGIN = 1
MIN_COARSE_RADIUS = 2
MIN_PYRAMID_SUBSET = 8
MIN_COARSE_STD_RATIO = 0.35  # integer_search.py:563


def block_downsample(vol, factor):
    """integer_search.py:235-243 (NumPy's float32 mean over the blocks)."""
    v = np.asarray(vol, dtype=np.float32)
    if factor <= 1:
        return np.ascontiguousarray(v)
    nz, ny, nx = (s // factor * factor for s in v.shape)
    v = v[:nz, :ny, :nx]
    out = v.reshape(nz // factor, factor, ny // factor, factor, nx // factor, factor).mean(axis=(1, 3, 5), dtype=np.float32)
    return np.ascontiguousarray(out, dtype=np.float32)


def phase_correlation_shift(f, g, max_dim=128):
    """integer_search.py:246-294 (whole volume): ``(dx, dy, dz)``."""
    f = np.asarray(f, dtype=np.float32)
    g = np.asarray(g, dtype=np.float32)
    factor = max(1, int(np.ceil(max(f.shape) / max_dim)))
    fs = block_downsample(f, factor).astype(np.float64)
    gs = block_downsample(g, factor).astype(np.float64)
    fs -= fs.mean()
    gs -= gs.mean()
    win = np.ones(fs.shape)
    for ax, n in enumerate(fs.shape):
        shape = [1, 1, 1]
        shape[ax] = n
        win = win * np.hanning(n).reshape(shape)
    R = np.fft.fftn(gs * win) * np.conj(np.fft.fftn(fs * win))
    mag = np.abs(R)
    R = R / np.where(mag > 1e-12, mag, 1.0)
    r = np.real(np.fft.ifftn(R))
    shift = np.array(np.unravel_index(int(np.argmax(r)), r.shape), dtype=np.float64)
    for ax, n in enumerate(r.shape):
        if shift[ax] > n // 2:
            shift[ax] -= n
    dz, dy, dx = shift * factor
    return np.array([dx, dy, dz])


def coarse_radius(radius, fac):
    """integer_search.py:566-572."""
    return tuple(max(MIN_COARSE_RADIUS, -(-int(r) // fac) + COARSE_RADIUS_MARGIN) for r in radius)


def auto_pyramid_levels(shape, subset, radius, max_levels=3, f=None, min_std_ratio=MIN_COARSE_STD_RATIO):
    """integer_search.py:575-614 (standard deviations in float64)."""
    levels = 0
    std0 = float(np.std(np.asarray(f, dtype=np.float64))) if f is not None else None
    for lv in range(1, max_levels + 1):
        fac = 2**lv
        fits = True
        for n, w, r in zip(tuple(shape)[::-1], subset, coarse_radius(radius, fac)):
            w_l = max(MIN_PYRAMID_SUBSET, (w // fac) // 2 * 2)
            if n // fac < w_l + 1 + 2 * max(r, 1) + 4:
                fits = False
        if fits and std0 is not None and std0 > 0:
            if float(np.std(block_downsample(f, fac).astype(np.float64))) < min_std_ratio * std0:
                fits = False
        if fits:
            levels = lv
        else:
            break
    return levels


def pyramid_search(f, g, coords, grid_shape, subset, radius, levels, clean, global_shift=None, outlier_threshold=2.0,
                   auto_expand=True, max_expand=3, fine_radius=PYRAMID_FINE_RADIUS, valid=None):
    """integer_search.py:617-694; ``levels <= 0`` selects the level count automatically
    (integer_search.py:648-650); ``clean(disp (N, 3), grid_shape, threshold)`` is the between-level
    median test + spring inpainting."""
    if levels <= 0:
        levels = auto_pyramid_levels(np.shape(f), subset, radius, f=f)
    coords = np.asarray(coords, dtype=np.int64)
    N = coords.shape[0]
    active = np.arange(N) if valid is None else np.flatnonzero(np.asarray(valid, dtype=bool))
    if active.size == 0:
        active = np.arange(N)
    disp_full = np.zeros((N, 3))
    if global_shift is not None:
        disp_full += np.asarray(global_shift, dtype=np.float64)[None, :]
    out = {"per_level": []}
    for lv in range(levels, -1, -1):
        fac = 2**lv
        fl = block_downsample(f, fac)
        gl = block_downsample(g, fac)
        subset_l = tuple(max(MIN_PYRAMID_SUBSET, (w // fac) // 2 * 2) for w in subset)
        rad_l = coarse_radius(radius, fac) if lv == levels else (int(fine_radius),) * 3
        coords_l = np.round(coords[active] / fac).astype(np.int64)
        shift_l = np.round(disp_full[active] / fac).astype(np.int64)
        sub = ncc_search_expanding(fl, gl, coords_l, subset_l, rad_l, shift_l, auto_expand, max_expand)
        res = {k: (np.full((N, 3), np.nan) if k == "disp" else (np.zeros(N, bool) if k in ("ok", "clipped") else np.full(N, np.nan)))
               for k in ("disp", "cc", "pce", "clipped", "ok")}  # fmt: skip
        for k in res:
            res[k][active] = sub[k]
        d = res["disp"] * fac
        d[~res["ok"]] = np.nan
        n_ok = int(res["ok"].sum())
        out["per_level"].append({"level": lv, "n_ok": n_ok, "radius": sub["radius"], "subset": subset_l})
        if n_ok == 0:
            continue
        disp_full = clean(d, grid_shape, outlier_threshold)
        out.update(ok=res["ok"], pce=res["pce"], cc=res["cc"], clipped=res["clipped"])
    out["disp"] = disp_full
    return out
