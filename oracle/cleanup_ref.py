"""ORACLE -- test infrastructure only, never product code.

Host NumPy/SciPy restatement of the reference's node-grid cleanup, the checker of the device median
test and spring fill (``aldvc_subpb1_cleanup``, ``aldvc_spring_fill``):

* ``universal_median_test`` / ``normalized_fluctuation``  utils/outlier_detection.py:24-170
  (Westerweel & Scarano 2005; the cut-mesh neighbourhoods of :60-132)
* ``fill_nan_grid`` / ``fill_nan_nodes``  utils/inpaint.py:35-126, the "spring" model of D'Errico's
  ``inpaint_nans`` (the discrete Laplace equation on the unknown nodes, known nodes as Dirichlet
  data; sparse direct solve up to 60,000 unknowns, CG beyond, nearest-neighbour past 300,000)
* ``fill_nan_nearest``  utils/inpaint.py:22-33

Pinned by the reference's own outputs (``tests/golden/golden_driver.npz``, written by
``tests/golden/make_golden_driver.py``) and by closed forms (``tests/test_oracle_cleanup.py``).
Grids are ``(nz, ny, nx)``; ``edge_ok`` is ``(nz, ny, nx, 3)`` with the +x, +y, +z edge of every
node (False where a masked boundary cut the mesh).
"""

from __future__ import annotations

import itertools
import warnings

import numpy as np
from scipy import sparse
from scipy.ndimage import distance_transform_edt
from scipy.sparse.linalg import cg, spsolve

SPRING_MAX_UNKNOWN = 300_000  # beyond this many unknowns the fill is nearest-neighbour (inpaint.py:20)
DIRECT_MAX_UNKNOWN = 60_000  # sparse direct solve up to here, conjugate gradients beyond (inpaint.py:117)
MIN_NEIGHBOURS = 6  # a node of a cut mesh with fewer reachable neighbours is not judged (outlier_detection.py:21)


# ---------------------------------------------------------------------------- shifts on a grid
def shifted(arr: np.ndarray, off, fill) -> np.ndarray:
    """``out[n] = arr[n + off]`` over the first three axes, ``fill`` beyond the grid."""
    out = np.full(arr.shape, fill, dtype=arr.dtype)
    dst, src = [], []
    for ax, o in enumerate(off):
        n = arr.shape[ax]
        dst.append(slice(max(0, -o), n - max(0, o)))
        src.append(slice(max(0, o), n - max(0, -o)))
    out[tuple(dst)] = arr[tuple(src)]
    return out


def _neighbour_stack(arr: np.ndarray, size: int) -> np.ndarray:
    """The ``size^3 - 1`` neighbours of every node (centre excluded), edge values repeated."""
    r = size // 2
    pad = np.pad(arr, r, mode="edge")
    nz, ny, nx = arr.shape
    views = [
        pad[r + dz : r + dz + nz, r + dy : r + dy + ny, r + dx : r + dx + nx]
        for dz, dy, dx in itertools.product(range(-r, r + 1), repeat=3)
        if (dz, dy, dx) != (0, 0, 0)
    ]
    return np.stack(views)


def _rank_median(arr: np.ndarray, size: int) -> np.ndarray:
    """Neighbourhood median as a rank filter: element ``(size^3 - 1) // 2`` of the sorted
    neighbours (the upper median for the even count), edges repeated -- the definition of
    ``scipy.ndimage.median_filter`` with a centre-less footprint and ``mode="nearest"``."""
    stack = _neighbour_stack(arr, size)
    k = stack.shape[0] // 2
    return np.partition(stack, k, axis=0)[k]


# ---------------------------------------------------------------------------- cut meshes
def reachable_offsets(edge_ok: np.ndarray) -> dict:
    """For each neighbour offset ``(dz, dy, dx)`` of the 3x3x3 window: where that neighbour is
    reached from the node through ok edges, taking the axis steps in some order."""
    E = np.asarray(edge_ok, dtype=bool)
    grid = E.shape[:3]
    step = {}
    for ax in range(3):  # array axis 0 = z, 1 = y, 2 = x; edge component 2 - ax
        plus = E[..., 2 - ax].copy()
        last = [slice(None)] * 3
        last[ax] = slice(grid[ax] - 1, grid[ax])
        plus[tuple(last)] = False  # no +edge out of the last node along the axis
        step[(ax, 1)] = plus
        minus_off = [0, 0, 0]
        minus_off[ax] = -1
        step[(ax, -1)] = shifted(plus, tuple(minus_off), False)  # the +edge of the node behind
    out = {}
    for off in itertools.product((-1, 0, 1), repeat=3):
        if off == (0, 0, 0):
            continue
        moves = [(ax, off[ax]) for ax in range(3) if off[ax]]
        reach = np.zeros(grid, dtype=bool)
        for order in itertools.permutations(moves):
            ok = np.ones(grid, dtype=bool)
            pos = [0, 0, 0]
            for ax, d in order:
                ok &= shifted(step[(ax, d)], tuple(pos), False)
                pos[ax] += d
            reach |= ok
        out[off] = reach
    return out


def _fluctuation_cut(arr: np.ndarray, edge_ok: np.ndarray, eps: float) -> np.ndarray:
    nan_mask = np.isnan(arr)
    reach = reachable_offsets(edge_ok)
    stack = np.stack([np.where(r, shifted(arr, off, np.nan), np.nan) for off, r in reach.items()])
    count = np.sum(np.isfinite(stack), axis=0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)  # neighbourhoods with no finite value
        med = np.nanmedian(stack, axis=0)
        absf = np.abs(arr - med)
        res = np.stack([np.where(r, shifted(absf, off, np.nan), np.nan) for off, r in reach.items()])
        med_res = np.nanmedian(res, axis=0)
    nf = absf / (med_res + eps)
    nf[nan_mask | ~np.isfinite(nf) | (count < MIN_NEIGHBOURS)] = 0.0
    return nf


# ---------------------------------------------------------------------------- median test
def normalized_fluctuation(field, valid=None, eps: float = 0.1, size: int = 3, edge_ok=None) -> np.ndarray:
    """``|u - med(u_nb)| / (med(|r_nb|) + eps)`` of a scalar ``(nz, ny, nx)`` field.

    Invalid or NaN nodes are inpainted first (so they do not bias the medians) and report 0; on a
    cut mesh (``edge_ok``) they are left out of the neighbourhoods instead, and only reachable
    neighbours count."""
    arr = np.array(field, dtype=np.float64, copy=True)
    if valid is not None:
        arr[~np.asarray(valid, dtype=bool)] = np.nan
    nan_mask = np.isnan(arr)
    if nan_mask.all():
        return np.zeros_like(arr)
    if edge_ok is not None:
        if size != 3:
            raise ValueError("the cut-mesh median test supports the 3 x 3 x 3 neighbourhood only")
        return _fluctuation_cut(arr, edge_ok, eps)
    if nan_mask.any():
        arr = fill_nan_grid(arr)
    fluct = arr - _rank_median(arr, size)
    absf = np.abs(fluct)
    nf = absf / (_rank_median(absf, size) + eps)
    nf[nan_mask] = 0.0
    return nf


def universal_median_test(field, valid, threshold: float, eps: float = 0.1, size: int = 3, edge_ok=None) -> np.ndarray:
    """Outlier flags ``(nz, ny, nx)`` of a vector ``(nz, ny, nx, C)`` (or scalar) grid field: the
    norm over components of the normalised fluctuation exceeds ``threshold``; only ``valid``
    nodes are flagged."""
    arr = np.asarray(field, dtype=np.float64)
    if arr.ndim == 3:
        arr = arr[..., None]
    if threshold <= 0:
        return np.zeros(arr.shape[:3], dtype=bool)
    mag2 = np.zeros(arr.shape[:3])
    for c in range(arr.shape[3]):
        nf = normalized_fluctuation(arr[..., c], valid, eps=eps, size=size, edge_ok=edge_ok)
        mag2 += nf * nf
    flag = np.sqrt(mag2) > threshold
    if valid is not None:
        flag &= np.asarray(valid, dtype=bool)
    return flag


# ---------------------------------------------------------------------------- inpainting
def fill_nan_nearest(grid) -> np.ndarray:
    """Every NaN node takes the value of the nearest finite node (Euclidean)."""
    arr = np.array(grid, dtype=np.float64, copy=True)
    nan = np.isnan(arr)
    if not nan.any():
        return arr
    if nan.all():
        warnings.warn("fill_nan_nearest: every node is NaN, returning zeros.", stacklevel=2)
        return np.zeros_like(arr)
    idx = distance_transform_edt(nan, return_distances=False, return_indices=True)
    return arr[tuple(idx)]


def fill_nan_grid(grid, method: str = "spring", edge_ok=None) -> np.ndarray:
    """Fill the NaN nodes of a grid by the spring model (harmonic extension over the 2·ndim
    neighbours; springs across cut edges removed), or nearest-neighbour."""
    arr = np.array(grid, dtype=np.float64, copy=True)
    nan = np.isnan(arr)
    n_unknown = int(nan.sum())
    if n_unknown == 0:
        return arr
    if nan.all():
        warnings.warn("fill_nan_grid: every node is NaN, returning zeros.", stacklevel=2)
        return np.zeros_like(arr)
    if method == "nearest" or n_unknown > SPRING_MAX_UNKNOWN:
        return fill_nan_nearest(arr)
    flat = arr.ravel()
    unknown = np.flatnonzero(nan.ravel())
    row_of = np.full(flat.size, -1, dtype=np.int64)
    row_of[unknown] = np.arange(n_unknown)
    ids = np.arange(flat.size).reshape(arr.shape)
    diag = np.zeros(n_unknown)
    rhs = np.zeros(n_unknown)
    rows, cols = [], []
    for ax in range(arr.ndim):
        if arr.shape[ax] < 2:
            continue
        lo = [slice(None)] * arr.ndim
        hi = [slice(None)] * arr.ndim
        lo[ax], hi[ax] = slice(0, -1), slice(1, None)
        a, b = ids[tuple(lo)].ravel(), ids[tuple(hi)].ravel()
        if edge_ok is not None and arr.ndim == 3:
            ok = np.asarray(edge_ok, dtype=bool)[..., 2 - ax][tuple(lo)].ravel()
            a, b = a[ok], b[ok]
        for p, q in ((a, b), (b, a)):  # the spring p-q in the equation of p, when p is unknown
            sel = row_of[p] >= 0
            rp, q = row_of[p[sel]], q[sel]
            np.add.at(diag, rp, 1.0)
            rq = row_of[q]
            known = rq < 0
            np.add.at(rhs, rp[known], flat[q[known]])
            rows.append(rp[~known])
            cols.append(rq[~known])
    r = np.concatenate(rows + [np.arange(n_unknown)])
    c = np.concatenate(cols + [np.arange(n_unknown)])
    v = np.concatenate([-np.ones(sum(x.size for x in rows)), diag + 1e-10])  # Tikhonov: isolated islands
    A = sparse.csr_matrix((v, (r, c)), shape=(n_unknown, n_unknown))
    try:
        if n_unknown <= DIRECT_MAX_UNKNOWN:
            x = spsolve(A.tocsc(), rhs)
        else:
            x, info = cg(A, rhs, rtol=1e-10, maxiter=2000)
            if info != 0:
                x = spsolve(A.tocsc(), rhs)
    except Exception:  # numerical safety net, as the reference's
        return fill_nan_nearest(arr)
    if not np.all(np.isfinite(x)):
        return fill_nan_nearest(arr)
    flat[unknown] = x
    return flat.reshape(arr.shape)


def fill_nan_nodes(values, grid_shape, method: str = "spring", edge_ok=None) -> np.ndarray:
    """Fill the NaNs of a per-node array ``(N,)`` or ``(N, C)`` laid out on ``grid_shape``."""
    v = np.asarray(values, dtype=np.float64)
    if v.ndim == 1:
        return fill_nan_grid(v.reshape(grid_shape), method, edge_ok).ravel()
    out = v.copy()
    for c in range(v.shape[1]):
        out[:, c] = fill_nan_grid(v[:, c].reshape(grid_shape), method, edge_ok).ravel()
    return out
