"""ORACLE -- test infrastructure only, never product code.

ctypes front-end of ``cpu_ref.c``, the float64 C restatement of the reference's CPU kernels.
The functions keep the reference's signatures and return conventions so that parity tests
read like the reference's own tests:

* ``sample_points``            solver/interp_kernels.py:139-152
* ``compute_gradients``        io/volume_ops.py:124-197 (stencil7, zero 3-voxel border)
* ``precompute_nodes``         solver/numba_kernels.py:499-509
* ``icgn_12dof_parallel``      solver/numba_kernels.py:1211-1278
* ``icgn_3dof_parallel``       solver/numba_kernels.py:1281-1346
* ``evaluate_zncc_parallel``   solver/numba_kernels.py:1354-1387
* ``zncc_direct``              solver/integer_search.py:44-151 (fills ``out`` in place)
* ``build_split_rows``         solver/numba_kernels.py:283-300
* ``compose_warp_inplace``     solver/numba_kernels.py:145-180
* ``cholesky12``               solver/numba_kernels.py:58-78

Only tests/, the smoke check and the CPU-baseline leg of the benchmark import this module.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libaldvc_oracle.so")
_lib = None

GRADIENT_BORDER = 3  # io/volume_ops.py:26


def build(force: bool = False) -> str:
    """Compile the oracle library in place with its Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    """The loaded oracle library (built on first use when missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name in (
            "orc_sample_points",
            "orc_gradients",
            "orc_precompute",
            "orc_icgn12",
            "orc_icgn3",
            "orc_evaluate_zncc",
            "orc_zncc_direct",
            "orc_split_rows",
            "orc_set_num_threads",
        ):
            getattr(L, name).restype = None
        L.orc_cholesky12.restype = ctypes.c_int
        L.orc_compose_warp.restype = ctypes.c_int
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def num_threads() -> int:
    return int(lib().orc_num_threads())


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(ctypes.c_int(int(n)))


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _i(v):
    return ctypes.c_int64(int(v))


def _c(v):
    return ctypes.c_int(int(v))


def _d(v):
    return ctypes.c_double(float(v))


def _vol(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype=dtype)


def _opt_grad(f, gx, gy, gz):
    """``(gx, gy, gz)`` as float32 arrays, or ``None`` for the 1x1x1 on-the-fly placeholders."""
    if gx is None or np.shape(gx) != np.shape(f):
        return None, None, None
    return _vol(gx), _vol(gy), _vol(gz)


def _opt_mask(f, mask):
    """The mask as uint8, or ``None`` for the 1x1x1 "no mask" placeholder."""
    if mask is None or np.shape(mask) != np.shape(f):
        return None
    return np.ascontiguousarray(mask, dtype=np.uint8)


def _split(n, split_index, split_keep):
    if split_index is None or split_keep is None:
        return None, None, 0
    si = np.ascontiguousarray(split_index, dtype=np.int64)
    sk = np.ascontiguousarray(split_keep, dtype=np.uint8)
    if si.shape != (n,):
        raise ValueError("split_index must have one entry per node")
    return si, sk, sk.shape[1] if sk.ndim == 2 else 0


def subset_count(hx, hy, hz, stride=1) -> int:
    """Sampled voxels of a subset (numba_kernels.py:52-55)."""
    return ((2 * hx) // stride + 1) * ((2 * hy) // stride + 1) * ((2 * hz) // stride + 1)


# --------------------------------------------------------------------------- interpolation


def sample_points(vol, zs, ys, xs, mode):
    vol = _vol(vol)
    zs = np.ascontiguousarray(zs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    out = np.empty(zs.shape[0])
    nz, ny, nx = vol.shape
    lib().orc_sample_points(_p(vol), _i(nz), _i(ny), _i(nx), _p(zs), _p(ys), _p(xs), _i(zs.shape[0]), _c(mode), _p(out))
    return out


def compute_gradients(f, border: int = GRADIENT_BORDER):
    f = _vol(f)
    gx = np.empty_like(f)
    gy = np.empty_like(f)
    gz = np.empty_like(f)
    nz, ny, nx = f.shape
    lib().orc_gradients(_p(f), _i(nz), _i(ny), _i(nx), _c(border), _p(gx), _p(gy), _p(gz))
    return gx, gy, gz


# --------------------------------------------------------------------------- small algebra


def cholesky12(H):
    H = np.ascontiguousarray(H, dtype=np.float64)
    L = np.zeros((12, 12))
    ok = lib().orc_cholesky12(_p(H), _p(L))
    return bool(ok), L


def compose_warp_inplace(P, dP) -> bool:
    """``W(P) <- W(P) W(dP)^-1`` on a contiguous float64 12-vector ``P`` (modified in place)."""
    if not (isinstance(P, np.ndarray) and P.dtype == np.float64 and P.flags.c_contiguous):
        raise TypeError("P must be a contiguous float64 array")
    dP = np.ascontiguousarray(dP, dtype=np.float64)
    return bool(lib().orc_compose_warp(_p(P), _p(dP)))


# --------------------------------------------------------------------------- precompute


def precompute_nodes(coords, hx, hy, hz, f, gx, gy, gz, mask, min_valid_ratio, cond_max, stride=1, split_index=None, split_keep=None):
    coords = np.ascontiguousarray(coords, dtype=np.int64)
    N = coords.shape[0]
    f = _vol(f)
    gx, gy, gz = _opt_grad(f, gx, gy, gz)
    m = _opt_mask(f, mask)
    si, sk, kb = _split(N, split_index, split_keep)
    H_all = np.zeros((N, 12, 12))
    L_all = np.zeros((N, 12, 12))
    meanf = np.zeros(N)
    bottomf = np.ones(N)
    nvalid = np.zeros(N, dtype=np.int64)
    valid = np.zeros(N, dtype=np.uint8)
    nz, ny, nx = f.shape
    lib().orc_precompute(
        _p(coords), _i(N), _c(hx), _c(hy), _c(hz), _c(stride),
        _p(f), _p(gx), _p(gy), _p(gz), _p(m), _i(nz), _i(ny), _i(nx),
        _d(min_valid_ratio), _d(cond_max), _p(si), _p(sk), _i(kb),
        _p(H_all), _p(L_all), _p(meanf), _p(bottomf), _p(nvalid), _p(valid),
    )  # fmt: skip
    return H_all, L_all, meanf, bottomf, nvalid, valid.astype(bool)


# --------------------------------------------------------------------------- 12-DOF IC-GN


def icgn_12dof_parallel(
    coords, P0, hx, hy, hz, f, gx, gy, gz, mask, g, mode, L_all, meanf_all, bottomf_all, valid,
    tol, dp_tol, max_iter, patience, stride=1, H_all=None, pattern=None, noise_gain=0.0,
    predictive=True, split_index=None, split_keep=None,
):  # fmt: skip
    """Returns ``(P_out (N,12), n_iter (N,) int32, status (N,) int8, zncc (N,))``."""
    coords = np.ascontiguousarray(coords, dtype=np.int64)
    N = coords.shape[0]
    P = np.array(P0, dtype=np.float64, copy=True, order="C").reshape(N, 12)
    f = _vol(f)
    g = _vol(g)
    if g.shape != f.shape:
        raise ValueError("f and g must have the same shape")
    gx, gy, gz = _opt_grad(f, gx, gy, gz)
    m = _opt_mask(f, mask)
    L_all = np.ascontiguousarray(L_all, dtype=np.float64)
    if H_all is None or pattern is None or noise_gain <= 0.0:
        H, pat, gain = None, None, 0.0
    else:
        H = np.ascontiguousarray(H_all, dtype=np.float64)
        pat = np.ascontiguousarray(pattern, dtype=np.float64)
        gain = float(noise_gain)
    vd = np.ascontiguousarray(valid, dtype=np.uint8)
    si, sk, kb = _split(N, split_index, split_keep)
    n_iter = np.zeros(N, dtype=np.int32)
    status = np.zeros(N, dtype=np.int8)
    zncc = np.zeros(N)
    nz, ny, nx = f.shape
    lib().orc_icgn12(
        _p(coords), _i(N), _p(P), _c(hx), _c(hy), _c(hz), _c(stride),
        _p(f), _p(gx), _p(gy), _p(gz), _p(m), _p(g), _i(nz), _i(ny), _i(nx), _c(mode),
        _p(L_all), _p(H), _p(pat), _d(gain), _p(vd), _d(tol), _d(dp_tol), _c(max_iter), _c(patience),
        _c(bool(predictive)), _p(si), _p(sk), _i(kb), _p(n_iter), _p(status), _p(zncc),
    )  # fmt: skip
    return P, n_iter, status, zncc


# --------------------------------------------------------------------------- 3-DOF IC-GN


def icgn_3dof_parallel(
    coords, U_old, F_fixed, vdual, hx, hy, hz, f, gx, gy, gz, mask, g, mode, H_all, meanf_all,
    bottomf_all, valid, mu, tol, dp_tol, max_iter, patience, stride=1, n_full=0.0, noise_gain=0.0,
    predictive=True, split_index=None, split_keep=None,
):  # fmt: skip
    """Returns ``(U_out (N,3), n_iter, status, zncc)``."""
    coords = np.ascontiguousarray(coords, dtype=np.int64)
    N = coords.shape[0]
    U_old = np.ascontiguousarray(U_old, dtype=np.float64).reshape(N, 3)
    U = U_old.copy()
    Ff = np.ascontiguousarray(F_fixed, dtype=np.float64).reshape(N, 9)
    vd3 = np.ascontiguousarray(vdual, dtype=np.float64).reshape(N, 3)
    f = _vol(f)
    g = _vol(g)
    if g.shape != f.shape:
        raise ValueError("f and g must have the same shape")
    gx, gy, gz = _opt_grad(f, gx, gy, gz)
    m = _opt_mask(f, mask)
    H = np.ascontiguousarray(H_all, dtype=np.float64)
    bf = np.ascontiguousarray(bottomf_all, dtype=np.float64)
    vd = np.ascontiguousarray(valid, dtype=np.uint8)
    si, sk, kb = _split(N, split_index, split_keep)
    n_iter = np.zeros(N, dtype=np.int32)
    status = np.zeros(N, dtype=np.int8)
    zncc = np.zeros(N)
    nz, ny, nx = f.shape
    lib().orc_icgn3(
        _p(coords), _i(N), _p(U_old), _p(U), _p(Ff), _p(vd3), _c(hx), _c(hy), _c(hz), _c(stride),
        _p(f), _p(gx), _p(gy), _p(gz), _p(m), _p(g), _i(nz), _i(ny), _i(nx), _c(mode),
        _p(H), _p(bf), _p(vd), _d(mu), _d(tol), _d(dp_tol), _c(max_iter), _c(patience),
        _d(noise_gain if noise_gain > 0.0 else 0.0), _c(bool(predictive)), _p(si), _p(sk), _i(kb),
        _p(n_iter), _p(status), _p(zncc),
    )  # fmt: skip
    return U, n_iter, status, zncc


def evaluate_zncc_parallel(coords, P_all, hx, hy, hz, f, mask, g, mode, meanf_all, bottomf_all, valid, stride=1):
    coords = np.ascontiguousarray(coords, dtype=np.int64)
    N = coords.shape[0]
    P = np.ascontiguousarray(P_all, dtype=np.float64).reshape(N, 12)
    f = _vol(f)
    g = _vol(g)
    m = _opt_mask(f, mask)
    vd = np.ascontiguousarray(valid, dtype=np.uint8)
    out = np.empty(N)
    nz, ny, nx = f.shape
    lib().orc_evaluate_zncc(
        _p(coords), _i(N), _p(P), _c(hx), _c(hy), _c(hz), _c(stride), _p(f), _p(m), _p(g),
        _i(nz), _i(ny), _i(nx), _c(mode), _p(vd), _p(out),
    )  # fmt: skip
    return out


# --------------------------------------------------------------------------- NCC map


def zncc_direct(f, g, coords, hx, hy, hz, rx, ry, rz, wx0, wy0, wz0, t_ok, out) -> None:
    """Fills ``out`` (n, 2rz+1, 2ry+1, 2rx+1) float64 in place."""
    f = _vol(f)
    g = _vol(g)
    coords = np.ascontiguousarray(coords, dtype=np.int64)
    n = coords.shape[0]
    wx0 = np.ascontiguousarray(wx0, dtype=np.int64)
    wy0 = np.ascontiguousarray(wy0, dtype=np.int64)
    wz0 = np.ascontiguousarray(wz0, dtype=np.int64)
    tk = np.ascontiguousarray(t_ok, dtype=np.uint8)
    buf = out if (out.dtype == np.float64 and out.flags.c_contiguous) else np.empty(out.shape)
    nz, ny, nx = f.shape
    lib().orc_zncc_direct(
        _p(f), _p(g), _i(nz), _i(ny), _i(nx), _p(coords), _i(n), _c(hx), _c(hy), _c(hz),
        _c(rx), _c(ry), _c(rz), _p(wx0), _p(wy0), _p(wz0), _p(tk), _p(buf),
    )  # fmt: skip
    if buf is not out:
        out[...] = buf


# --------------------------------------------------------------------------- split rows


def build_split_rows(coords, cand, hx, hy, hz, stride, mask, margin=0):
    """``(rows (M, ceil(S/8)) uint8, n_keep (M,), n_inmask (M,))``."""
    coords = np.ascontiguousarray(coords, dtype=np.int64)
    cand = np.ascontiguousarray(cand, dtype=np.int64)
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    M = cand.shape[0]
    S = subset_count(hx, hy, hz, stride)
    kb = (S + 7) // 8
    rows = np.zeros((M, kb), dtype=np.uint8)
    n_keep = np.zeros(M, dtype=np.int64)
    n_inmask = np.zeros(M, dtype=np.int64)
    nz, ny, nx = mask.shape
    lib().orc_split_rows(
        _p(coords), _p(cand), _i(M), _c(hx), _c(hy), _c(hz), _c(stride), _p(mask),
        _i(nz), _i(ny), _i(nx), _c(margin), _i(kb), _p(rows), _p(n_keep), _p(n_inmask),
    )  # fmt: skip
    return rows, n_keep, n_inmask
