"""Reference-signature entry points of the local DVC step, backed by the HIP library.

Each function keeps the argument list, argument meaning and return convention of the reference
backend function it replaces (al_dvc 1.2.1):

* :func:`precompute_nodes`        solver/numba_kernels.py:499-509 (``precompute_nodes_cuda``, cuda_kernels.py:1635-1710)
* :func:`icgn_12dof_parallel`     solver/numba_kernels.py:1211-1278 (``icgn_12dof_cuda``, cuda_kernels.py:1397-1515)
* :func:`icgn_3dof_parallel`      solver/numba_kernels.py:1281-1346 (``icgn_3dof_cuda``, cuda_kernels.py:1518-1632)
* :func:`evaluate_zncc_parallel`  solver/numba_kernels.py:1354-1387
* :func:`zncc_direct`             solver/integer_search.py:44-151 (``zncc_direct_cuda``, cuda_kernels.py:1713-1755)
* :func:`sample_points`           solver/interp_kernels.py:139-152

Arrays may be host numpy arrays (the reference's convention: fresh host arrays come back) or
torch tensors on the GPU (device-resident: results stay on the device, nothing is copied to the
host). The 1x1x1 placeholders of the reference mean "no mask" (``mask``) and "gradients on the
fly" (``gx, gy, gz``). Per-node failures are status codes, never exceptions; a failed library
call raises :class:`~pyaldvc_amd._lib.HipLibraryError`.
"""

from __future__ import annotations

import ctypes
import threading
from collections import OrderedDict

import numpy as np
import torch

from ._lib import check, lib

INTERP_LINEAR = 0
INTERP_CUBIC = 1
INTERP_BSPLINE = 2
INTERP_MODE_BY_NAME = {"linear": INTERP_LINEAR, "cubic": INTERP_CUBIC, "bspline": INTERP_BSPLINE}

STATUS_CONVERGED = 0
STATUS_MAX_ITER = 1
STATUS_OUT_OF_BOUNDS = 2
STATUS_INVALID_SUBSET = 3
STATUS_SINGULAR = 4
STATUS_NAN = 5
STATUS_SKIPPED = 6
STATUS_STALLED = 7

VARIANTS = {"plain": 0, "fused": 1, "fused_gather": 2}
DEVICE_CACHE_SIZE = 8  # uploaded host volumes kept per process (cuda_kernels.py:70)


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("the HIP backend needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(stream: torch.cuda.Stream | None = None) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class _VolumeCache:
    """Device copies of host volumes, keyed by the host array (cuda_kernels.py:184-214)."""

    def __init__(self, size: int = DEVICE_CACHE_SIZE):
        self.size = size
        self._d: OrderedDict = OrderedDict()
        self._lock = threading.Lock()

    def get(self, arr: np.ndarray, dtype: torch.dtype) -> torch.Tensor:
        key = (id(arr), arr.__array_interface__["data"][0], arr.shape, arr.dtype.str, dtype)
        with self._lock:
            hit = self._d.get(key)
            if hit is not None and hit[0] is arr:
                self._d.move_to_end(key)
                return hit[1]
        t = torch.from_numpy(np.ascontiguousarray(arr)).to(device=device(), dtype=dtype)
        with self._lock:
            self._d[key] = (arr, t)
            while len(self._d) > self.size:
                self._d.popitem(last=False)
        return t

    def clear(self) -> None:
        with self._lock:
            self._d.clear()


_cache = _VolumeCache()


def clear_device_cache() -> None:
    _cache.clear()


_NP_TO_TORCH = {
    np.float32: torch.float32,
    np.float64: torch.float64,
    np.int64: torch.int64,
    np.int32: torch.int32,
    np.int8: torch.int8,
    np.uint8: torch.uint8,
}


class _Args:
    """Moves arguments to the device and remembers whether results go back to the host."""

    def __init__(self, anchor):
        self.host = not isinstance(anchor, torch.Tensor)

    @staticmethod
    def vol(x, dtype=np.float32) -> torch.Tensor:
        """A volume-sized array (cached upload for host arrays)."""
        td = _NP_TO_TORCH[dtype]
        if isinstance(x, torch.Tensor):
            return x.to(device=device(), dtype=td).contiguous()
        a = np.asarray(x)
        if a.dtype != dtype:
            a = a.astype(dtype)
        return _cache.get(a, td)

    @staticmethod
    def arr(x, dtype) -> torch.Tensor:
        td = _NP_TO_TORCH[dtype]
        if isinstance(x, torch.Tensor):
            return x.to(device=device(), dtype=td).contiguous()
        return torch.from_numpy(np.ascontiguousarray(x, dtype=dtype)).to(device())

    def out(self, t: torch.Tensor):
        return t.cpu().numpy() if self.host else t


def _shape(x) -> tuple:
    return tuple(x.shape)


def _ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _grads(A: _Args, f, gx, gy, gz):
    """Stored gradient volumes, or ``(None, None, None)`` for the on-the-fly placeholders."""
    if gx is None or _shape(gx) != _shape(f):
        return None, None, None
    return A.vol(gx), A.vol(gy), A.vol(gz)


def _mask(A: _Args, f, mask):
    if mask is None or _shape(mask) != _shape(f):
        return None
    return A.vol(mask, np.uint8)


def _split(A: _Args, n: int, split_index, split_keep):
    if split_index is None or split_keep is None:
        return None, None, 0
    si = A.arr(split_index, np.int64)
    sk = A.arr(split_keep, np.uint8)
    if tuple(si.shape) != (n,):
        raise ValueError("split_index must have one entry per node")
    if sk.dim() != 2:
        raise ValueError("split_keep must be a 2-D array of packed rows")
    if n and int(si.max()) >= int(sk.shape[0]):
        raise ValueError("split_index refers to a row beyond split_keep")
    return si, sk, int(sk.shape[1])


def _dims(f) -> tuple[int, int, int]:
    s = _shape(f)
    if len(s) != 3:
        raise ValueError(f"expected a 3-D volume, got shape {s}")
    return int(s[0]), int(s[1]), int(s[2])


def subset_count(hx, hy, hz, stride=1) -> int:
    """Sampled voxels of a subset (numba_kernels.py:52-55)."""
    return ((2 * hx) // stride + 1) * ((2 * hy) // stride + 1) * ((2 * hz) // stride + 1)


# --------------------------------------------------------------------------- gradients / sampling


def compute_gradients(f, border: int = 3):
    """7-point central differences ``(gx, gy, gz)``, zero within ``border`` voxels of a face
    (io/volume_ops.py:124-197)."""
    A = _Args(f)
    fd = A.vol(f)
    nz, ny, nx = _dims(fd)
    gx, gy, gz = (torch.empty_like(fd) for _ in range(3))
    check(
        lib().aldvc_gradients(_ptr(fd), nz, ny, nx, int(border), _ptr(gx), _ptr(gy), _ptr(gz), stream_handle()),
        "aldvc_gradients",
    )
    return A.out(gx), A.out(gy), A.out(gz)


def sample_points(vol, zs, ys, xs, mode):
    """Interpolate ``vol`` at the points; NaN outside ``1 <= c <= n - 3`` (float32 arithmetic)."""
    A = _Args(vol)
    v = A.vol(vol)
    nz, ny, nx = _dims(v)
    z = A.arr(zs, np.float64)
    y = A.arr(ys, np.float64)
    x = A.arr(xs, np.float64)
    out = torch.empty(z.shape[0], dtype=torch.float32, device=v.device)
    check(
        lib().aldvc_sample_points(_ptr(v), nz, ny, nx, _ptr(z), _ptr(y), _ptr(x), int(z.shape[0]), int(mode), _ptr(out), stream_handle()),
        "aldvc_sample_points",
    )
    return A.out(out.double())


# --------------------------------------------------------------------------- precompute


def precompute_nodes(coords, hx, hy, hz, f, gx, gy, gz, mask, min_valid_ratio, cond_max, stride=1, split_index=None,
                     split_keep=None, separable=None):  # fmt: skip
    """``(H_all (N,12,12), L_all (N,12,12), meanf (N,), bottomf (N,), n_valid (N,), valid (N,))``.
    ``separable``: the separable pass for the nodes without keep rows (None: by
    ``separable_wanted``)."""
    A = _Args(f)
    c = A.arr(coords, np.int64)
    N = int(c.shape[0])
    fd = A.vol(f)
    nz, ny, nx = _dims(fd)
    gxd, gyd, gzd = _grads(A, f, gx, gy, gz)
    md = _mask(A, f, mask)
    si, sk, kb = _split(A, N, split_index, split_keep)
    dev = fd.device
    ext = _separable_extent(c, None, hx, hy, hz, stride, (nz, ny, nx), gxd is not None, separable)
    ws_bytes = 0 if ext is None else int(lib().aldvc_precompute_row_bytes(*ext, ny, nx))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    H = torch.zeros((N, 12, 12), dtype=torch.float64, device=dev)
    L = torch.zeros((N, 12, 12), dtype=torch.float64, device=dev)
    meanf = torch.zeros(N, dtype=torch.float64, device=dev)
    bottomf = torch.ones(N, dtype=torch.float64, device=dev)
    nvalid = torch.zeros(N, dtype=torch.int64, device=dev)
    valid = torch.zeros(N, dtype=torch.uint8, device=dev)
    check(
        lib().aldvc_precompute_tuned(
            _ptr(c), N, int(hx), int(hy), int(hz), int(stride), _ptr(fd), _ptr(gxd), _ptr(gyd), _ptr(gzd), _ptr(md),
            nz, ny, nx, float(min_valid_ratio), float(cond_max), _ptr(si), _ptr(sk), kb,
            _ptr(H), _ptr(L), _ptr(meanf), _ptr(bottomf), _ptr(nvalid), _ptr(valid),
            TUNE_ROW_MOMENTS if ws_bytes else 0, _ptr(ws), ws_bytes, stream_handle(),
        ),
        "aldvc_precompute_tuned",
    )  # fmt: skip
    return A.out(H), A.out(L), A.out(meanf), A.out(bottomf), A.out(nvalid), A.out(valid.bool())


# --------------------------------------------------------------------------- subset splitting


def subset_valid_fraction(mask, coordinates, winsize):
    """In-mask fraction ``(N,)`` of each node's window ``[c - w//2, c + w//2]`` per axis, clipped
    to the volume, over the nominal window volume (mesh/grid_mesh.py:135-188). ``coordinates``
    are rounded to voxels; ``winsize`` is ``(wx, wy, wz)``."""
    A = _Args(mask)
    md = A.vol(mask, np.uint8)
    nz, ny, nx = _dims(md)
    if isinstance(coordinates, torch.Tensor):
        c = torch.round(coordinates.to(device(), torch.float64)).to(torch.int64).contiguous()
    else:
        c = A.arr(np.round(np.asarray(coordinates, dtype=np.float64)).astype(np.int64), np.int64)
    N = int(c.shape[0])
    hx, hy, hz = (int(w) // 2 for w in winsize)
    frac = torch.zeros(N, dtype=torch.float64, device=md.device)
    ws = torch.empty(int(lib().aldvc_mask_table_bytes(nz, ny, nx)), dtype=torch.uint8, device=md.device)
    check(
        lib().aldvc_subset_valid_fraction(_ptr(md), nz, ny, nx, _ptr(c), N, hx, hy, hz, _ptr(frac), _ptr(ws),
                                          int(ws.numel()), stream_handle()),
        "aldvc_subset_valid_fraction",
    )  # fmt: skip
    return A.out(frac)


def build_split_rows(coords, cand, hx, hy, hz, stride, mask, margin=0):
    """Packed keep rows ``(rows (M, ceil(S/8)) uint8, n_keep (M,), n_inmask (M,))`` of the
    candidate nodes ``cand`` (numba_kernels.py:271-289): bit i of a row is set when sampled voxel
    i (kernel order) belongs to the 6-connected in-mask component holding the subset centre."""
    A = _Args(mask)
    md = A.vol(mask, np.uint8)
    nz, ny, nx = _dims(md)
    c = A.arr(coords, np.int64)
    cd = A.arr(cand, np.int64)
    M = int(cd.shape[0])
    if M and (int(cd.min()) < 0 or int(cd.max()) >= int(c.shape[0])):
        raise ValueError("cand holds an index outside coords")
    kb = (subset_count(hx, hy, hz, stride) + 7) // 8
    rows = torch.zeros((M, kb), dtype=torch.uint8, device=md.device)
    n_keep = torch.zeros(M, dtype=torch.int64, device=md.device)
    n_inmask = torch.zeros(M, dtype=torch.int64, device=md.device)
    if M:
        check(
            lib().aldvc_split_rows(_ptr(c), _ptr(cd), M, int(hx), int(hy), int(hz), int(stride), _ptr(md), nz, ny, nx,
                                   int(margin), kb, _ptr(rows), _ptr(n_keep), _ptr(n_inmask), stream_handle()),
            "aldvc_split_rows",
        )  # fmt: skip
    return A.out(rows), A.out(n_keep), A.out(n_inmask)


# --------------------------------------------------------------------------- 12-DOF IC-GN

TUNE_ROW_MOMENTS = 32  # aldvc_icgn12_tuned, aldvc_icgn3_tuned, aldvc_precompute_tuned: the separable passes
ROW_MOMENTS_MIN_READS = 500_000_000  # subset-voxel reads of the per-node pass (~0.8 ms of it) below which it stays


def separable_wanted(n_nodes, half, stride, dims, stored) -> bool:
    """Whether the separable passes (reference moments, precompute) pay for ``n_nodes`` nodes of
    half-widths ``half`` in a volume of shape ``dims`` (z, y, x): stored gradients, stride 1, and at
    least ``ROW_MOMENTS_MIN_READS`` subset-voxel reads of the per-node pass and four times the voxel
    count (the separable pass's cost follows the volume, the per-node pass's the node count; below
    that, its extra launches and the count read are not repaid). A caller that cuts a volume into
    boxes decides once for the whole and passes the answer to every box (``separable=``), so that
    boxes and whole take the same pass."""
    if not stored or int(stride) != 1 or n_nodes <= 0:
        return False
    nz, ny, nx = dims
    hx, hy, hz = (int(h) for h in half)
    reads = int(n_nodes) * (2 * hx + 1) * (2 * hy + 1) * (2 * hz + 1)
    return reads >= ROW_MOMENTS_MIN_READS and reads >= 4 * nz * ny * nx


def _separable_extent(c, vd, hx, hy, hz, stride, dims, stored, separable=None):
    """``(ncx, ncy, nplanes)`` of the separable passes over the nodes ``c`` (device ``(N, 3)``
    int64, x y z) whose subsets lie inside the volume (and are valid, when ``vd`` is given), or None
    where the per-node pass runs: on-the-fly gradients, stride > 1, ``separable`` False, or (None)
    ``separable_wanted`` says no. One device-to-host read of four counts."""
    if not stored or int(stride) != 1 or c.shape[0] == 0 or separable is False:
        return None
    nz, ny, nx = dims
    if separable is None and not separable_wanted(int(c.shape[0]), (hx, hy, hz), stride, dims, stored):
        return None
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    ok = (x >= hx) & (x < nx - hx) & (y >= hy) & (y < ny - hy) & (z >= hz) & (z < nz - hz)
    if vd is not None:
        ok &= vd != 0
    mx = torch.zeros(nx + 1, dtype=torch.int64, device=c.device)
    my = torch.zeros(ny + 1, dtype=torch.int64, device=c.device)
    mx[torch.where(ok, x, nx)] = 1
    my[torch.where(ok, y, ny)] = 1
    zlo = torch.where(ok, z, nz).min()
    zhi = torch.where(ok, z, -1).max()
    ncx, ncy, zlo, zhi = (int(v) for v in torch.stack([mx[:nx].sum(), my[:ny].sum(), zlo, zhi]).tolist())
    if zhi < zlo:
        return None
    return ncx, ncy, zhi - zlo + 2 * int(hz) + 1


def row_moments_bytes(c, vd, hx, hy, hz, stride, dims, stored, separable=None) -> int:
    """Extra workspace of the separable reference-moment pass (flag ``TUNE_ROW_MOMENTS``) for the
    nodes ``c``, 0 when they take the per-node pass (``_separable_extent``). The pass weights the
    sums by the reference mask; nodes with keep rows take the per-node pass."""
    ext = _separable_extent(c, vd, hx, hy, hz, stride, dims, stored, separable)
    return 0 if ext is None else int(lib().aldvc_icgn12_row_moments_bytes(*ext, dims[1], dims[2]))



def icgn_12dof_parallel(
    coords, P0, hx, hy, hz, f, gx, gy, gz, mask, g, mode, L_all, meanf_all, bottomf_all, valid,
    tol, dp_tol, max_iter, patience, stride=1, H_all=None, pattern=None, noise_gain=0.0,
    predictive=True, split_index=None, split_keep=None, variant: str = "fused", stream=None, tuning=None,
    separable=None,
):  # fmt: skip
    """12-DOF IC-GN over all nodes: ``(P_out (N,12), n_iter (N,) int32, status (N,) int8, zncc (N,))``.

    ``variant="plain"`` runs the reference GPU kernel's design (two passes through a scratch row,
    2048-node launches); ``"fused"`` (default) runs the one-pass kernel, which stages the deformed
    volume in LDS bricks where the warped subset fits one; ``"fused_gather"`` is the one-pass
    kernel with memory gathers only; ``"auto"`` lets the autotuner pick the fused launch shape for
    this subset shape on this device (timed once on a node sample, then cached). ``tuning`` (an
    ``autotune.Tuning``) sets the launch shape explicitly. ``separable``: the reference moments of
    the nodes without keep rows by the separable pass (None: by ``separable_wanted``).
    ``meanf_all`` feeds the fused kernel's accumulation pilot only; ``bottomf_all`` is accepted
    for signature parity (the 12-DOF solver renormalises on the dynamic voxel set,
    numba_kernels.py:680-708).
    """
    if variant == "auto" and tuning is None:
        from . import autotune

        args = (coords, P0, hx, hy, hz, f, gx, gy, gz, mask, g, mode, L_all, meanf_all, bottomf_all, valid, tol, dp_tol,
                max_iter, patience)  # fmt: skip
        kw = dict(stride=stride, H_all=H_all, pattern=pattern, noise_gain=noise_gain, predictive=predictive,
                  split_index=split_index, split_keep=split_keep, stream=stream)  # fmt: skip
        tuning = autotune.choose_icgn12(args, kw)
    if tuning is not None:
        variant = {v: k for k, v in VARIANTS.items()}[tuning.variant]
    if variant not in VARIANTS:
        raise ValueError(f"variant must be one of {sorted(VARIANTS) + ['auto']}")
    lds_budget, flags = (int(tuning.lds_budget), int(tuning.flags)) if tuning is not None else (0, 0)
    A = _Args(f)
    c = A.arr(coords, np.int64)
    N = int(c.shape[0])
    P = A.arr(P0, np.float64).reshape(N, 12).clone()
    fd = A.vol(f)
    gd = A.vol(g)
    if _shape(gd) != _shape(fd):
        raise ValueError("f and g must have the same shape")
    nz, ny, nx = _dims(fd)
    gxd, gyd, gzd = _grads(A, f, gx, gy, gz)
    md = _mask(A, f, mask)
    Ld = A.arr(L_all, np.float64)
    use_noise = H_all is not None and pattern is not None and noise_gain > 0.0
    Hd = A.arr(H_all, np.float64) if use_noise else None
    pat = A.arr(pattern, np.float64) if use_noise else None
    mf = A.arr(meanf_all, np.float64)
    vd = A.arr(valid, np.uint8)
    si, sk, kb = _split(A, N, split_index, split_keep)
    dev = fd.device
    n_iter = torch.zeros(N, dtype=torch.int32, device=dev)
    status = torch.zeros(N, dtype=torch.int8, device=dev)
    zncc = torch.zeros(N, dtype=torch.float64, device=dev)
    var = VARIANTS[variant]
    ws_bytes = int(lib().aldvc_icgn12_workspace_bytes(N, int(hx), int(hy), int(hz), int(stride), var))
    extra = 0 if var == 0 else row_moments_bytes(c, vd, hx, hy, hz, stride, (nz, ny, nx), gxd is not None, separable)
    if extra:
        flags |= TUNE_ROW_MOMENTS
        ws_bytes += extra
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    check(
        lib().aldvc_icgn12_tuned(
            _ptr(c), N, _ptr(P), int(hx), int(hy), int(hz), int(stride),
            _ptr(fd), _ptr(gxd), _ptr(gyd), _ptr(gzd), _ptr(md), _ptr(gd), nz, ny, nx, int(mode),
            _ptr(Ld), _ptr(Hd), _ptr(pat), float(noise_gain) if use_noise else 0.0, _ptr(mf), _ptr(vd),
            float(tol), float(dp_tol), int(max_iter), int(patience), int(bool(predictive)),
            _ptr(si), _ptr(sk), kb, _ptr(n_iter), _ptr(status), _ptr(zncc), var, lds_budget, flags, _ptr(ws), ws_bytes,
            stream_handle(stream),
        ),
        "aldvc_icgn12",
    )  # fmt: skip
    return A.out(P), A.out(n_iter), A.out(status), A.out(zncc)


# --------------------------------------------------------------------------- 3-DOF IC-GN


def icgn_3dof_parallel(
    coords, U_old, F_fixed, vdual, hx, hy, hz, f, gx, gy, gz, mask, g, mode, H_all, meanf_all,
    bottomf_all, valid, mu, tol, dp_tol, max_iter, patience, stride=1, n_full=0.0, noise_gain=0.0,
    predictive=True, split_index=None, split_keep=None, stream=None, variant: str = "auto", tuning=None,
    separable=None,
):  # fmt: skip
    """ADMM subproblem-1 translation IC-GN: ``(U_out (N,3), n_iter, status, zncc)``.

    The fused kernel's translation-only form. ``variant="auto"`` (default) takes the autotuner's
    launch shape for this subset shape on this device (timed once on a node sample, then cached);
    ``"fused"``: LDS bricks where a chunk fits one; ``"fused_gather"``: memory gathers only.
    ``tuning`` (an ``autotune.Tuning``) sets the launch shape explicitly; ``separable`` as for
    ``icgn_12dof_parallel``."""
    if tuning is None:
        if variant == "auto":
            from . import autotune

            args = (coords, U_old, F_fixed, vdual, hx, hy, hz, f, gx, gy, gz, mask, g, mode, H_all, meanf_all,
                    bottomf_all, valid, mu, tol, dp_tol, max_iter, patience)  # fmt: skip
            kw = dict(stride=stride, n_full=n_full, noise_gain=noise_gain, predictive=predictive,
                      split_index=split_index, split_keep=split_keep, stream=stream)  # fmt: skip
            tuning = autotune.choose_icgn3(args, kw)
        elif variant in ("fused", "fused_gather"):
            from .autotune import Tuning

            tuning = Tuning(VARIANTS[variant])
        else:
            raise ValueError("variant must be 'auto', 'fused' or 'fused_gather'")
    A = _Args(f)
    c = A.arr(coords, np.int64)
    N = int(c.shape[0])
    Uo = A.arr(U_old, np.float64).reshape(N, 3)
    Fd = A.arr(F_fixed, np.float64).reshape(N, 9)
    vdd = A.arr(vdual, np.float64).reshape(N, 3)
    fd = A.vol(f)
    gd = A.vol(g)
    if _shape(gd) != _shape(fd):
        raise ValueError("f and g must have the same shape")
    nz, ny, nx = _dims(fd)
    gxd, gyd, gzd = _grads(A, f, gx, gy, gz)
    md = _mask(A, f, mask)
    Hd = A.arr(H_all, np.float64)
    mf = A.arr(meanf_all, np.float64)
    bf = A.arr(bottomf_all, np.float64)
    vd = A.arr(valid, np.uint8)
    si, sk, kb = _split(A, N, split_index, split_keep)
    dev = fd.device
    U = torch.empty_like(Uo)
    n_iter = torch.zeros(N, dtype=torch.int32, device=dev)
    status = torch.zeros(N, dtype=torch.int8, device=dev)
    zncc = torch.zeros(N, dtype=torch.float64, device=dev)
    ws_bytes = int(lib().aldvc_icgn3_workspace_bytes(N))
    extra = row_moments_bytes(c, vd, hx, hy, hz, stride, (nz, ny, nx), gxd is not None, separable)
    ws_bytes += extra
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    check(
        lib().aldvc_icgn3_tuned(
            _ptr(c), N, _ptr(Uo), _ptr(U), _ptr(Fd), _ptr(vdd), int(hx), int(hy), int(hz), int(stride),
            _ptr(fd), _ptr(gxd), _ptr(gyd), _ptr(gzd), _ptr(md), _ptr(gd), nz, ny, nx, int(mode),
            _ptr(Hd), _ptr(mf), _ptr(bf), _ptr(vd), float(mu), float(tol), float(dp_tol), int(max_iter),
            int(patience), float(noise_gain) if noise_gain > 0.0 else 0.0, int(bool(predictive)),
            _ptr(si), _ptr(sk), kb, _ptr(n_iter), _ptr(status), _ptr(zncc), None, int(tuning.variant),
            int(tuning.lds_budget), int(tuning.flags) | (TUNE_ROW_MOMENTS if extra else 0), _ptr(ws), ws_bytes,
            stream_handle(stream),
        ),
        "aldvc_icgn3_tuned",
    )  # fmt: skip
    return A.out(U), A.out(n_iter), A.out(status), A.out(zncc)


# --------------------------------------------------------------------------- ZNCC of given warps


def evaluate_zncc_parallel(coords, P_all, hx, hy, hz, f, mask, g, mode, meanf_all, bottomf_all, valid, stride=1):
    """ZNCC of each node's subset under ``P_all[n]`` (NaN if invalid or not sampled)."""
    A = _Args(f)
    c = A.arr(coords, np.int64)
    N = int(c.shape[0])
    P = A.arr(P_all, np.float64).reshape(N, 12)
    fd = A.vol(f)
    gd = A.vol(g)
    nz, ny, nx = _dims(fd)
    md = _mask(A, f, mask)
    vd = A.arr(valid, np.uint8)
    out = torch.empty(N, dtype=torch.float64, device=fd.device)
    check(
        lib().aldvc_evaluate_zncc(
            _ptr(c), N, _ptr(P), int(hx), int(hy), int(hz), int(stride), _ptr(fd), _ptr(md), _ptr(gd),
            nz, ny, nx, int(mode), _ptr(vd), _ptr(out), stream_handle(),
        ),
        "aldvc_evaluate_zncc",
    )  # fmt: skip
    return A.out(out)


# --------------------------------------------------------------------------- NCC maps


def zncc_direct(f, g, coords, hx, hy, hz, rx, ry, rz, wx0, wy0, wz0, t_ok, out=None, precision: str = "fp32"):
    """ZNCC maps ``(n, 2rz+1, 2ry+1, 2rx+1)`` float32; with a host ``out`` array it is filled in place
    (the reference's contract), otherwise the map is returned."""
    A = _Args(f if out is None else out)
    fd = A.vol(f)
    gd = A.vol(g)
    nz, ny, nx = _dims(fd)
    c = A.arr(coords, np.int64)
    n = int(c.shape[0])
    prec = {"fp32": 0, "bf16": 1}[precision]
    maps = torch.empty((n, 2 * rz + 1, 2 * ry + 1, 2 * rx + 1), dtype=torch.float32, device=fd.device)
    if n:
        # every device argument is bound to a name: a temporary's memory would go back to the
        # caching allocator (and be handed to the next temporary) before the launch
        x0d, y0d, z0d = A.arr(wx0, np.int64), A.arr(wy0, np.int64), A.arr(wz0, np.int64)
        okd = A.arr(t_ok, np.uint8)
        ws_bytes = int(lib().aldvc_zncc_map_workspace_bytes(n, hx, hy, hz, rx, ry, rz, prec))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=fd.device)
        check(
            lib().aldvc_zncc_map(
                _ptr(fd), _ptr(gd), nz, ny, nx, _ptr(c), n, int(hx), int(hy), int(hz), int(rx), int(ry), int(rz),
                _ptr(x0d), _ptr(y0d), _ptr(z0d), _ptr(okd), _ptr(maps), prec, _ptr(ws), ws_bytes, stream_handle(),
            ),
            "aldvc_zncc_map",
        )  # fmt: skip
    if out is None:
        return A.out(maps)
    if isinstance(out, torch.Tensor):
        out.copy_(maps)
    else:
        out[...] = maps.cpu().numpy()
    return out


def ncc_peaks(maps, rx, ry, rz, pinv):
    """Peak statistics of ZNCC maps: ``(peak_index, peak_value, pce, sub (n, 3) [x, y, z])``."""
    A = _Args(maps)
    m = A.arr(maps, np.float32)
    n = int(m.shape[0])
    dev = m.device
    pk = torch.empty(n, dtype=torch.int64, device=dev)
    pv = torch.empty(n, dtype=torch.float64, device=dev)
    pce = torch.empty(n, dtype=torch.float64, device=dev)
    sub = torch.empty((n, 3), dtype=torch.float64, device=dev)
    pinv_d = A.arr(pinv, np.float64)
    check(
        lib().aldvc_ncc_peaks(_ptr(m), n, int(rx), int(ry), int(rz), _ptr(pinv_d), _ptr(pk), _ptr(pv), _ptr(pce), _ptr(sub), stream_handle()),
        "aldvc_ncc_peaks",
    )
    return A.out(pk), A.out(pv), A.out(pce), A.out(sub)
