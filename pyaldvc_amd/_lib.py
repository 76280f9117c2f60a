"""ctypes binding of ``libaldvc_hip.so``, the C ABI of the HIP kernels.

torch is imported first on purpose: it loads its HIP runtime (``libamdhip64.so.7``), and the
library then resolves to that same runtime, so device pointers and streams are shared with
torch tensors. The library is required: there is no CPU fallback, and a missing or stale
build raises :class:`HipLibraryError` (stale: the source hash compiled into the library differs
from the hash of ``csrc/`` on disk).
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (must precede the CDLL load, see the module docstring)

from . import build

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaldvc_hip.so")
ABI_VERSION = 2

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int
F64 = ctypes.c_double
SZ = ctypes.c_size_t

_D, _I, _P, _Z, _L = ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int64


class AdmmLoopDesc(ctypes.Structure):
    """``AldvcAdmmLoop`` of include/aldvc_hip.h, field for field (``struct_bytes`` lets the library
    refuse a layout that differs from its own)."""

    _fields_ = [
        ("struct_bytes", _Z), ("coords", _P), ("n_nodes", _L), ("hx", _I), ("hy", _I), ("hz", _I), ("stride", _I),
        ("f", _P), ("gx", _P), ("gy", _P), ("gz", _P), ("mask", _P), ("g", _P),
        ("vol_nz", _L), ("vol_ny", _L), ("vol_nx", _L), ("mode", _I),
        ("H_all", _P), ("meanf", _P), ("bottomf", _P), ("valid", _P),
        ("icgn_tol", _D), ("icgn_dp_tol", _D), ("icgn_max_iter", _I), ("icgn_patience", _I), ("noise_gain", _D),
        ("predictive", _I), ("split_index", _P), ("split_keep", _P), ("keep_bytes", _L),
        ("variant", _I), ("lds_budget", _I), ("flags", _I),
        ("nx", _I), ("ny", _I), ("nz", _I), ("edge_ok", _P),
        ("outlier_threshold", _D), ("median_eps", _D), ("fill_tol", _D), ("fill_budget", _I), ("fill_limit", _I),
        ("method", _I), ("mats", _P), ("Ae", _P), ("elem_active", _P), ("node_active", _P), ("codes", _P),
        ("spacing_x", _D), ("spacing_y", _D), ("spacing_z", _D),
        ("beta", _D), ("mu", _D), ("alpha", _D), ("pcg_tol", _D), ("pcg_budget", _I), ("pcg_limit", _I),
        ("accumulate", _I), ("admm_tol", _D), ("max_steps", _I),
        ("U1", _P), ("U2", _P), ("v", _P), ("U1raw", _P), ("U1n", _P), ("U2n", _P), ("F2", _P), ("W", _P), ("F2n", _P),
        ("status", _P), ("zncc", _P), ("n_iter", _P), ("bad", _P), ("history", _P),
        ("pcg_iterations", _P), ("fill_iterations", _P), ("stop", _P), ("steps", _P), ("capped", _P),
        ("ws_icgn3", _P), ("ws_cleanup", _P), ("ws_global", _P), ("ws_dual", _P),
        ("ws_icgn3_bytes", _Z), ("ws_cleanup_bytes", _Z), ("ws_global_bytes", _Z), ("ws_dual_bytes", _Z),
    ]  # fmt: skip


# name -> (restype, argtypes); mirrors include/aldvc_hip.h
SIGNATURES = {
    "aldvc_abi_version": (I32, []),
    "aldvc_source_hash": (ctypes.c_char_p, []),
    "aldvc_last_error": (ctypes.c_char_p, []),
    "aldvc_kernel_timing": (I32, [I32]),
    "aldvc_last_kernel_ms": (ctypes.c_float, []),
    "aldvc_workgroup_clock": (I32, [P, I64, ctypes.POINTER(I32)]),
    "aldvc_gradients": (I32, [P, I64, I64, I64, I32, P, P, P, P]),
    "aldvc_sample_points": (I32, [P, I64, I64, I64, P, P, P, I64, I32, P, P]),
    "aldvc_precompute": (
        I32,
        [P, I64, I32, I32, I32, I32, P, P, P, P, P, I64, I64, I64, F64, F64, P, P, I64, P, P, P, P, P, P, P],
    ),
    "aldvc_precompute_tuned": (
        I32,
        [P, I64, I32, I32, I32, I32, P, P, P, P, P, I64, I64, I64, F64, F64, P, P, I64, P, P, P, P, P, P,
         I32, P, SZ, P],
    ),  # fmt: skip
    "aldvc_precompute_row_bytes": (SZ, [I64, I64, I64, I64, I64]),
    "aldvc_icgn12_workspace_bytes": (SZ, [I64, I32, I32, I32, I32, I32]),
    "aldvc_icgn12_row_moments_bytes": (SZ, [I64, I64, I64, I64, I64]),
    "aldvc_icgn12": (
        I32,
        [
            P, I64, P, I32, I32, I32, I32,  # coords, N, P, hx, hy, hz, stride
            P, P, P, P, P, P, I64, I64, I64, I32,  # f, gx, gy, gz, mask, g, nz, ny, nx, mode
            P, P, P, F64, P, P,  # L, H, pattern, noise_gain, meanf, valid
            F64, F64, I32, I32, I32,  # tol, dp_tol, max_iter, patience, predictive
            P, P, I64,  # split_index, split_keep, keep_bytes
            P, P, P, I32, P, SZ, P,  # n_iter, status, zncc, variant, workspace, bytes, stream
        ],
    ),  # fmt: skip
    "aldvc_icgn12_tuned": (
        I32,
        [
            P, I64, P, I32, I32, I32, I32,  # coords, N, P, hx, hy, hz, stride
            P, P, P, P, P, P, I64, I64, I64, I32,  # f, gx, gy, gz, mask, g, nz, ny, nx, mode
            P, P, P, F64, P, P,  # L, H, pattern, noise_gain, meanf, valid
            F64, F64, I32, I32, I32,  # tol, dp_tol, max_iter, patience, predictive
            P, P, I64,  # split_index, split_keep, keep_bytes
            P, P, P, I32, I32, I32, P, SZ, P,  # n_iter, status, zncc, variant, lds_budget, flags, workspace, bytes, stream
        ],
    ),  # fmt: skip
    "aldvc_icgn12_noalloc": (
        I32,
        [
            P, I64, P, I32, I32, I32, I32,  # coords, N, P, hx, hy, hz, stride
            P, P, P, P, P, P, I64, I64, I64, I32,  # f, gx, gy, gz, mask, g, nz, ny, nx, mode
            P, P, P, F64, P, P,  # L, H, pattern, noise_gain, meanf, valid
            F64, F64, I32, I32, I32,  # tol, dp_tol, max_iter, patience, predictive
            P, P, I64,  # split_index, split_keep, keep_bytes
            P, P, P, I32, P,  # n_iter, status, zncc, variant, stream
        ],
    ),  # fmt: skip
    "aldvc_icgn3_workspace_bytes": (SZ, [I64]),
    "aldvc_icgn3": (
        I32,
        [
            P, I64, P, P, P, P, I32, I32, I32, I32,  # coords, N, U_old, U, F, vdual, hx, hy, hz, stride
            P, P, P, P, P, P, I64, I64, I64, I32,  # f, gx, gy, gz, mask, g, nz, ny, nx, mode
            P, P, P, P,  # H, meanf, bottomf, valid
            F64, F64, F64, I32, I32, F64, I32,  # mu, tol, dp_tol, max_iter, patience, noise_gain, predictive
            P, P, I64,  # split
            P, P, P, P, SZ, P,  # n_iter, status, zncc, workspace, bytes, stream
        ],
    ),  # fmt: skip
    "aldvc_icgn3_gated": (
        I32,
        [
            P, I64, P, P, P, P, I32, I32, I32, I32,  # coords, N, U_old, U, F, vdual, hx, hy, hz, stride
            P, P, P, P, P, P, I64, I64, I64, I32,  # f, gx, gy, gz, mask, g, nz, ny, nx, mode
            P, P, P, P,  # H, meanf, bottomf, valid
            F64, F64, F64, I32, I32, F64, I32,  # mu, tol, dp_tol, max_iter, patience, noise_gain, predictive
            P, P, I64,  # split
            P, P, P, P, P, SZ, P,  # n_iter, status, zncc, gate, workspace, bytes, stream
        ],
    ),  # fmt: skip
    "aldvc_icgn3_tuned": (
        I32,
        [
            P, I64, P, P, P, P, I32, I32, I32, I32,  # coords, N, U_old, U, F, vdual, hx, hy, hz, stride
            P, P, P, P, P, P, I64, I64, I64, I32,  # f, gx, gy, gz, mask, g, nz, ny, nx, mode
            P, P, P, P,  # H, meanf, bottomf, valid
            F64, F64, F64, I32, I32, F64, I32,  # mu, tol, dp_tol, max_iter, patience, noise_gain, predictive
            P, P, I64,  # split
            P, P, P, P,  # n_iter, status, zncc, gate
            I32, I32, I32, P, SZ, P,  # variant, lds_budget, flags, workspace, bytes, stream
        ],
    ),  # fmt: skip
    "aldvc_evaluate_zncc": (
        I32,
        [P, I64, P, I32, I32, I32, I32, P, P, P, I64, I64, I64, I32, P, P, P],
    ),
    "aldvc_zncc_map": (
        I32,
        [P, P, I64, I64, I64, P, I64, I32, I32, I32, I32, I32, I32, P, P, P, P, P, I32, P, SZ, P],
    ),
    "aldvc_zncc_map_workspace_bytes": (SZ, [I64, I32, I32, I32, I32, I32, I32, I32]),
    "aldvc_ncc_peaks": (I32, [P, I64, I32, I32, I32, P, P, P, P, P, P]),
    "aldvc_mask_table_bytes": (SZ, [I64, I64, I64]),
    "aldvc_subset_valid_fraction": (I32, [P, I64, I64, I64, P, I64, I32, I32, I32, P, P, SZ, P]),
    "aldvc_split_rows": (I32, [P, P, I64, I32, I32, I32, I32, P, I64, I64, I64, I32, I64, P, P, P, P]),
    "aldvc_voi_moments_workspace_bytes": (SZ, [I64, I64]),
    "aldvc_voi_moments": (I32, [P, I32, I64, I64, I64, I32, I32, I32, I32, I32, I32, F64, P, P, SZ, P]),
    "aldvc_normalize": (I32, [P, I32, I64, F64, F64, P, P]),
    "aldvc_bspline_prefilter": (I32, [P, I64, I64, I64, P, P, P]),
    "aldvc_mask_nan": (I32, [P, P, I64, P]),
    "aldvc_block_downsample": (I32, [P, I64, I64, I64, I32, P, P]),
    "aldvc_bridging_elements": (I32, [P, I64, I64, I64, P, I64, I32, P, P]),
    "aldvc_displacement_uncertainty": (I32, [P, P, P, P, P, P, I64, P, F64, I32, P, P, P]),
    "aldvc_fem_apply": (I32, [P, I32, I32, I32, P, P, P, P, P]),
    "aldvc_fem_rhs": (I32, [P, I32, I32, I32, P, P, F64, F64, F64, P, P, P, P, P]),
    "aldvc_fem_nodal_gradient": (I32, [P, I32, I32, I32, P, P, P, P, P]),
    "aldvc_pcg_workspace_bytes": (SZ, [I32, I32, I32]),
    "aldvc_pcg_fem": (I32, [P, I32, I32, I32, P, P, P, P, F64, I32, ctypes.POINTER(I32), P, SZ, P]),
    "aldvc_pcg_fem_async": (I32, [P, I32, I32, I32, P, P, P, P, F64, I32, P, P, P, SZ, P]),
    "aldvc_admm_global_workspace_bytes": (SZ, [I32, I32, I32]),
    "aldvc_admm_global_step": (
        I32,
        [P, P, I32, I32, I32, P, P, F64, F64, F64, P, P, P, P, F64, I32, P, P, P, P, I32, P, P, SZ, P],
    ),
    "aldvc_fd_codes": (I32, [I32, I32, I32, P, P, P, P]),
    "aldvc_fd_apply": (I32, [P, I32, I32, I32, F64, F64, F64, F64, F64, P, P, P]),
    "aldvc_fd_rhs": (I32, [P, I32, I32, I32, F64, F64, F64, F64, F64, F64, P, P, P, P, P, P]),
    "aldvc_fd_nodal_gradient": (I32, [P, I32, I32, I32, F64, F64, F64, P, P, P]),
    "aldvc_pcg_fd": (I32, [P, I32, I32, I32, F64, F64, F64, F64, F64, P, P, F64, I32, ctypes.POINTER(I32), P, SZ, P]),
    "aldvc_admm_global_step_fd": (
        I32,
        [P, I32, I32, I32, F64, F64, F64, F64, F64, F64, P, P, P, P, F64, I32, P, P, P, P, I32, P, P, SZ, P],
    ),
    "aldvc_admm_loop_enqueue": (I32, [ctypes.POINTER(AdmmLoopDesc), I32, P]),
    "aldvc_admm_loop_graph": (I32, [ctypes.POINTER(AdmmLoopDesc), P, ctypes.POINTER(P)]),
    "aldvc_graph_launch": (I32, [P, P]),
    "aldvc_graph_destroy": (I32, [P]),
    "aldvc_subpb1_cleanup_workspace_bytes": (SZ, [I32, I32, I32]),
    "aldvc_subpb1_cleanup": (I32, [I32, I32, I32, P, P, P, P, F64, F64, F64, I32, P, P, P, I32, P, P, P, SZ, P]),
    "aldvc_spring_fill_workspace_bytes": (SZ, [I32, I32, I32]),
    "aldvc_spring_fill": (I32, [I32, I32, I32, P, P, F64, I32, ctypes.POINTER(I32), P, SZ, P]),
    "aldvc_dual_update_workspace_bytes": (SZ, [I64]),
    "aldvc_dual_update": (I32, [I64, P, P, P, P, P, P, P, P, I32, I32, F64, I32, P, P, P, P, SZ, P]),
    "aldvc_warp_znssd": (
        I32,
        [P, P, I64, I64, I64, I32, P, I64, I32, I32, I32, P, I32, I32, I32, F64, F64, F64, F64, F64, F64, P, P, P],
    ),
    "aldvc_warp_backward_workspace_bytes": (SZ, [I64, I64, I64, I64, I32, I32, I32, I32, I32, I32]),
    "aldvc_warp_znssd_backward": (
        I32,
        [P, P, I64, I64, I64, I32, P, I64, I32, I32, I32, P, I32, I32, I32, F64, F64, F64, F64, F64, F64,
         P, P, P, P, P, SZ, P],
    ),  # fmt: skip
}


class HipLibraryError(RuntimeError):
    """The HIP library is missing, stale, or a call into it failed."""


_lib = None
_lock = threading.Lock()


def lib():
    """The loaded library (raises :class:`HipLibraryError` when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise HipLibraryError(
                f"{LIB_PATH} is not built; run `python -m pyaldvc_amd.build` (or __graft_entry__.build())"
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                raise HipLibraryError(f"{LIB_PATH} does not export {name}; rebuild it")
            fn.restype = res
            fn.argtypes = args
        v = L.aldvc_abi_version()
        if v != ABI_VERSION:
            raise HipLibraryError(f"ABI version {v} of {LIB_PATH} != {ABI_VERSION} expected; rebuild it")
        built = (L.aldvc_source_hash() or b"").decode()
        if os.path.isdir(build.CSRC):
            disk = build.source_hash()
            if built != disk:
                raise HipLibraryError(
                    f"{LIB_PATH} was built from other sources (hash {built}, sources on disk {disk}); "
                    "rebuild it with `python -m pyaldvc_amd.build`"
                )
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    """Raise with the library's message when a call returned a nonzero code."""
    if rc != 0:
        msg = lib().aldvc_last_error()
        msg = msg.decode() if msg else ""
        raise HipLibraryError(f"{what} failed ({rc}): {msg}")
