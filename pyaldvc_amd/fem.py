"""ADMM subproblem 2 on the GPU: the operators of the node grid (hex8 FEM or finite
differences), the right-hand side, the float64 Jacobi-PCG solve and the nodal gradient
(solver/global_operators.py, solver/subpb2_solver.py, mesh/hex8.py).

Both discretisations are applied matrix-free by the HIP library (``csrc/aldvc_global.hip``). The
grid is uniform, so for FEM every element is the same box and an operator is one 8x8 element matrix
summed over the active elements; for FD (``subpb2_method="fd"``) a node's rows are fixed by one
byte of neighbour flags (valid neighbours, cut edges) and the spacing. The API mirrors the
reference: ``build_global_operators(mesh, method, order)``,
``build_global_system(ops, beta, mu, alpha, para)``, ``global_rhs(...)``,
``solve_subpb2(system, U1, F1, W, v)`` and ``nodal_gradient(ops, U)``, host arrays in and out.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import kernels as K
from ._lib import check, lib

# natural coordinates of the hex8 corners in the grid's element order (mesh/hex8.py:16-29)
HEX8_SIGNS = np.array(
    [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64
)
DIRECT_MAX_ACTIVE = 5_000  # "auto" is the direct solver up to here (subpb2_solver.py:27)
DIRECT_TOL = 1e-13  # the "direct" solver: CG driven to round-off


def gauss_legendre(order: int):
    """1-D Gauss-Legendre points and weights on [-1, 1] for 1, 2 or 3 points."""
    if order == 1:
        return np.array([0.0]), np.array([2.0])
    if order == 2:
        g = 1.0 / np.sqrt(3.0)
        return np.array([-g, g]), np.array([1.0, 1.0])
    if order == 3:
        g = np.sqrt(0.6)
        return np.array([-g, 0.0, g]), np.array([5.0, 8.0, 5.0]) / 9.0
    raise ValueError(f"Gauss order {order} not supported (use 1, 2 or 3)")


def hex8_box_matrices(spacing, order: int = 2) -> dict:
    """Element stiffness ``K = int grad N grad N^T``, mass ``M = int N N^T`` and gradient-mass
    ``G_j[a, b] = int dN_a/dx_j N_b`` of an axis-aligned box with edge lengths ``spacing``
    (x, y, z), by tensor Gauss quadrature (mesh/hex8.py:74-105)."""
    h = np.asarray(spacing, dtype=np.float64)
    half = h / 2.0
    pts, wts = gauss_legendre(order)
    Km = np.zeros((8, 8))
    Mm = np.zeros((8, 8))
    Gm = np.zeros((3, 8, 8))
    s = HEX8_SIGNS
    for wk, zeta in zip(wts, pts):
        for wj, eta in zip(wts, pts):
            for wi, ksi in zip(wts, pts):
                xi = np.array([ksi, eta, zeta])
                f = 1.0 + s * xi  # (8, 3) one-dimensional factors
                N = 0.125 * f.prod(axis=1)
                dN = np.empty((8, 3))
                for d in range(3):
                    others = [e for e in range(3) if e != d]
                    dN[:, d] = 0.125 * s[:, d] * f[:, others[0]] * f[:, others[1]] / half[d]
                w = wi * wj * wk * half.prod()
                Km += w * dN @ dN.T
                Mm += w * np.outer(N, N)
                for d in range(3):
                    Gm[d] += w * np.outer(dN[:, d], N)
    return {"K": Km, "M": Mm, "G": Gm, "volume": float(h.prod())}


def hex8_elements(grid_shape, node_valid=None) -> np.ndarray:
    """``(E, 8)`` corner nodes of every grid cell, cells in (z, y, x) order, corners in the hex8
    order of the grid (mesh/grid_mesh.py:79-101); rows of -1 for cells with an invalid corner
    (grid_mesh.py:217-219)."""
    nz, ny, nx = (int(v) for v in grid_shape)
    iz, iy, ix = (a.ravel() for a in np.meshgrid(np.arange(nz - 1), np.arange(ny - 1), np.arange(nx - 1), indexing="ij"))
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    el = np.column_stack([(iz + dz) * ny * nx + (iy + dy) * nx + (ix + dx) for dx, dy, dz in corners]).astype(np.int64)
    if node_valid is not None:
        bad = ~np.asarray(node_valid, dtype=bool)[el].all(axis=1)
        el[bad, :] = -1
    return el


@dataclass
class GlobalOperators:
    """The operators of a node grid on the GPU: for FEM the element matrices and the element and
    node activity, for FD the per-node codes and the spacing."""

    mats: torch.Tensor  # FEM: (5, 8, 8) float64: K, M, G_x, G_y, G_z
    elem_active: torch.Tensor  # FEM: (nz-1, ny-1, nx-1) uint8
    node_active: torch.Tensor  # (nz, ny, nx) uint8
    grid_shape: tuple
    active: np.ndarray  # (N,) bool, nodes that carry equations
    method: str
    n_nodes: int
    codes: torch.Tensor = None  # FD: (nz, ny, nx) uint8 neighbour flags (aldvc_fd_codes)
    spacing: tuple = None  # FD: (hx, hy, hz)

    @property
    def n_active(self) -> int:
        return int(self.active.sum())


def _fd_edges(mesh, n: int):
    """The +x, +y, +z edge flags of a cut mesh, or ``None`` when nothing was cut (the reference
    takes them only from a cut mesh, global_operators.py:126)."""
    e = getattr(mesh, "edge_ok", None)
    if e is None:
        return None
    e = np.asarray(e, dtype=bool)
    return e.reshape(n, 3) if e.size == 3 * n and not e.all() else None


def build_fd_operators(mesh) -> GlobalOperators:
    """Finite-difference operators of the node grid (global_operators.py:74-137): ``mesh`` gives
    ``grid_shape``, ``spacing`` (x, y, z), ``node_valid`` (all valid when it does not hold one flag
    per node) and, for a cut mesh, ``edge_ok`` (N, 3)."""
    nz, ny, nx = (int(v) for v in mesh.grid_shape)
    n = nz * ny * nx
    nv = getattr(mesh, "node_valid", None)
    valid = np.asarray(nv, dtype=bool).ravel() if nv is not None and np.size(nv) == n else np.ones(n, dtype=bool)
    edges = _fd_edges(mesh, n)
    dev = K.device()
    valid_d = torch.from_numpy(valid.astype(np.uint8)).to(dev)
    edges_d = None if edges is None else torch.from_numpy(np.ascontiguousarray(edges.astype(np.uint8))).to(dev)
    codes = torch.empty(n, dtype=torch.uint8, device=dev)
    check(lib().aldvc_fd_codes(nx, ny, nz, K._ptr(valid_d), K._ptr(edges_d), K._ptr(codes), K.stream_handle()),
          "aldvc_fd_codes")  # fmt: skip
    spacing = tuple(float(h) for h in mesh.spacing)
    if min(spacing) <= 0.0:
        raise ValueError("the grid spacing must be positive")
    return GlobalOperators(
        mats=None, elem_active=None, node_active=valid_d.reshape(nz, ny, nx), grid_shape=(nz, ny, nx),
        active=valid.copy(), method="fd", n_nodes=n, codes=codes.reshape(nz, ny, nx), spacing=spacing,
    )  # fmt: skip


def build_global_operators(mesh, method: str = "fem", gauss_pt_order: int = 2) -> GlobalOperators:
    """The operators of ``subpb2_method`` ``method`` (global_operators.py:140-144): FEM over the
    active elements of ``mesh`` (``mesh.elements`` (E, 8) with rows of -1 for dropped elements, in
    the grid's element order; :58-70), or finite differences (``build_fd_operators``)."""
    if method == "fd":
        return build_fd_operators(mesh)
    if method != "fem":
        raise ValueError(f"unknown subpb2 method {method!r}")
    nz, ny, nx = (int(v) for v in mesh.grid_shape)
    elements = np.asarray(mesh.elements)
    if elements.shape != ((nz - 1) * (ny - 1) * (nx - 1), 8):
        raise ValueError("elements must hold one row per grid cell, in the grid's cell order")
    ea = elements[:, 0] >= 0
    if not ea.any():
        raise ValueError("Mesh has no active elements; cannot build FEM operators.")
    active = np.zeros(nz * ny * nx, dtype=bool)
    active[np.unique(elements[ea])] = True
    box = hex8_box_matrices(mesh.spacing, gauss_pt_order)
    mats = np.stack([box["K"], box["M"], *box["G"]])
    dev = K.device()
    return GlobalOperators(
        mats=torch.from_numpy(np.ascontiguousarray(mats)).to(dev),
        elem_active=torch.from_numpy(ea.astype(np.uint8).reshape(nz - 1, ny - 1, nx - 1)).to(dev),
        node_active=torch.from_numpy(active.astype(np.uint8).reshape(nz, ny, nx)).to(dev),
        grid_shape=(nz, ny, nx),
        active=active,
        method="fem",
        n_nodes=nz * ny * nx,
    )


@dataclass
class GlobalSystem:
    ops: GlobalOperators
    beta: float
    mu: float
    alpha: float
    Ae: torch.Tensor  # FEM: (8, 8) element matrix of (beta + alpha) K + mu M; None for FD
    solver: str
    pcg_tol: float
    pcg_max_iter: int


def build_global_system(ops: GlobalOperators, beta: float, mu: float, alpha: float, para) -> GlobalSystem:
    """``A = (beta + alpha) K + mu M`` on the active nodes (subpb2_solver.py:45-74). The
    reference's sparse LU for small systems ("direct") is conjugate gradients to round-off here."""
    Ae = None if ops.method == "fd" else ((beta + alpha) * ops.mats[0] + mu * ops.mats[1]).contiguous()
    solver = getattr(para, "global_solver", "auto")
    if solver == "auto":
        solver = "direct" if ops.n_active <= DIRECT_MAX_ACTIVE else "pcg"
    return GlobalSystem(ops, float(beta), float(mu), float(alpha), Ae, solver, float(para.pcg_tol), int(para.pcg_max_iter))


def _grid(ops):
    nz, ny, nx = ops.grid_shape
    return nx, ny, nz, K._ptr(ops.elem_active), K._ptr(ops.node_active)


def _dev(a, shape):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shape))).to(K.device())


def _fd(ops):
    """The FD grid arguments of the C ABI: codes, nx, ny, nz, hx, hy, hz."""
    nz, ny, nx = ops.grid_shape
    return (K._ptr(ops.codes), nx, ny, nz, *ops.spacing)


def _rhs_device(ops, beta, mu, alpha, U1, F1, W, v) -> torch.Tensor:
    n = ops.n_nodes
    if ops.method == "fd":
        args = [_dev(a, s) for a, s in ((U1, (n, 3)), (F1, (n, 3, 3)), (W, (n, 3, 3)), (v, (n, 3)))]
        b = torch.empty((n, 3), dtype=torch.float64, device=args[0].device)
        check(lib().aldvc_fd_rhs(*_fd(ops), float(beta), float(mu), float(alpha), *(K._ptr(a) for a in args), K._ptr(b),
                                 K.stream_handle()), "aldvc_fd_rhs")  # fmt: skip
        return b
    FmW = np.asarray(F1, dtype=np.float64).reshape(n, 3, 3) - np.asarray(W, dtype=np.float64).reshape(n, 3, 3)
    FmW_j = _dev(np.transpose(FmW, (2, 0, 1)), (3, n, 3))  # [j][node][i] = (F - W)[node, i, j]
    Umv = _dev(np.asarray(U1, dtype=np.float64).reshape(n, 3) - np.asarray(v, dtype=np.float64).reshape(n, 3), (n, 3))
    U = _dev(U1, (n, 3))
    b = torch.empty((n, 3), dtype=torch.float64, device=U.device)
    check(
        lib().aldvc_fem_rhs(K._ptr(ops.mats), *_grid(ops), float(beta), float(mu), float(alpha), K._ptr(FmW_j),
                            K._ptr(Umv), K._ptr(U), K._ptr(b), K.stream_handle()),
        "aldvc_fem_rhs",
    )  # fmt: skip
    return b


def global_rhs(ops, beta, mu, alpha, U1, F1, W, v) -> np.ndarray:
    """``b_i = beta sum_j G_j (F1 - W)_ij + mu M (U1 - v)_i + alpha K U1_i`` as ``(N, 3)``
    (subpb2_solver.py:108-127)."""
    return _rhs_device(ops, beta, mu, alpha, U1, F1, W, v).cpu().numpy()


def pcg_fem(system: GlobalSystem, B: torch.Tensor, X0: torch.Tensor, tol: float, max_iter: int):
    """Jacobi-preconditioned CG on the active nodes for the three columns of ``B`` at once
    (subpb2_solver.py:77-105), for either discretisation: ``(X, iterations)``."""
    ops = system.ops
    nz, ny, nx = ops.grid_shape
    X = X0.clone()
    it = ctypes.c_int(0)
    ws = torch.empty(int(lib().aldvc_pcg_workspace_bytes(nx, ny, nz)), dtype=torch.uint8, device=X.device)
    if ops.method == "fd":
        check(lib().aldvc_pcg_fd(*_fd(ops), system.beta + system.alpha, system.mu, K._ptr(B), K._ptr(X), float(tol),
                                 int(max_iter), ctypes.byref(it), K._ptr(ws), int(ws.numel()), K.stream_handle()),
              "aldvc_pcg_fd")  # fmt: skip
        return X, int(it.value)
    check(
        lib().aldvc_pcg_fem(K._ptr(system.Ae), *_grid(ops), K._ptr(B), K._ptr(X), float(tol), int(max_iter),
                            ctypes.byref(it), K._ptr(ws), int(ws.numel()), K.stream_handle()),
        "aldvc_pcg_fem",
    )  # fmt: skip
    return X, int(it.value)


def solve_subpb2(system: GlobalSystem, U1, F1, W, v):
    """``u_hat`` (N, 3) of the global step; inactive nodes keep ``U1`` (subpb2_solver.py:130-155)."""
    t0 = time.perf_counter()
    ops = system.ops
    n = ops.n_nodes
    b = _rhs_device(ops, system.beta, system.mu, system.alpha, U1, F1, W, v)
    act = torch.from_numpy(ops.active).to(b.device)
    b[~act] = 0.0
    X0 = _dev(U1, (n, 3))
    if system.solver == "direct":
        X, n_it = pcg_fem(system, b, X0, DIRECT_TOL, max(system.pcg_max_iter, 10 * ops.n_active))
        n_it = 0
    else:
        X, n_it = pcg_fem(system, b, X0, system.pcg_tol, system.pcg_max_iter)
    U2 = np.asarray(U1, dtype=np.float64).reshape(n, 3).copy()
    U2[ops.active] = X.cpu().numpy()[ops.active]
    return U2, {"solver": system.solver, "iterations": int(n_it), "time": time.perf_counter() - t0}


def nodal_gradient(ops: GlobalOperators, U) -> np.ndarray:
    """``(N, 3, 3)`` displacement gradient, NaN at inactive nodes (global_operators.py:151-170):
    FEM the lumped-L2 gradient ``F_ij = (G_j^T u_i) / m_L``, FD ``F_ij = D_j u_i``."""
    n = ops.n_nodes
    Ud = _dev(U, (n, 3))
    F = torch.empty((n, 3, 3), dtype=torch.float64, device=Ud.device)
    if ops.method == "fd":
        check(lib().aldvc_fd_nodal_gradient(*_fd(ops), K._ptr(Ud), K._ptr(F), K.stream_handle()), "aldvc_fd_nodal_gradient")
        return F.cpu().numpy()
    check(lib().aldvc_fem_nodal_gradient(K._ptr(ops.mats), *_grid(ops), K._ptr(Ud), K._ptr(F), K.stream_handle()),
          "aldvc_fem_nodal_gradient")  # fmt: skip
    return F.cpu().numpy()
