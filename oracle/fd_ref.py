"""Test infrastructure (the oracle): the finite-difference global step in float64 SciPy.

Restates solver/global_operators.py:74-137 (``_difference_operator``, ``build_fd_operators``), the
FD branch of ``nodal_gradient`` (:151-170) and solver/subpb2_solver.py:45-155 (``global_rhs``, the
system ``(beta + alpha) Kg + mu M`` on the active nodes, its Jacobi-preconditioned CG and the direct
solve); only the tests, ``__graft_entry__.smoke()`` and the bench's CPU leg use it. Pinned in
tests/test_oracle_fd.py against the reference's own outputs on three grids (masked, cut, plain;
``tests/golden/golden_fd.npz``) and by closed forms: every difference row is exact on a linear
field, ``Kg`` annihilates constants, and ``Kg`` equals ``sum_j D_j^T D_j`` built independently as
dense matrices from the row rule.
"""

from __future__ import annotations

import numpy as np
from scipy import sparse
from scipy.sparse.linalg import spsolve


def neighbour_flags(grid_shape, axis: int, valid, edge_ok=None):
    """``(has_prev, has_next)`` per node along ``axis`` (0 = x, 1 = y, 2 = z): the neighbour is on
    the grid, valid, and (with ``edge_ok``, the +axis edge flag of every node) not across a cut
    edge (global_operators.py:90-104)."""
    nz, ny, nx = grid_shape
    n = nz * ny * nx
    stride = (1, nx, nx * ny)[axis]
    size = (nx, ny, nz)[axis]
    pos = (np.arange(n) // stride) % size
    v = np.asarray(valid, dtype=bool).ravel()
    prev = np.zeros(n, dtype=bool)
    nxt = np.zeros(n, dtype=bool)
    prev[pos > 0] = v[np.flatnonzero(pos > 0) - stride]
    nxt[pos < size - 1] = v[np.flatnonzero(pos < size - 1) + stride]
    if edge_ok is not None:
        e = np.asarray(edge_ok, dtype=bool).ravel()
        nxt &= e
        prev[pos > 0] &= e[np.flatnonzero(pos > 0) - stride]
    return prev, nxt


def difference_operator(grid_shape, axis: int, h: float, valid, edge_ok=None) -> sparse.csr_matrix:
    """Row ``a`` of ``D``: ``(x[a+s] - x[a-s]) / 2h`` with both neighbours, the one-sided difference
    with one, zero with none or at an invalid node (global_operators.py:74-120)."""
    nz, ny, nx = grid_shape
    n = nz * ny * nx
    stride = (1, nx, nx * ny)[axis]
    prev, nxt = neighbour_flags(grid_shape, axis, valid, edge_ok)
    v = np.asarray(valid, dtype=bool).ravel()
    a = np.arange(n)
    rows, cols, vals = [], [], []
    for sel, lo, hi, w in ((v & prev & nxt, -stride, stride, 0.5 / h), (v & nxt & ~prev, 0, stride, 1.0 / h),
                           (v & prev & ~nxt, -stride, 0, 1.0 / h)):  # fmt: skip
        r = a[sel]
        rows += [r, r]
        cols += [r + lo, r + hi]
        vals += [np.full(r.size, -w), np.full(r.size, w)]
    return sparse.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


class FdOperators:
    """``D_j``, ``Kg = sum_j D_j^T D_j``, ``M = diag(valid)``, ``G_j = D_j^T``, ``active = valid``
    (global_operators.py:123-137)."""

    def __init__(self, grid_shape, spacing, valid, edge_ok=None):
        self.grid_shape = tuple(int(s) for s in grid_shape)
        n = int(np.prod(self.grid_shape))
        self.n_nodes = n
        self.active = np.asarray(valid, dtype=bool).ravel().copy()
        eo = None if edge_ok is None else np.asarray(edge_ok, dtype=bool).reshape(n, 3)
        self.D = [difference_operator(self.grid_shape, j, float(spacing[j]), self.active, None if eo is None else eo[:, j])
                  for j in range(3)]  # fmt: skip
        self.Kg = sum((d.T @ d for d in self.D), sparse.csr_matrix((n, n))).tocsr()
        self.M = sparse.diags(self.active.astype(np.float64)).tocsr()
        self.G = [d.T.tocsr() for d in self.D]


def nodal_gradient(ops: FdOperators, U) -> np.ndarray:
    """``F[:, i, j] = D_j U[:, i]``, NaN at inactive nodes (global_operators.py:151-170)."""
    U = np.asarray(U, dtype=np.float64).reshape(ops.n_nodes, 3)
    F = np.stack([np.stack([ops.D[j] @ U[:, i] for j in range(3)], 1) for i in range(3)], 1)
    F[~ops.active] = np.nan
    return F


def global_rhs(ops: FdOperators, beta, mu, alpha, U1, F1, W, v) -> np.ndarray:
    """``mu M (U1 - v) + beta sum_j G_j (F1 - W)[:, :, j] + alpha Kg U1`` (subpb2_solver.py:108-127)."""
    n = ops.n_nodes
    FmW = np.asarray(F1, np.float64).reshape(n, 3, 3) - np.asarray(W, np.float64).reshape(n, 3, 3)
    U1 = np.asarray(U1, np.float64).reshape(n, 3)
    b = mu * (ops.M @ (U1 - np.asarray(v, np.float64).reshape(n, 3)))
    for j in range(3):
        b = b + beta * (ops.G[j] @ FmW[:, :, j])
    if alpha != 0.0:
        b = b + alpha * (ops.Kg @ U1)
    return b


def solve_subpb2(ops: FdOperators, beta, mu, alpha, U1, F1, W, v, solver="pcg", tol=1e-8, max_iter=500):
    """``(U2, iterations)``: the system on the active nodes solved directly or by Jacobi-PCG from
    ``U1`` for the three columns at once; inactive nodes keep ``U1`` (subpb2_solver.py:45-155)."""
    n = ops.n_nodes
    U1 = np.asarray(U1, np.float64).reshape(n, 3)
    idx = np.flatnonzero(ops.active)
    A = ((beta + alpha) * ops.Kg + mu * ops.M).tocsr()[idx][:, idx].tocsr()
    b = global_rhs(ops, beta, mu, alpha, U1, F1, W, v)[idx]
    if solver == "direct":
        X, it = np.column_stack([spsolve(A.tocsc(), b[:, c]) for c in range(3)]), 0
    else:
        d = A.diagonal()
        inv = np.where(np.abs(d) > 1e-300, 1.0 / np.where(d == 0, 1.0, d), 1.0)
        X = U1[idx].copy()
        R = b - A @ X
        Z = inv[:, None] * R
        P = Z.copy()
        rz = np.sum(R * Z, axis=0)
        bn = np.linalg.norm(b, axis=0)
        bn[bn == 0] = 1.0
        it = 0
        for it in range(1, max_iter + 1):
            AP = A @ P
            den = np.sum(P * AP, axis=0)
            al = rz / np.where(den == 0, np.inf, den)
            X += P * al
            R -= AP * al
            if np.all(np.linalg.norm(R, axis=0) / bn < tol):
                break
            Z = inv[:, None] * R
            rz_new = np.sum(R * Z, axis=0)
            P = Z + P * (rz_new / np.where(rz == 0, np.inf, rz))
            rz = rz_new
    U2 = U1.copy()
    U2[idx] = X
    return U2, it
