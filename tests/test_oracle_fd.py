"""The oracle's finite-difference global step (oracle/fd_ref.py) against the reference.

Golden vectors: al_dvc 1.2.1's ``build_fd_operators``, ``nodal_gradient``, ``global_rhs`` and
``solve_subpb2`` (direct and PCG) on a masked, a cut and a plain grid (``golden_fd.npz``, written by
``tests/golden/make_golden_fd.py``). Closed forms: difference rows are exact on linear fields,
``Kg`` annihilates constants, and ``Kg`` equals ``sum_j D_j^T D_j`` of dense operators built
node by node from the row rule.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import fd_ref as FD

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = (("a", 1e-3, 1e-3, 0.0), ("b", 0.5, 2e-3, 0.05))


@pytest.fixture(scope="module")
def gf():
    return np.load(os.path.join(HERE, "golden", "golden_fd.npz"))


def _ops(gf, key):
    eo = gf[f"{key}_edge_ok"] if bool(gf[f"{key}_is_cut"]) else None
    return FD.FdOperators(gf[f"{key}_grid_shape"], gf[f"{key}_spacing"], gf[f"{key}_node_valid"], eo)


def _close(a, b, rel=1e-12):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    assert np.max(np.abs(a - b)) <= rel * scale


@pytest.mark.parametrize("key", ["m", "c", "u"])
def test_operators_match_reference(gf, key):
    ops = _ops(gf, key)
    X = gf[f"{key}_X"]
    assert np.array_equal(ops.active, gf[f"{key}_active"])
    for j in range(3):
        _close(ops.D[j] @ X, gf[f"{key}_DX"][j])
        _close(ops.G[j] @ X, gf[f"{key}_GX"][j])
    _close(ops.Kg @ X, gf[f"{key}_KX"])
    _close(ops.M @ X, gf[f"{key}_MX"])


@pytest.mark.parametrize("key", ["m", "c", "u"])
def test_gradient_rhs_and_solve_match_reference(gf, key):
    ops = _ops(gf, key)
    args = [gf[f"{key}_{k}"] for k in ("U1", "F1", "W", "v")]
    F = FD.nodal_gradient(ops, args[0])
    ref = gf[f"{key}_grad"]
    assert np.array_equal(np.isnan(F), np.isnan(ref))
    ok = ~np.isnan(ref)
    _close(F[ok], ref[ok])
    for tag, beta, mu, alpha in CASES:
        _close(FD.global_rhs(ops, beta, mu, alpha, *args), gf[f"{key}_rhs_{tag}"])
        for solver in ("direct", "pcg"):
            U2, it = FD.solve_subpb2(ops, beta, mu, alpha, *args, solver=solver)
            assert np.max(np.abs(U2 - gf[f"{key}_U2_{tag}_{solver}"])) < 1e-9
            if solver == "pcg":
                assert it == int(gf[f"{key}_iter_{tag}_pcg"])


def _dense_rows(shape, spacing, valid, edge_ok):
    """D_j node by node: a row looks one step back and one step forward along the axis."""
    nz, ny, nx = shape
    n = nz * ny * nx
    D = np.zeros((3, n, n))
    for a in range(n):
        ix, iy, iz = a % nx, (a // nx) % ny, a // (nx * ny)
        if not valid[a]:
            continue
        for j, (p, size, s) in enumerate(((ix, nx, 1), (iy, ny, nx), (iz, nz, nx * ny))):
            back = p > 0 and valid[a - s] and edge_ok[a - s, j]
            fwd = p < size - 1 and valid[a + s] and edge_ok[a, j]
            h = spacing[j]
            if back and fwd:
                D[j, a, a - s], D[j, a, a + s] = -0.5 / h, 0.5 / h
            elif fwd:
                D[j, a, a], D[j, a, a + s] = -1.0 / h, 1.0 / h
            elif back:
                D[j, a, a - s], D[j, a, a] = -1.0 / h, 1.0 / h
    return D


@pytest.mark.parametrize("key", ["m", "c", "u"])
def test_closed_forms(gf, key):
    ops = _ops(gf, key)
    shape = tuple(int(s) for s in gf[f"{key}_grid_shape"])
    valid = gf[f"{key}_node_valid"]
    D = _dense_rows(shape, gf[f"{key}_spacing"], valid, gf[f"{key}_edge_ok"])
    for j in range(3):
        assert np.array_equal(ops.D[j].toarray(), D[j])
    assert np.allclose(ops.Kg.toarray(), sum(D[j].T @ D[j] for j in range(3)), rtol=0, atol=1e-15)
    assert np.allclose(ops.Kg @ np.ones(ops.n_nodes), 0.0, atol=1e-12)
    # a linear field: every nonzero difference row is exact
    A = np.array([[0.01, -0.02, 0.03], [0.005, 0.0, -0.01], [0.02, 0.015, -0.005]])
    U = gf[f"{key}_coords"] @ A.T + np.array([0.3, -0.1, 0.2])
    F = FD.nodal_gradient(ops, U)
    for j in range(3):
        live = np.abs(D[j]).sum(axis=1) > 0
        assert np.max(np.abs(F[live, :, j] - A[None, :, j])) < 1e-12
    assert np.all(np.isnan(F[~valid]))
