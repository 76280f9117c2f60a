"""The finite-difference global step on the GPU against the reference (needs an MI355X).

Golden vectors: al_dvc 1.2.1's ``build_fd_operators`` (by its action on seeded vectors),
``nodal_gradient``, ``global_rhs`` and ``solve_subpb2`` (direct and PCG) on a masked, a cut and a
plain grid (``golden_fd.npz``). The reference's own closed forms (tests/test_global_step.py): the
gradient of a linear field is exact, ``Kg`` annihilates constants, a large ``mu`` returns ``U1`` and a
large ``beta`` with a compatible ``F`` returns a field of that gradient. On a 23 x 17 x 13 grid with
scattered invalid nodes and cut edges, the device path against the oracle's restatement.
"""

from __future__ import annotations

import os
from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FEM = pytest.importorskip("pyaldvc_amd.fem")
K = pytest.importorskip("pyaldvc_amd.kernels")
from oracle import fd_ref as FDR  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = (("a", 1e-3, 1e-3, 0.0), ("b", 0.5, 2e-3, 0.05))


@pytest.fixture(scope="module")
def gf():
    return np.load(os.path.join(HERE, "golden", "golden_fd.npz"))


def _mesh(gf, key):
    eo = gf[f"{key}_edge_ok"] if bool(gf[f"{key}_is_cut"]) else None
    return SimpleNamespace(grid_shape=tuple(int(s) for s in gf[f"{key}_grid_shape"]), spacing=tuple(gf[f"{key}_spacing"]),
                           node_valid=gf[f"{key}_node_valid"], edge_ok=eo)  # fmt: skip


def _apply(ops, cK, cM, X):
    import torch

    nz, ny, nx = ops.grid_shape
    Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float64)).to(K.device())
    Y = torch.empty_like(Xd)
    K.check(K.lib().aldvc_fd_apply(K._ptr(ops.codes), nx, ny, nz, *ops.spacing, float(cK), float(cM), K._ptr(Xd),
                                   K._ptr(Y), K.stream_handle()), "aldvc_fd_apply")  # fmt: skip
    return Y.cpu().numpy()


def _close(a, b, rel=1e-12):
    assert np.max(np.abs(a - b)) <= rel * max(float(np.max(np.abs(b))), 1e-300)


@pytest.mark.parametrize("key", ["m", "c", "u"])
def test_operators_match_reference(gf, key):
    ops = FEM.build_global_operators(_mesh(gf, key), "fd")
    assert ops.method == "fd" and np.array_equal(ops.active, gf[f"{key}_active"])
    X = gf[f"{key}_X"]
    _close(_apply(ops, 1.0, 0.0, X), gf[f"{key}_KX"])
    _close(_apply(ops, 0.0, 1.0, X), gf[f"{key}_MX"])
    _close(_apply(ops, 0.7, 0.3, X), 0.7 * gf[f"{key}_KX"] + 0.3 * gf[f"{key}_MX"])
    F = FEM.nodal_gradient(ops, X)  # F[:, i, j] = D_j X[:, i] at valid nodes
    live = ops.active
    for j in range(3):
        _close(F[live, :, j], gf[f"{key}_DX"][j][live])


@pytest.mark.parametrize("key", ["m", "c", "u"])
def test_gradient_rhs_and_solve_match_reference(gf, key):
    ops = FEM.build_global_operators(_mesh(gf, key), "fd")
    args = [gf[f"{key}_{k}"] for k in ("U1", "F1", "W", "v")]
    F = FEM.nodal_gradient(ops, args[0])
    ref = gf[f"{key}_grad"]
    assert np.array_equal(np.isnan(F), np.isnan(ref))
    ok = ~np.isnan(ref)
    _close(F[ok], ref[ok])
    for tag, beta, mu, alpha in CASES:
        _close(FEM.global_rhs(ops, beta, mu, alpha, *args), gf[f"{key}_rhs_{tag}"])
        for solver in ("pcg", "direct"):
            para = SimpleNamespace(global_solver=solver, pcg_tol=1e-8, pcg_max_iter=500)
            U2, info = FEM.solve_subpb2(FEM.build_global_system(ops, beta, mu, alpha, para), *args)
            assert np.array_equal(U2[~ops.active], args[0][~ops.active])
            assert np.max(np.abs(U2 - gf[f"{key}_U2_{tag}_{solver}"])) < 1e-9, (tag, solver)
            if solver == "pcg":
                assert info["iterations"] == int(gf[f"{key}_iter_{tag}_pcg"])


def _plain_mesh(shape=(6, 6, 6), h=8.0):
    ax = [np.arange(n) * h for n in shape[::-1]]  # x, y, z
    gz, gy, gx = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    coords = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
    return SimpleNamespace(grid_shape=shape, spacing=(h, h, h), node_valid=np.ones(coords.shape[0], bool)), coords


def test_reference_closed_forms():
    """tests/test_global_step.py:17-51 of the reference, with the FD operators."""
    mesh, coords = _plain_mesh()
    ops = FEM.build_global_operators(mesh, "fd")
    A = np.array([[0.01, 0.02, -0.03], [0.04, -0.05, 0.06], [0.07, 0.08, 0.09]])
    F = FEM.nodal_gradient(ops, coords @ A.T + np.array([1.0, 2.0, 3.0]))
    assert np.allclose(F, A, atol=1e-10)
    assert np.allclose(_apply(ops, 1.0, 0.0, np.ones((ops.n_nodes, 3))), 0.0, atol=1e-10)
    assert ops.active.all()
    rng = np.random.default_rng(0)
    A = np.diag([0.01, -0.02, 0.03])
    U_lin = coords @ A.T
    U1 = U_lin + 0.05 * rng.standard_normal(U_lin.shape)
    F1 = np.tile(A, (ops.n_nodes, 1, 1))
    W, v = np.zeros_like(F1), np.zeros_like(U1)
    para = SimpleNamespace(global_solver="direct", pcg_tol=1e-8, pcg_max_iter=500)
    U2, _ = FEM.solve_subpb2(FEM.build_global_system(ops, 1e-12, 1.0, 0.0, para), U1, F1, W, v)
    assert np.allclose(U2, U1, atol=1e-8)
    U2, _ = FEM.solve_subpb2(FEM.build_global_system(ops, 1e6, 1e-6, 0.0, para), U1, F1, W, v)
    assert np.allclose(FEM.nodal_gradient(ops, U2), A, atol=1e-4)


def test_larger_cut_grid_matches_the_oracle():
    shape = (13, 17, 23)  # (nz, ny, nx)
    n = int(np.prod(shape))
    rng = np.random.default_rng(5)
    valid = rng.random(n) > 0.08
    edge_ok = rng.random((n, 3)) > 0.05
    mesh = SimpleNamespace(grid_shape=shape, spacing=(4.0, 5.0, 3.0), node_valid=valid, edge_ok=edge_ok)
    ops = FEM.build_global_operators(mesh, "fd")
    ref = FDR.FdOperators(shape, mesh.spacing, valid, edge_ok)
    X = rng.normal(size=(n, 3))
    _close(_apply(ops, 1.0, 0.0, X), ref.Kg @ X)
    U1, v = rng.normal(0, 1.5, (n, 3)), rng.normal(0, 0.1, (n, 3))
    F1, W = rng.normal(0, 0.02, (n, 3, 3)), rng.normal(0, 0.01, (n, 3, 3))
    F1[~valid] = np.nan  # the nodal gradient's output at invalid nodes: never read
    F = FEM.nodal_gradient(ops, U1)
    Fr = FDR.nodal_gradient(ref, U1)
    assert np.array_equal(np.isnan(F), np.isnan(Fr))
    _close(F[valid], Fr[valid])
    beta, mu, alpha = 0.5, 2e-3, 0.05
    b = FEM.global_rhs(ops, beta, mu, alpha, U1, F1, W, v)
    assert np.all(np.isfinite(b))
    _close(b, FDR.global_rhs(ref, beta, mu, alpha, U1, np.nan_to_num(F1), W, v))
    para = SimpleNamespace(global_solver="pcg", pcg_tol=1e-10, pcg_max_iter=2000)
    U2, info = FEM.solve_subpb2(FEM.build_global_system(ops, beta, mu, alpha, para), U1, F1, W, v)
    U2r, it = FDR.solve_subpb2(ref, beta, mu, alpha, U1, np.nan_to_num(F1), W, v, tol=1e-10, max_iter=2000)
    assert abs(info["iterations"] - it) <= 1
    assert np.max(np.abs(U2 - U2r)) < 1e-7
