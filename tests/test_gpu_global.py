"""ADMM global step on the GPU against the reference (needs an MI355X).

``global_rhs``, ``solve_subpb2`` (direct and PCG) and ``nodal_gradient`` of al_dvc 1.2.1 on a
7x6x5 node grid with dropped elements (``golden_driver.npz``); closed forms: the lumped gradient of
a trilinear-exact (linear) field is its gradient at every active node, and the PCG solution
satisfies the system to its tolerance.
"""

from __future__ import annotations

import os
from types import SimpleNamespace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FEM = pytest.importorskip("pyaldvc_amd.fem")
K = pytest.importorskip("pyaldvc_amd.kernels")
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def gd():
    return np.load(os.path.join(HERE, "golden", "golden_driver.npz"))


@pytest.fixture(scope="module")
def ops(gd):
    mesh = SimpleNamespace(grid_shape=tuple(gd["gs_grid_shape"]), elements=gd["gs_elements"], spacing=tuple(gd["gs_spacing"]))
    return FEM.build_global_operators(mesh, "fem", 2)


def _args(gd):
    return gd["gs_U1"], gd["gs_F1"], gd["gs_W"], gd["gs_v"]


def test_operators_and_rhs(gd, ops):
    assert np.array_equal(ops.active, gd["gs_active"])
    for beta, mu, alpha, tag in ((1e-3, 1e-3, 0.0, "a"), (0.5, 2e-3, 0.05, "b")):
        b = FEM.global_rhs(ops, beta, mu, alpha, *_args(gd))
        ref = gd[f"gs_rhs_{tag}"]
        assert np.max(np.abs(b - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("tag,beta,mu,alpha", [("a", 1e-3, 1e-3, 0.0), ("b", 0.5, 2e-3, 0.05)])
@pytest.mark.parametrize("solver", ["pcg", "direct"])
def test_solve_subpb2(gd, ops, tag, beta, mu, alpha, solver):
    para = SimpleNamespace(global_solver=solver, pcg_tol=1e-8, pcg_max_iter=500)
    system = FEM.build_global_system(ops, beta, mu, alpha, para)
    U2, info = FEM.solve_subpb2(system, *_args(gd))
    ref = gd[f"gs_U2_{tag}_{solver}"]
    inactive = ~ops.active
    assert np.array_equal(U2[inactive], gd["gs_U1"][inactive])
    if solver == "pcg":  # the same iteration: same count, same iterate to round-off
        assert info["iterations"] == int(gd[f"gs_iter_{tag}_pcg"])
        assert np.max(np.abs(U2 - ref)) < 1e-9
    else:  # the reference's sparse LU, here CG to round-off
        assert np.max(np.abs(U2 - ref)) < 1e-9


def test_pcg_solves_the_system(gd, ops):
    import torch

    para = SimpleNamespace(global_solver="pcg", pcg_tol=1e-10, pcg_max_iter=1000)
    system = FEM.build_global_system(ops, 0.3, 1e-3, 0.0, para)
    rng = np.random.default_rng(0)
    B = rng.standard_normal((ops.n_nodes, 3))
    B[~ops.active] = 0.0
    Bd = torch.from_numpy(B).to(K.device())
    X, it = FEM.pcg_fem(system, Bd, torch.zeros_like(Bd), 1e-10, 1000)
    AX = torch.empty_like(X)
    nz, ny, nx = ops.grid_shape
    K.check(K.lib().aldvc_fem_apply(K._ptr(system.Ae), nx, ny, nz, K._ptr(ops.elem_active), K._ptr(ops.node_active),
                                         K._ptr(X), K._ptr(AX), K.stream_handle()), "apply")  # fmt: skip
    r = (AX - Bd).cpu().numpy()
    assert np.all(np.linalg.norm(r, axis=0) / np.linalg.norm(B, axis=0) < 1e-10)
    assert 0 < it < 1000


def test_nodal_gradient(gd, ops):
    F = FEM.nodal_gradient(ops, gd["gs_U1"])
    ref = gd["gs_grad"]
    assert np.array_equal(np.isnan(F), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.max(np.abs(F[ok] - ref[ok])) <= 1e-12 * np.max(np.abs(ref[ok]))
    # a linear field: the lumped L2 gradient is exact at every active node
    A = np.array([[0.01, -0.02, 0.03], [0.005, 0.0, -0.01], [0.02, 0.015, -0.005]])
    U = gd["gs_coords"] @ A.T + np.array([0.3, -0.1, 0.2])
    F = FEM.nodal_gradient(ops, U)
    assert np.max(np.abs(F[ops.active] - A)) < 1e-12
