"""Write golden vectors of the reference's finite-difference global step.

Run once, in the build container only (the reference is not available anywhere else):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_fd.py

It imports ``al_dvc`` 1.2.1 from ``/root/reference/src`` and records, on three node grids,
``build_fd_operators`` (``solver/global_operators.py:74-137``: the difference operators ``D_j``,
``Kg``, ``M``, ``G_j``, ``active``), ``nodal_gradient`` (``:151-170``, the FD branch), ``global_rhs``
and ``solve_subpb2`` with the direct and the PCG solver (``solver/subpb2_solver.py:45-155``):

* ``m``: a 7 x 6 x 5 grid (spacing 5, 6, 7) with a masked corner column and a masked slab, so some
  nodes are invalid (``apply_mask_to_mesh``, ``mesh/grid_mesh.py:191-253``);
* ``c``: the same grid with a thin masked slit cut into it (``cut_bridging=True``), so some +x edges
  between valid nodes are cut (``edge_ok``) and the operators must take one-sided differences there;
* ``u``: an unmasked 9 x 4 x 6 grid (spacing 4, 3, 5).

The operators are recorded by their action on seeded vectors (``D_j X``, ``Kg X``, ``M X``, ``G_j X``
for a random ``X`` of 3 columns). Inputs are seeded NumPy arrays stored next to the outputs in
``tests/golden/golden_fd.npz``. Nothing here is imported by the tests.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, "/root/reference/src")
sys.dont_write_bytecode = True

from al_dvc.core.config import DVCPara  # noqa: E402
from al_dvc.mesh import grid_mesh as rgm  # noqa: E402
from al_dvc.solver import global_operators as rgo  # noqa: E402
from al_dvc.solver import subpb2_solver as rsp  # noqa: E402

CASES = (("a", 1e-3, 1e-3, 0.0), ("b", 0.5, 2e-3, 0.05))


def meshes():
    x0 = 10.0 + 5.0 * np.arange(7)
    y0 = 12.0 + 6.0 * np.arange(6)
    z0 = 9.0 + 7.0 * np.arange(5)
    base = rgm.mesh_setup(x0, y0, z0)
    mask = np.ones((60, 60, 60), dtype=np.uint8)
    mask[40:, 40:, :] = 0  # a corner column of nodes
    mask[:14, :, 30:36] = 0
    yield "m", rgm.apply_mask_to_mesh(base, mask, (10, 10, 10), 0.5)
    slit = np.ones((60, 60, 60), dtype=np.uint8)
    slit[:, 20:40, 22:24] = 0  # between the x = 20 and x = 25 node planes, over part of y
    slit[40:, 40:, :] = 0
    yield "c", rgm.apply_mask_to_mesh(base, slit, (10, 10, 10), 0.5, cut_bridging=True)
    yield "u", rgm.mesh_setup(12.0 + 4.0 * np.arange(9), 8.0 + 3.0 * np.arange(4), 10.0 + 5.0 * np.arange(6))


def main() -> None:
    rng = np.random.default_rng(29)
    out = {}
    for key, mesh in meshes():
        n = mesh.n_nodes
        ops = rgo.build_fd_operators(mesh)
        out[f"{key}_grid_shape"] = np.array(mesh.grid_shape)
        out[f"{key}_spacing"] = np.array(mesh.spacing)
        out[f"{key}_coords"] = mesh.coordinates
        out[f"{key}_node_valid"] = np.asarray(mesh.node_valid, dtype=bool)
        out[f"{key}_edge_ok"] = mesh.edges_ok() if mesh.is_cut else np.ones((n, 3), dtype=bool)
        out[f"{key}_is_cut"] = np.array(bool(mesh.is_cut))
        out[f"{key}_active"] = ops.active
        X = rng.normal(0.0, 1.0, (n, 3))
        out[f"{key}_X"] = X
        out[f"{key}_DX"] = np.stack([ops.D[j] @ X for j in range(3)])
        out[f"{key}_KX"] = ops.Kg @ X
        out[f"{key}_MX"] = ops.M @ X
        out[f"{key}_GX"] = np.stack([ops.G[j] @ X for j in range(3)])
        U1 = rng.normal(0, 1.5, (n, 3))
        F1 = rng.normal(0, 0.02, (n, 3, 3))
        W = rng.normal(0, 0.01, (n, 3, 3))
        v = rng.normal(0, 0.1, (n, 3))
        out.update({f"{key}_U1": U1, f"{key}_F1": F1, f"{key}_W": W, f"{key}_v": v})
        out[f"{key}_grad"] = rgo.nodal_gradient(ops, U1)
        for tag, beta, mu, alpha in CASES:
            out[f"{key}_rhs_{tag}"] = rsp.global_rhs(ops, beta, mu, alpha, U1, F1, W, v)
            for solver in ("direct", "pcg"):
                para = DVCPara(global_solver=solver, pcg_tol=1e-8, pcg_max_iter=500)
                system = rsp.build_global_system(ops, beta, mu, alpha, para)
                U2, info = rsp.solve_subpb2(system, U1, F1, W, v)
                out[f"{key}_U2_{tag}_{solver}"] = U2
                out[f"{key}_iter_{tag}_{solver}"] = np.array(info["iterations"])
        cut = int((~out[f"{key}_edge_ok"]).sum())
        print(f"{key}: {n} nodes, {int(ops.active.sum())} active, {cut} cut edges, pcg iterations "
              f"{int(out[f'{key}_iter_a_pcg'])}/{int(out[f'{key}_iter_b_pcg'])}")  # fmt: skip
    path = os.path.join(HERE, "golden_fd.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


if __name__ == "__main__":
    main()
