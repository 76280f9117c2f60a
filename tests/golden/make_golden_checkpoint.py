"""Write a checkpoint directory with the reference's own writer, for the format interop test.

Run once, in the build container only (the reference is not available anywhere else):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden_checkpoint.py

It imports ``al_dvc`` 1.2.1 from ``/root/reference/src`` and calls ``Checkpoint.prepare`` and
``Checkpoint.save`` (``core/checkpoint.py:52-120``) on a 3 x 2 x 4 node grid with a dropped element,
a cut edge, a parameter set with a VOI range and a frame schedule, and a frame result carrying every
optional field and an ADMM record with two local passes and a beta sweep. The output goes to
``tests/golden/ckpt_ref/`` (``meta.json``, ``frame_0001.npz``); the inputs are seeded NumPy arrays.
Nothing here is imported by the tests.
"""

from __future__ import annotations

import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, "/root/reference/src")
sys.dont_write_bytecode = True

from al_dvc.core.checkpoint import Checkpoint  # noqa: E402
from al_dvc.core.config import DVCPara  # noqa: E402
from al_dvc.core.data_structures import ADMMInfo, FrameResult, FrameSchedule, LocalSolveInfo, VOIRange  # noqa: E402
from al_dvc.mesh.grid_mesh import mesh_setup  # noqa: E402

OUT = os.path.join(HERE, "ckpt_ref")


def main() -> None:
    rng = np.random.default_rng(20261204)
    mesh = mesh_setup(np.array([8.0, 16.0, 24.0, 32.0]), np.array([8.0, 20.0]), np.array([6.0, 14.0, 22.0]))
    n = mesh.n_nodes
    elements = np.array(mesh.elements, copy=True)
    elements[2] = -1
    edge_ok = np.ones((n, 3), dtype=bool)
    edge_ok[5, 0] = False
    mesh = type(mesh)(**{**mesh.__dict__, "elements": elements, "edge_ok": edge_ok,
                         "node_valid": np.arange(n) % 7 != 3, "boundary_nodes": np.array([0, 1, 2, 9], np.int64)})  # fmt: skip
    para = DVCPara(winsize=(16, 16, 12), winstepsize=(8, 12, 8), voi=VOIRange(x=(2, 40), y=(0, -1), z=(1, 30)),
                   frame_schedule=FrameSchedule(ref_indices=(0, 0)), mu=2e-3, verbose=True)  # fmt: skip
    info = [LocalSolveInfo(n_iter=rng.integers(0, 9, n).astype(np.int32), status=rng.integers(0, 8, n).astype(np.int8),
                           zncc=rng.random(n), solve_time=0.125 * (i + 1), n_bad=i + 2) for i in range(2)]  # fmt: skip
    admm = ADMMInfo(beta=1.5e-3, mu=2e-3, n_steps=4, update_global=(0.1, 0.02, 0.003), update_local=(0.2, 0.03, 0.004),
                    primal_residual_u=(0.3, 0.04), primal_residual_f=(0.5, 0.06), local_info=tuple(info),
                    subpb2_time=0.75, beta_sweep={"betas": np.array([1e-4, 1e-3, 1e-2]), "err1": rng.random(3),
                                                  "err2": rng.random(3), "score": rng.random(3), "k_best": 1})  # fmt: skip
    fr = FrameResult(U=rng.normal(size=(n, 3)), F=rng.normal(size=(n, 3, 3)), ref_frame=0,
                     U_accum=rng.normal(size=(n, 3)), U_local=rng.normal(size=(n, 3)), F_local=rng.normal(size=(n, 3, 3)),
                     U0=rng.normal(size=(n, 3)), zncc=rng.random(n), U_std=rng.random((n, 3)),
                     status=rng.integers(0, 8, n).astype(np.int8), admm=admm, split_fraction=rng.random(n).astype(np.float32),
                     outlier=rng.random(n) > 0.7)  # fmt: skip
    if os.path.isdir(OUT):
        shutil.rmtree(OUT)
    ck = Checkpoint(OUT)
    ck.prepare(para, (44, 36, 48), para.frame_schedule, mesh, 3, resume=False)
    ck.save(1, fr, mesh)
    print("wrote", sorted(os.listdir(OUT)))


if __name__ == "__main__":
    main()
