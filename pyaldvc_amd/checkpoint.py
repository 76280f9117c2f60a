"""Per-frame checkpoints in the reference's format (core/checkpoint.py:35-205, al_dvc 1.2.1).

A directory holds ``meta.json`` (parameters, volume shape, frame schedule, node grid) and one
``frame_<k>.npz`` per solved frame pair with the arrays of a ``FrameResult`` and the node validity
of the mesh, under the reference's key names, so that a run of this path can be resumed by the
reference and the other way round. Files are written to a temporary name and renamed, and read
with ``allow_pickle=False``. A directory written by a different run (parameters, volume shape,
schedule or node grid) is refused with ``CheckpointMismatch``.
"""

from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

META_NAME = "meta.json"
FORMAT_VERSION = 1  # checkpoint.py:25


class CheckpointMismatch(ValueError):
    """The checkpoint directory belongs to a different run."""


@dataclass
class FrameResult:
    """The reference's per-pair result (data_structures.py:386-423)."""

    U: np.ndarray
    F: np.ndarray
    ref_frame: int = 0
    U_accum: np.ndarray | None = None
    U_local: np.ndarray | None = None
    F_local: np.ndarray | None = None
    U0: np.ndarray | None = None
    zncc: np.ndarray | None = None
    U_std: np.ndarray | None = None
    status: np.ndarray | None = None
    admm: Any = None
    split_fraction: np.ndarray | None = None
    outlier: np.ndarray | None = None


OPTIONAL = ("U_accum", "U_local", "F_local", "U0", "zncc", "status", "U_std", "split_fraction", "outlier")
ADMM_SERIES = ("update_global", "update_local", "primal_residual_u", "primal_residual_f")


def grid_axes(mesh) -> tuple:
    """``(x0, y0, z0)`` lattice axes of a full node grid ``(nz, ny, nx)``."""
    nz, ny, nx = (int(v) for v in mesh.grid_shape)
    c = np.asarray(mesh.coordinates, dtype=np.float64).reshape(nz, ny, nx, 3)
    return c[0, 0, :, 0], c[0, :, 0, 1], c[:, 0, 0, 2]


def meta_for(para: dict, volume_shape, ref_indices, mesh, n_frames: int) -> dict:
    """The ``meta.json`` content (checkpoint.py:35-50); ``para`` is the parameter dictionary."""
    d = dict(para)
    d.pop("verbose", None)
    d.pop("n_threads", None)
    x0, y0, z0 = grid_axes(mesh)
    return {"format": FORMAT_VERSION, "para": d, "volume_shape": [int(s) for s in volume_shape],
            "n_frames": int(n_frames), "ref_indices": [int(r) for r in ref_indices],
            "grid_shape": [int(s) for s in mesh.grid_shape], "x0": [float(v) for v in x0],
            "y0": [float(v) for v in y0], "z0": [float(v) for v in z0]}  # fmt: skip


def _norm(value):
    """Lists for tuples, floats to 12 decimals (checkpoint.py:199-207), so that a JSON round trip
    compares equal."""
    if isinstance(value, (list, tuple)):
        return [_norm(v) for v in value]
    if isinstance(value, dict):
        return {k: _norm(v) for k, v in value.items()}
    if isinstance(value, float):
        return round(value, 12)
    return value


def _meta_diff(a: dict, b: dict) -> list:
    """Names of the entries that differ, parameters key by key (checkpoint.py:186-205)."""
    diff = []
    for key in sorted(set(a) | set(b)):
        if key == "para":
            pa, pb = a.get("para", {}), b.get("para", {})
            diff += [f"para.{k}" for k in sorted(set(pa) | set(pb)) if _norm(pa.get(k)) != _norm(pb.get(k))]
        elif _norm(a.get(key)) != _norm(b.get(key)):
            diff.append(key)
    return diff


class Checkpoint:
    """Reader/writer of a checkpoint directory (checkpoint.py:53-183)."""

    def __init__(self, directory):
        self.dir = Path(directory)

    def frame_path(self, k: int) -> Path:
        return self.dir / f"frame_{k:04d}.npz"

    def prepare(self, meta: dict, resume: bool) -> None:
        """Create the directory and ``meta.json``; with ``resume`` an existing meta must match."""
        self.dir.mkdir(parents=True, exist_ok=True)
        path = self.dir / META_NAME
        if path.exists() and resume:
            diff = _meta_diff(json.loads(path.read_text(encoding="utf-8")), meta)
            if diff:
                raise CheckpointMismatch(f"checkpoint directory {self.dir} was written by a different run "
                                         f"({', '.join(diff)}); use resume=False (or another directory) to start over")  # fmt: skip
            return
        for p in self.dir.glob("frame_*.npz"):
            p.unlink()
        path.write_text(json.dumps(meta, indent=1), encoding="utf-8")

    def has(self, k: int) -> bool:
        return self.frame_path(k).is_file()

    def completed_frames(self) -> list:
        return sorted(int(p.stem.split("_")[1]) for p in self.dir.glob("frame_*.npz"))

    def save(self, k: int, fr: FrameResult, mesh) -> Path:
        """``frame_<k>.npz`` with the result and the mesh's validity, elements, boundary and cut edges."""
        arrays: dict = {
            "U": np.asarray(fr.U), "F": np.asarray(fr.F), "ref_frame": np.int64(fr.ref_frame),
            "node_valid": np.asarray(mesh.node_valid, dtype=bool),
            "elements": np.asarray(getattr(mesh, "elements", np.zeros((0, 8))), dtype=np.int64),
            "boundary_nodes": np.asarray(getattr(mesh, "boundary_nodes", np.zeros(0)), dtype=np.int64),
            "edge_ok": np.asarray(mesh.edge_ok if getattr(mesh, "edge_ok", None) is not None else np.zeros((0, 3)),
                                  dtype=bool),
        }  # fmt: skip
        for name in OPTIONAL:
            val = getattr(fr, name)
            if val is not None:
                arrays[name] = np.asarray(val)
        a = fr.admm
        if a is not None:
            arrays["admm_scalars"] = np.array([a.beta, a.mu, a.n_steps, a.subpb2_time], dtype=np.float64)
            for name in ADMM_SERIES:
                arrays["admm_" + name] = np.asarray(getattr(a, name), dtype=np.float64)
            for i, li in enumerate(a.local_info):
                arrays[f"admm_local_{i}_n_iter"] = np.asarray(li.n_iter)
                arrays[f"admm_local_{i}_status"] = np.asarray(li.status)
                arrays[f"admm_local_{i}_zncc"] = np.asarray(li.zncc)
                arrays[f"admm_local_{i}_scalars"] = np.array([li.solve_time, li.n_bad], dtype=np.float64)
            if getattr(a, "beta_sweep", None) is not None:
                for key, val in a.beta_sweep.items():
                    arrays["admm_sweep_" + key] = np.asarray(val)
        path = self.frame_path(k)
        tmp = path.with_suffix(".tmp.npz")
        np.savez_compressed(tmp, **arrays)
        tmp.replace(path)
        return path

    def load(self, k: int):
        """``(FrameResult, mesh arrays dict)`` of frame ``k``; the ADMM diagnostics come back as the
        device loop's ``ADMMInfo`` with ``LocalSolveInfo`` entries."""
        from .admm import ADMMInfo
        from .local import LocalSolveInfo

        with np.load(self.frame_path(k), allow_pickle=False) as z:
            d = {key: z[key] for key in z.files}
        mesh = {"node_valid": d["node_valid"].astype(bool), "elements": d["elements"].astype(np.int64),
                "boundary_nodes": d["boundary_nodes"].astype(np.int64),
                "edge_ok": d["edge_ok"].astype(bool) if "edge_ok" in d else np.zeros((0, 3), dtype=bool)}  # fmt: skip
        admm = None
        if "admm_scalars" in d:
            beta, mu, n_steps, subpb2_time = (float(v) for v in d["admm_scalars"])
            local_info, i = [], 0
            while f"admm_local_{i}_n_iter" in d:
                st, nb = d[f"admm_local_{i}_scalars"]
                local_info.append(LocalSolveInfo(n_iter=d[f"admm_local_{i}_n_iter"], status=d[f"admm_local_{i}_status"],
                                                 zncc=d[f"admm_local_{i}_zncc"], solve_time=float(st), n_bad=int(nb)))  # fmt: skip
                i += 1
            sweep = {key[len("admm_sweep_"):]: d[key] for key in d if key.startswith("admm_sweep_")} or None
            if sweep is not None:
                if "k_best" in sweep:
                    sweep["k_best"] = int(sweep["k_best"])
                for key in ("beta", "criterion"):
                    if key in sweep and np.ndim(sweep[key]) == 0:
                        sweep[key] = sweep[key].item()
            admm = ADMMInfo(beta=beta, mu=mu, n_steps=int(n_steps),
                            **{name: tuple(float(v) for v in d["admm_" + name]) for name in ADMM_SERIES},
                            local_info=tuple(local_info), subpb2_time=subpb2_time, beta_sweep=sweep)  # fmt: skip
        fr = FrameResult(U=d["U"], F=d["F"], ref_frame=int(d["ref_frame"]), admm=admm,
                         **{name: d.get(name) for name in OPTIONAL if name != "outlier"},
                         outlier=None if d.get("outlier") is None else np.asarray(d["outlier"], dtype=bool))  # fmt: skip
        return fr, mesh
