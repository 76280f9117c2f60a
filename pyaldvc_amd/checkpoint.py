"""Per-frame checkpoints, readable and writable in the reference's on-disk format
(al_dvc 1.2.1, core/checkpoint.py:35-120).

On disk (the contract shared with the reference, so either side can resume the other's run):

* ``meta.json``: ``format`` 1, ``para`` (every parameter as JSON, ``verbose`` and ``n_threads``
  left out; a VOI range as ``{"x", "y", "z"}`` lists, a frame schedule as ``{"ref_indices"}``),
  ``volume_shape``, ``n_frames``, ``ref_indices``, ``grid_shape`` and the lattice axes ``x0``,
  ``y0``, ``z0``;
* ``frame_NNNN.npz`` per solved pair: the result arrays under their field names, the mesh state
  (``node_valid``, ``elements``, ``boundary_nodes``, ``edge_ok``), ``ref_frame``, and the ADMM record
  flattened into ``admm_scalars`` = [beta, mu, n_steps, subpb2_time], one array per residual series,
  ``admm_local_<i>_{n_iter,status,zncc,scalars}`` per local pass (scalars = [solve_time, n_bad]) and
  ``admm_sweep_<key>`` per entry of the beta sweep.

Here the format is a table of fields and an encoder/decoder pair around it (:func:`encode_frame`,
:func:`decode_frame`); :class:`Checkpoint` only moves files. Frames are written to a partial file
and renamed into place; files are read with ``allow_pickle=False``. Resuming compares the stored
``meta.json`` with this run's (floats to 12 decimals) and refuses a different run.
"""

from __future__ import annotations

import dataclasses
import json
import os
import re
import tempfile
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

META_NAME = "meta.json"
FORMAT_VERSION = 1
LEFT_OUT_PARAMS = ("verbose", "n_threads")  # per-machine settings, not part of a run's identity
FRAME_NAME = re.compile(r"^frame_(\d{4,})\.npz$")
LOCAL_PASS = re.compile(r"^admm_local_(\d+)_n_iter$")
SWEEP = "admm_sweep_"

# (field, dtype) of the per-node arrays of a frame; the first two are always present
RESULT_FIELDS = (("U", np.float64), ("F", np.float64), ("U_accum", np.float64), ("U_local", np.float64),
                 ("F_local", np.float64), ("U0", np.float64), ("zncc", np.float64), ("status", np.int8),
                 ("U_std", np.float64), ("split_fraction", np.float32), ("outlier", np.bool_))  # fmt: skip
MESH_FIELDS = (("node_valid", np.bool_), ("elements", np.int64), ("boundary_nodes", np.int64), ("edge_ok", np.bool_))
RESIDUALS = ("update_global", "update_local", "primal_residual_u", "primal_residual_f")


class CheckpointMismatch(ValueError):
    """The checkpoint directory belongs to a different run."""


@dataclass
class FrameResult:
    """A solved frame pair (the reference's data_structures.py:386-423 fields)."""

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


# ---------------------------------------------------------------------------- meta.json
def _plain(value):
    """A parameter value as JSON: VOI-like objects (x, y, z) and schedules (ref_indices) as dicts,
    tuples as lists, NumPy scalars as Python numbers."""
    if isinstance(value, dict):
        return {k: _plain(v) for k, v in value.items()}
    if all(hasattr(value, a) for a in ("x", "y", "z")) and not isinstance(value, (str, bytes)):
        return {a: _plain(getattr(value, a)) for a in ("x", "y", "z")}
    if hasattr(value, "ref_indices"):
        return {"ref_indices": [int(i) for i in value.ref_indices]}
    if isinstance(value, (tuple, list)):
        return [_plain(v) for v in value]
    if isinstance(value, np.generic):
        return value.item()
    return value


def para_dict(para) -> dict:
    """The parameter set as the JSON dictionary ``meta.json`` stores: a mapping, a dataclass
    (``DVCPara``) or any object with attributes (a namespace)."""
    if isinstance(para, dict):
        items = dict(para)
    elif dataclasses.is_dataclass(para):
        items = {f.name: getattr(para, f.name) for f in dataclasses.fields(para)}
    else:
        items = dict(vars(para))
    return {k: _plain(v) for k, v in items.items() if k not in LEFT_OUT_PARAMS}


def lattice_axes(mesh) -> tuple:
    """``(x0, y0, z0)``: the mesh's own axes when it carries them, else read off its full node grid."""
    if all(hasattr(mesh, a) for a in ("x0", "y0", "z0")):
        return tuple(np.asarray(getattr(mesh, a), dtype=np.float64) for a in ("x0", "y0", "z0"))
    nz, ny, nx = (int(v) for v in mesh.grid_shape)
    grid = np.asarray(mesh.coordinates, dtype=np.float64).reshape(nz, ny, nx, 3)
    return grid[0, 0, :, 0], grid[0, :, 0, 1], grid[:, 0, 0, 2]


def run_meta(para, volume_shape, schedule, mesh, n_frames: int) -> dict:
    """The ``meta.json`` of a run; ``schedule`` is a frame schedule or its list of reference indices."""
    refs = schedule.ref_indices if hasattr(schedule, "ref_indices") else schedule
    axes = lattice_axes(mesh)
    meta = {"format": FORMAT_VERSION, "para": para_dict(para), "volume_shape": [int(s) for s in volume_shape],
            "n_frames": int(n_frames), "ref_indices": [int(r) for r in refs],
            "grid_shape": [int(s) for s in mesh.grid_shape]}  # fmt: skip
    meta.update({name: [float(v) for v in ax] for name, ax in zip(("x0", "y0", "z0"), axes)})
    return meta


meta_for = run_meta


def _canonical(value):
    if isinstance(value, float):
        return round(value, 12)
    if isinstance(value, (list, tuple)):
        return [_canonical(v) for v in value]
    if isinstance(value, dict):
        return {k: _canonical(v) for k, v in value.items()}
    return value


def meta_differences(stored: dict, current: dict) -> list:
    """The entries in which two ``meta.json`` contents differ (``para.<name>`` for parameters)."""
    a, b = _canonical(json.loads(json.dumps(stored))), _canonical(json.loads(json.dumps(current)))
    pa, pb = a.pop("para", {}), b.pop("para", {})
    out = [f"para.{k}" for k in sorted(pa.keys() | pb.keys()) if pa.get(k) != pb.get(k)]
    return out + [k for k in sorted(a.keys() | b.keys()) if a.get(k) != b.get(k)]


# ---------------------------------------------------------------------------- frame files
def encode_frame(result: FrameResult, mesh) -> dict:
    """The arrays of one ``frame_NNNN.npz``."""
    out: dict = {"ref_frame": np.int64(result.ref_frame)}
    for name, dtype in RESULT_FIELDS:
        value = getattr(result, name)
        if value is not None:
            out[name] = np.asarray(value, dtype=dtype)
    for name, dtype in MESH_FIELDS:
        value = getattr(mesh, name, None)
        out[name] = np.asarray(value if value is not None else np.zeros((0, 3) if name == "edge_ok" else 0), dtype=dtype)
    rec = result.admm
    if rec is not None:
        out["admm_scalars"] = np.array([rec.beta, rec.mu, rec.n_steps, rec.subpb2_time], dtype=np.float64)
        out.update({f"admm_{name}": np.asarray(getattr(rec, name), dtype=np.float64) for name in RESIDUALS})
        for i, loc in enumerate(rec.local_info):
            out.update({f"admm_local_{i}_n_iter": np.asarray(loc.n_iter, dtype=np.int32),
                        f"admm_local_{i}_status": np.asarray(loc.status, dtype=np.int8),
                        f"admm_local_{i}_zncc": np.asarray(loc.zncc, dtype=np.float64),
                        f"admm_local_{i}_scalars": np.array([loc.solve_time, loc.n_bad], dtype=np.float64)})  # fmt: skip
        for key, value in (rec.beta_sweep or {}).items():
            out[SWEEP + key] = np.asarray(value)
    return out


def decode_frame(arrays: dict) -> tuple:
    """``(FrameResult, mesh fields)`` from the arrays of one frame file."""
    from .admm import ADMMInfo
    from .local import LocalSolveInfo

    rec = None
    if "admm_scalars" in arrays:
        beta, mu, n_steps, t2 = (float(v) for v in arrays["admm_scalars"])
        passes = sorted(int(m.group(1)) for m in map(LOCAL_PASS.match, arrays) if m)
        local = []
        for i in passes:
            solve_time, n_bad = (float(v) for v in arrays[f"admm_local_{i}_scalars"])
            local.append(LocalSolveInfo(n_iter=arrays[f"admm_local_{i}_n_iter"], status=arrays[f"admm_local_{i}_status"],
                                        zncc=arrays[f"admm_local_{i}_zncc"], solve_time=solve_time, n_bad=int(n_bad)))  # fmt: skip
        sweep = {k[len(SWEEP):]: (v.item() if np.ndim(v) == 0 else v) for k, v in arrays.items() if k.startswith(SWEEP)}
        series = {name: tuple(float(v) for v in arrays[f"admm_{name}"]) for name in RESIDUALS}
        rec = ADMMInfo(beta=beta, mu=mu, n_steps=int(n_steps), local_info=tuple(local), subpb2_time=t2,
                       beta_sweep=sweep or None, **series)  # fmt: skip
    fields = {name: arrays.get(name) for name, _ in RESULT_FIELDS}
    if fields["outlier"] is not None:
        fields["outlier"] = fields["outlier"].astype(bool)
    result = FrameResult(ref_frame=int(arrays["ref_frame"]), admm=rec, **fields)
    mesh = {name: arrays[name].astype(dtype) for name, dtype in MESH_FIELDS if name in arrays}
    mesh.setdefault("edge_ok", np.zeros((0, 3), dtype=bool))
    return result, mesh


class Checkpoint:
    """A checkpoint directory: ``meta.json`` plus one ``frame_NNNN.npz`` per solved pair."""

    def __init__(self, directory):
        self.dir = Path(directory)

    def frame_path(self, k: int) -> Path:
        return self.dir / f"frame_{int(k):04d}.npz"

    def prepare(self, meta: dict, resume: bool) -> None:
        """Start (or, with ``resume``, continue) the run described by ``meta``: a stored ``meta.json``
        of another run is refused on resume; a fresh start clears the frames and writes ``meta``."""
        self.dir.mkdir(parents=True, exist_ok=True)
        stored = self.dir / META_NAME
        if resume and stored.is_file():
            diff = meta_differences(json.loads(stored.read_text(encoding="utf-8")), meta)
            if diff:
                raise CheckpointMismatch(f"{self.dir} holds frames of another run (differs in {', '.join(diff)}); "
                                         "start afresh (resume=False) or choose another directory")  # fmt: skip
            return
        for k in self.completed_frames():
            self.frame_path(k).unlink()
        stored.write_text(json.dumps(meta, indent=1), encoding="utf-8")

    def prepare_run(self, para, volume_shape, schedule, mesh, n_frames: int, resume: bool) -> None:
        """:meth:`prepare` with the ``meta.json`` of these run parameters."""
        self.prepare(run_meta(para, volume_shape, schedule, mesh, n_frames), resume)

    def has(self, k: int) -> bool:
        return self.frame_path(k).is_file()

    def completed_frames(self) -> list:
        if not self.dir.is_dir():
            return []
        return sorted(int(m.group(1)) for m in (FRAME_NAME.match(p.name) for p in self.dir.iterdir()) if m)

    def save(self, k: int, result: FrameResult, mesh) -> Path:
        """Write frame ``k`` (to a partial file first, then renamed into place)."""
        path = self.frame_path(k)
        fd, tmp = tempfile.mkstemp(prefix=path.name, suffix=".partial", dir=self.dir)
        try:
            with os.fdopen(fd, "wb") as fh:
                np.savez_compressed(fh, **encode_frame(result, mesh))
            os.replace(tmp, path)
        finally:
            if os.path.exists(tmp):
                os.unlink(tmp)
        return path

    def load(self, k: int) -> tuple:
        """``(FrameResult, mesh fields)`` of frame ``k``."""
        with np.load(self.frame_path(k), allow_pickle=False) as z:
            return decode_frame({name: z[name] for name in z.files})
