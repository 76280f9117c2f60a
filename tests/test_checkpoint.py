"""Per-frame checkpoints in the reference's format (pyaldvc_amd/checkpoint.py, core/checkpoint.py)."""

from __future__ import annotations

import json
from types import SimpleNamespace

import numpy as np
import pytest

from pyaldvc_amd import checkpoint as C
from pyaldvc_amd.admm import ADMMInfo
from pyaldvc_amd.local import LocalSolveInfo
from pyaldvc_amd.mesh_cut import MaskedMesh, hex8_elements


def _mesh():
    ax = np.array([8.0, 16.0, 24.0])
    Z, Y, X = np.meshgrid(ax, ax[:2], ax, indexing="ij")
    coords = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])
    grid = (3, 2, 3)
    el = hex8_elements(grid)
    el[1] = -1
    edge = np.ones((coords.shape[0], 3), dtype=bool)
    edge[4, 0] = False
    return MaskedMesh(coords, grid, np.arange(coords.shape[0]) % 5 != 0, edge, el, np.array([0, 1, 2], np.int64))


def _result(n, rng, admm=True):
    info = [LocalSolveInfo(n_iter=rng.integers(0, 9, n).astype(np.int32), status=rng.integers(0, 3, n).astype(np.int8),
                           zncc=rng.random(n), solve_time=0.25, n_bad=3) for _ in range(2)]  # fmt: skip
    a = ADMMInfo(beta=1e-3, mu=1e-3, n_steps=3, update_global=(0.1, 0.01), update_local=(0.2, 0.02),
                 primal_residual_u=(0.3, 0.03), primal_residual_f=(0.4, 0.04), local_info=tuple(info), subpb2_time=0.5,
                 beta_sweep={"beta": np.array([1e-4, 1e-3]), "k_best": np.int64(1), "criterion": np.array([2.0, 1.0])})  # fmt: skip
    return C.FrameResult(U=rng.normal(size=(n, 3)), F=rng.normal(size=(n, 3, 3)), ref_frame=2, U0=rng.normal(size=(n, 3)),
                         zncc=rng.random(n), status=rng.integers(0, 3, n).astype(np.int8), U_std=rng.random((n, 3)),
                         outlier=rng.random(n) > 0.8, admm=a if admm else None)  # fmt: skip


def test_frame_round_trip_with_the_reference_keys(tmp_path):
    mesh = _mesh()
    fr = _result(mesh.n_nodes, np.random.default_rng(0))
    ck = C.Checkpoint(tmp_path)
    ck.prepare(C.meta_for({"winsize": [16, 16, 16], "mu": 1e-3, "verbose": True}, (40, 32, 40), [0, 0], mesh, 3), resume=False)
    path = ck.save(1, fr, mesh)
    assert path.name == "frame_0001.npz" and ck.has(1) and ck.completed_frames() == [1]
    with np.load(path, allow_pickle=False) as z:
        keys = set(z.files)
    # the names core/checkpoint.py:88-117 writes
    assert {"U", "F", "ref_frame", "node_valid", "elements", "boundary_nodes", "edge_ok", "U0", "zncc", "status",
            "U_std", "outlier", "admm_scalars", "admm_update_global", "admm_update_local", "admm_primal_residual_u",
            "admm_primal_residual_f", "admm_local_0_n_iter", "admm_local_1_scalars", "admm_sweep_beta",
            "admm_sweep_k_best"} <= keys  # fmt: skip
    got, m = ck.load(1)
    for name in ("U", "F", "U0", "zncc", "status", "U_std", "outlier"):
        assert np.array_equal(getattr(got, name), getattr(fr, name)), name
    assert got.ref_frame == 2 and got.U_local is None
    assert np.array_equal(m["elements"], mesh.elements) and np.array_equal(m["edge_ok"], mesh.edge_ok)
    assert np.array_equal(m["node_valid"], mesh.node_valid)
    a, b = got.admm, fr.admm
    assert (a.beta, a.mu, a.n_steps, a.subpb2_time) == (b.beta, b.mu, b.n_steps, b.subpb2_time)
    assert a.update_global == b.update_global and a.primal_residual_f == b.primal_residual_f
    assert len(a.local_info) == 2 and a.local_info[1].n_bad == 3 and np.array_equal(a.local_info[0].zncc, b.local_info[0].zncc)
    assert a.beta_sweep["k_best"] == 1 and np.array_equal(a.beta_sweep["beta"], b.beta_sweep["beta"])
    meta = json.loads((tmp_path / C.META_NAME).read_text())
    assert meta["format"] == 1 and "verbose" not in meta["para"] and meta["grid_shape"] == [3, 2, 3]
    assert meta["x0"] == [8.0, 16.0, 24.0] and meta["y0"] == [8.0, 16.0] and meta["z0"] == [8.0, 16.0, 24.0]


def test_resume_checks_the_run_and_a_fresh_start_clears_frames(tmp_path):
    mesh = _mesh()
    ck = C.Checkpoint(tmp_path)
    meta = C.meta_for({"winsize": [16, 16, 16], "mu": 1e-3}, (40, 32, 40), [0, 0], mesh, 3)
    ck.prepare(meta, resume=False)
    ck.save(0, _result(mesh.n_nodes, np.random.default_rng(1), admm=False), mesh)
    ck.prepare(json.loads(json.dumps(meta)), resume=True)  # the same run: frames kept
    assert ck.completed_frames() == [0]
    other = C.meta_for({"winsize": [16, 16, 16], "mu": 2e-3}, (40, 32, 40), [0, 0], mesh, 3)
    with pytest.raises(C.CheckpointMismatch, match="para.mu"):
        ck.prepare(other, resume=True)
    ck.prepare(other, resume=False)
    assert ck.completed_frames() == []
    fr, _ = ck.load(0) if ck.has(0) else (None, None)
    assert fr is None


# ---------------------------------------------------------------------------- interop with the reference
REF_DIR = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "ckpt_ref")


def _ref_arrays():
    with np.load(f"{REF_DIR}/frame_0001.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_reads_a_frame_the_reference_wrote():
    """tests/golden/ckpt_ref was written by the reference's Checkpoint.save (make_golden_checkpoint.py)."""
    raw = _ref_arrays()
    fr, mesh = C.Checkpoint(REF_DIR).load(1)
    for name in ("U", "F", "U_accum", "U_local", "F_local", "U0", "zncc", "U_std", "status", "split_fraction", "outlier"):
        assert np.array_equal(getattr(fr, name), raw[name]), name
    assert fr.ref_frame == 0 and fr.outlier.dtype == bool
    for name in ("node_valid", "elements", "boundary_nodes", "edge_ok"):
        assert np.array_equal(mesh[name], raw[name]), name
    a = fr.admm
    assert (a.beta, a.mu, a.n_steps, a.subpb2_time) == tuple(raw["admm_scalars"][[0, 1]]) + (4, 0.75)
    assert a.update_global == tuple(raw["admm_update_global"]) and a.primal_residual_f == tuple(raw["admm_primal_residual_f"])
    assert len(a.local_info) == 2
    for i, li in enumerate(a.local_info):
        assert np.array_equal(li.n_iter, raw[f"admm_local_{i}_n_iter"]) and np.array_equal(li.status, raw[f"admm_local_{i}_status"])
        assert (li.solve_time, li.n_bad) == (0.125 * (i + 1), i + 2)
    assert a.beta_sweep["k_best"] == 1 and isinstance(a.beta_sweep["k_best"], int)
    assert np.array_equal(a.beta_sweep["betas"], raw["admm_sweep_betas"])


def test_writes_the_reference_layout(tmp_path):
    """A frame written here has the reference file's keys, dtypes, shapes and values."""
    raw = _ref_arrays()
    fr, mesh = C.Checkpoint(REF_DIR).load(1)
    ck = C.Checkpoint(tmp_path)
    path = ck.save(1, fr, SimpleNamespace(**mesh))
    with np.load(path, allow_pickle=False) as z:
        mine = {k: z[k] for k in z.files}
    assert sorted(mine) == sorted(raw)
    for k in raw:
        assert (mine[k].dtype, mine[k].shape) == (raw[k].dtype, raw[k].shape), k
        assert np.array_equal(mine[k], raw[k]), k
    assert [p.name for p in tmp_path.iterdir()] == ["frame_0001.npz"]  # no partial file left behind


def test_resumes_on_the_reference_meta_with_attribute_style_parameters(tmp_path):
    import json
    import shutil

    stored = json.loads(open(f"{REF_DIR}/meta.json").read())
    p = dict(stored["para"])
    # the parameter object as a caller holds it: attributes, a VOI range object, a schedule object, tuples
    p["voi"] = SimpleNamespace(**{a: tuple(v) for a, v in p["voi"].items()})
    p["frame_schedule"] = SimpleNamespace(ref_indices=tuple(p["frame_schedule"]["ref_indices"]))
    p = {k: tuple(v) if isinstance(v, list) else v for k, v in p.items()}
    p["verbose"] = False  # left out of the comparison, as the reference leaves it out of meta.json
    para = SimpleNamespace(**p)
    mesh = SimpleNamespace(grid_shape=tuple(stored["grid_shape"]), x0=stored["x0"], y0=stored["y0"], z0=stored["z0"])
    meta = C.run_meta(para, tuple(stored["volume_shape"]), para.frame_schedule, mesh, stored["n_frames"])
    assert C.meta_differences(stored, meta) == []
    assert meta["para"]["voi"] == stored["para"]["voi"] and "verbose" not in meta["para"]
    shutil.copytree(REF_DIR, tmp_path / "run")
    ck = C.Checkpoint(tmp_path / "run")
    ck.prepare(meta, resume=True)
    assert ck.completed_frames() == [1]
    other = C.run_meta(SimpleNamespace(**{**p, "mu": 5e-3}), tuple(stored["volume_shape"]), [0, 0], mesh, 3)
    with pytest.raises(C.CheckpointMismatch, match="para.mu"):
        ck.prepare(other, resume=True)
