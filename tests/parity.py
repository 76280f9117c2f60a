"""Shared helpers of the GPU parity tests: bead-pair cases and the reference's parity criteria."""

from __future__ import annotations

import numpy as np

from pyaldvc_amd import synthetic as syn

MASK_NONE = np.ones((1, 1, 1), dtype=np.uint8)
OTF = np.zeros((1, 1, 1), dtype=np.float32)


def assert_icgn_parity(ref, got, tol_u=1e-3, tol_F=1e-4, tol_zncc=1e-3, min_iter_agree=0.95):
    """The reference's GPU-vs-CPU criteria (tests/test_cuda_backend.py:149-161), tightened to the
    1e-3 voxel sub-voxel bar: identical status codes, >= 95 % identical iteration counts,
    converged translations within ``tol_u`` voxel and F within ``tol_F``, ZNCC within ``tol_zncc``
    where both are finite."""
    Pc, itc, stc, zc = ref
    Pg, itg, stg, zg = got
    assert np.array_equal(stc, stg), f"status codes differ at {np.flatnonzero(stc != stg)[:10]}: {stc[stc != stg][:10]} vs {stg[stc != stg][:10]}"
    agree = np.mean(itc == itg) if itc.size else 1.0
    assert agree >= min_iter_agree, f"iteration counts agree at only {agree:.3f} of the nodes"
    ok = stc == 0
    if ok.any():
        if Pc.ndim == 2 and Pc.shape[1] == 12:
            du = np.max(np.abs(Pg[ok, 9:] - Pc[ok, 9:]))
            dF = np.max(np.abs(Pg[ok, :9] - Pc[ok, :9]))
            assert du < tol_u, f"max |du| = {du:.3g} voxel"
            assert dF < tol_F, f"max |dF| = {dF:.3g}"
        else:
            du = np.max(np.abs(Pg[ok] - Pc[ok]))
            assert du < tol_u, f"max |du| = {du:.3g} voxel"
    fin = np.isfinite(zc) & np.isfinite(zg)
    if fin.any():
        assert np.max(np.abs(zg[fin] - zc[fin])) < tol_zncc


def bead_case(shape=(64, 72, 80), density=0.006, sigma=1.5, F=None, t=(1.3, -0.7, 0.9), seed=5, noise=0.01):
    """A normalised bead pair under an affine field with its stored gradients."""
    from oracle import cpu_ref

    F = np.array([[0.02, 0.01, 0.0], [0.0, -0.015, 0.005], [0.003, 0.0, 0.01]]) if F is None else np.asarray(F)
    centre = tuple((s - 1) / 2 for s in shape[::-1])
    disp = syn.affine(F, centre, t)
    n_beads = int(density * np.prod(shape))
    f_raw, g_raw = syn.bead_pair_np(shape, n_beads, sigma, disp, seed=seed, noise=noise)
    f = syn.normalize_np(f_raw)
    g = syn.normalize_np(g_raw)
    gx, gy, gz = cpu_ref.compute_gradients(f)
    return {"f": f, "g": g, "gx": gx, "gy": gy, "gz": gz, "disp": disp, "F": F, "shape": shape}


def lattice(shape, winsize, step):
    return syn.node_lattice(shape, (winsize,) * 3, (step,) * 3)


def truth_at(case, coords):
    u, v, w = case["disp"](coords[:, 0].astype(float), coords[:, 1].astype(float), coords[:, 2].astype(float))
    return np.column_stack([u, v, w])
