"""Displacement uncertainty on the GPU (pyaldvc_amd/uncertainty.py, aldvc_displacement_uncertainty)
against the oracle (oracle/uncertainty_ref.py) and against the empirical error of noisy solves, as
the reference's tests/test_uncertainty.py:16-45 checks it."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import uncertainty_ref as U
from pyaldvc_amd import kernels as K
from pyaldvc_amd import synthetic as syn
from pyaldvc_amd.uncertainty import displacement_uncertainty

pytestmark = pytest.mark.gpu


def _ctx(H, valid, n_valid, bottomf, half=(8, 8, 8), stride=1):
    return SimpleNamespace(H_all=H, valid=valid, n_valid=n_valid, bottomf=bottomf, half=half, stride=stride)


def _random_case(n, seed, big_noise=False):
    rng = np.random.default_rng(seed)
    rows = 60
    J = rng.normal(size=(n, rows, 12)) * np.array([4.0] * 9 + [1.0, 1.3, 0.8])
    H = np.einsum("nri,nrj->nij", J, J)
    zncc = rng.uniform(0.3 if big_noise else 0.9, 0.999, n)
    n_valid = rng.integers(10, 5000, n)
    bottomf = rng.uniform(0.5, 3.0, n)
    valid = rng.random(n) > 0.05
    status = np.where(rng.random(n) > 0.1, 0, rng.integers(1, 6, n)).astype(np.int8)
    zncc[rng.random(n) < 0.02] = np.nan
    return H, zncc, valid, n_valid, bottomf, status


@pytest.mark.parametrize("noise_correction", [True, False])
@pytest.mark.parametrize("big_noise", [False, True])  # True: the correction's 10 % floor binds often
def test_uncertainty_matches_oracle_on_random_hessians(noise_correction, big_noise):
    H, zncc, valid, n_valid, bottomf, status = _random_case(3001, 3 + big_noise, big_noise)
    ctx = _ctx(H, valid, n_valid, bottomf, half=(6, 7, 5))
    got = displacement_uncertainty(ctx, zncc, status, noise_correction)
    ref = U.displacement_uncertainty(H, zncc, valid, n_valid, bottomf, (6, 7, 5), status=status,
                                     noise_correction=noise_correction)  # fmt: skip
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    fin = np.isfinite(ref)
    assert fin.sum() > 2000
    assert np.max(np.abs(got[fin] - ref[fin]) / ref[fin]) < 1e-9
    # without a status every valid node with a finite ZNCC gets a value
    got2 = displacement_uncertainty(ctx, zncc, None, noise_correction)
    ref2 = U.displacement_uncertainty(H, zncc, valid, n_valid, bottomf, (6, 7, 5), noise_correction=noise_correction)
    assert np.array_equal(np.isnan(got2), np.isnan(ref2))


def test_singular_hessian_batch_takes_the_pseudo_inverse():
    H, zncc, valid, n_valid, bottomf, status = _random_case(257, 9)
    H[5, 2, :] = 0.0
    H[5, :, 2] = 0.0
    valid[5], status[5], zncc[5], n_valid[5] = True, 0, 0.95, 3000
    with pytest.raises(np.linalg.LinAlgError):  # the uncorrected Hessians: exactly singular at node 5
        np.linalg.inv(H)
    got = displacement_uncertainty(_ctx(H, valid, n_valid, bottomf), zncc, status, noise_correction=False)
    ref = U.displacement_uncertainty(H, zncc, valid, n_valid, bottomf, (8, 8, 8), status=status, noise_correction=False)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    fin = np.isfinite(ref)
    assert np.isfinite(got[5]).all()
    assert np.max(np.abs(got[fin] - ref[fin]) / ref[fin]) < 1e-8


def test_device_tensors_in_device_tensor_out():
    H, zncc, valid, n_valid, bottomf, status = _random_case(100, 11)
    dev = K.device()
    ctx = SimpleNamespace(device={"H": torch.from_numpy(H).to(dev), "valid": torch.from_numpy(valid.astype(np.uint8)).to(dev),
                                  "n_valid": torch.from_numpy(n_valid.astype(np.int64)).to(dev),
                                  "bottomf": torch.from_numpy(bottomf).to(dev)}, half=(8, 8, 8), stride=1)  # fmt: skip
    got = displacement_uncertainty(ctx, torch.from_numpy(zncc).to(dev), torch.from_numpy(status).to(dev))
    assert isinstance(got, torch.Tensor) and got.is_cuda
    ref = U.displacement_uncertainty(H, zncc, valid, n_valid, bottomf, (8, 8, 8), status=status)
    g = got.cpu().numpy()
    fin = np.isfinite(ref)
    assert np.array_equal(np.isnan(g), np.isnan(ref)) and np.max(np.abs(g[fin] - ref[fin]) / ref[fin]) < 1e-9


def _solve_noisy(f_raw, g_raw, coords, U_true, half):
    f, g = syn.normalize_np(f_raw), syn.normalize_np(g_raw)
    gx, gy, gz = K.compute_gradients(f)
    nomask = np.ones((1, 1, 1), np.uint8)
    H, L, mf, bf, nv, valid = K.precompute_nodes(coords, half, half, half, f, gx, gy, gz, nomask, 0.5, 1e12)
    P0 = np.zeros((coords.shape[0], 12))
    P0[:, 9:] = np.round(U_true)
    P, it, st, zn = K.icgn_12dof_parallel(coords, P0, half, half, half, f, gx, gy, gz, nomask, g, 1, L, mf, bf, valid,
                                          1e-2, 1e-3, 100, 5)  # fmt: skip
    ctx = _ctx(H, valid.astype(bool), nv, bf, half=(half,) * 3)
    return P[:, 9:], st, displacement_uncertainty(ctx, zn, st), (H, zn, valid, nv, bf, st)


def test_predicted_std_tracks_the_empirical_error():
    """tests/test_uncertainty.py:16-45 on a bead pair under an affine field: the predicted standard
    deviation of u, v, w is within (0.6, 1.5) of the RMS error over interior converged nodes, grows
    with the noise, and equals the oracle on the same Hessians; nodes that did not converge carry
    no estimate."""
    shape = (80, 80, 80)
    F = syn.strain_rotation_F(0.01, 2.0)
    disp = syn.affine(F, tuple((s - 1) / 2 for s in shape[::-1]), t=(0.37, -0.62, 1.15))
    coords = syn.node_lattice(shape, (16, 16, 16), (8, 8, 8))
    u, v, w = disp(*(coords[:, a].astype(float) for a in range(3)))
    U_true = np.column_stack([u, v, w])
    interior = np.all((coords >= 16) & (coords <= 64), axis=1)
    preds = []
    # noise levels where the random error dominates the interpolation bias of a bead texture (about
    # 0.004 voxel in x and y and 0.007 in z at noise 0.02, where the ratio drops to 0.4-0.6)
    for k, noise in enumerate((0.05, 0.1, 0.16)):
        f_raw, g_raw = syn.bead_pair_np(shape, 3500, 2.0, disp, seed=30 + k, noise=noise)
        Uu, st, std, (H, zn, valid, nv, bf, st2) = _solve_noisy(f_raw, g_raw, coords, U_true, 8)
        ref = U.displacement_uncertainty(H, zn, valid.astype(bool), nv, bf, (8, 8, 8), status=st2)
        fin = np.isfinite(ref)
        assert np.array_equal(np.isnan(std), np.isnan(ref))
        assert np.max(np.abs(std[fin] - ref[fin]) / ref[fin]) < 1e-9
        ok = interior & (st == 0) & np.all(np.isfinite(std), axis=1)
        assert ok.sum() > 40
        emp = np.sqrt(np.mean((Uu - U_true)[ok] ** 2, axis=0))
        pred = np.sqrt(np.mean(std[ok] ** 2, axis=0))
        ratio = pred / emp
        assert np.all((0.6 < ratio) & (ratio < 1.5)), (noise, emp, pred)
        preds.append(float(pred.mean()))
        assert np.all(np.isnan(std[st != 0]))
    assert preds[0] < preds[1] < preds[2]
