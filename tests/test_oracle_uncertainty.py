"""The oracle's displacement uncertainty (oracle/uncertainty_ref.py) pinned by hand-derived closed
forms and by a Monte-Carlo least-squares experiment (solver/uncertainty.py:83-133)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import uncertainty_ref as U
from pyaldvc_amd import noise


def test_moment_matrix_and_pattern_equal_the_pinned_noise_module():
    for half, stride in (((8, 8, 8), 1), ((6, 4, 10), 2), ((3, 5, 2), 3)):
        assert np.array_equal(U.subset_moment_matrix(half, stride), noise.subset_moment_matrix(half, stride))
        assert np.array_equal(U.noise_hessian_pattern(half, stride), noise.noise_hessian_pattern(half, stride))
    # by hand: half 1 per axis, stride 1 -> 27 voxels, sum X^2 = 2 * 9 = 18
    assert np.array_equal(np.diagonal(U.subset_moment_matrix((1, 1, 1))), [18.0, 18.0, 18.0, 27.0])


def _diag_case(noise_level):
    h = np.array([50.0, 51.0, 52.0, 53.0, 54.0, 55.0, 56.0, 57.0, 58.0, 0.3, 0.4, 0.5])
    n = 4
    H = np.zeros((n, 12, 12))
    for i in range(n):
        H[i] = np.diag(h * (1.0 + 0.1 * i))
    zncc = np.array([1.0 - noise_level, 1.0 - 2 * noise_level, 0.95, 0.99])
    n_valid = np.array([4913, 4000, 2500, 4913])
    bottomf = np.array([1.0, 2.0, 0.5, 3.0])
    return H, zncc, n_valid, bottomf


@pytest.mark.parametrize("noise_level", [1e-3, 0.3])  # 0.3: the 10 % floor of the correction binds
def test_diagonal_hessian_closed_form(noise_level):
    """For diagonal H: var_t = f / H0_t + c s^4 n_valid / H0_t^2 with
    H0_p = H_p - c s^2 (n_valid / n) D_p lim and lim = min(1, min_t 0.9 H_t / (c s^2 n_valid))."""
    half = (8, 8, 8)
    H, zncc, n_valid, bottomf = _diag_case(noise_level)
    D = np.diagonal(U.noise_hessian_pattern(half))
    n = D[9]
    c = U.STENCIL_NOISE_GAIN
    valid = np.ones(4, bool)
    got_plain = U.displacement_uncertainty(H, zncc, valid, n_valid, bottomf, half, noise_correction=False)
    got = U.displacement_uncertainty(H, zncc, valid, n_valid, bottomf, half)
    bound = False
    for i in range(4):
        om = 1.0 - zncc[i]
        f = 2.0 * om / (n_valid[i] - 12) * bottomf[i] ** 2
        s2 = om * bottomf[i] ** 2 / n_valid[i]
        Hd = np.diagonal(H[i])
        for k, t in enumerate((9, 10, 11)):
            assert got_plain[i, k] == pytest.approx(np.sqrt(f / Hd[t]), rel=1e-12)
        cd = c * s2 * n_valid[i]
        lim = min(1.0, min(0.9 * Hd[t] / cd for t in (9, 10, 11)))
        bound |= lim < 1.0
        H0 = Hd - c * s2 * (n_valid[i] / n) * D * lim
        for k, t in enumerate((9, 10, 11)):
            want = f / H0[t] + c * s2**2 * n_valid[i] / H0[t] ** 2
            assert got[i, k] == pytest.approx(np.sqrt(want), rel=1e-12)
    assert bound == (noise_level > 0.1)


def test_nodes_without_an_estimate_are_nan():
    H, zncc, n_valid, bottomf = _diag_case(1e-3)
    zncc = zncc.copy()
    zncc[1] = np.nan
    n_valid = n_valid.copy()
    n_valid[2] = 14  # needs more than 12 + 2 voxels
    valid = np.array([True, True, True, False])
    st = np.zeros(4, np.int8)
    out = U.displacement_uncertainty(H, zncc, valid, n_valid, bottomf, (8, 8, 8), status=st)
    assert np.isfinite(out[0]).all() and np.isnan(out[1:]).all()
    st[0] = 2
    assert np.isnan(U.displacement_uncertainty(H, zncc, valid, n_valid, bottomf, (8, 8, 8), status=st)).all()


def test_singular_hessian_takes_the_pseudo_inverse():
    """An exactly singular H makes the batched inverse fail; the pseudo-inverse of a diagonal matrix
    with a zero F entry leaves the translation block's closed form unchanged."""
    H, zncc, n_valid, bottomf = _diag_case(1e-3)
    H[2, 0, 0] = 0.0
    valid = np.ones(4, bool)
    with pytest.raises(np.linalg.LinAlgError):
        np.linalg.inv(H)
    out = U.displacement_uncertainty(H, zncc, valid, n_valid, bottomf, (8, 8, 8), noise_correction=False)
    f = 2.0 * (1.0 - zncc[2]) / (n_valid[2] - 12) * bottomf[2] ** 2
    assert out[2, 0] == pytest.approx(np.sqrt(f / H[2, 9, 9]), rel=1e-10)


def test_uncorrected_model_matches_a_monte_carlo_least_squares_covariance():
    """y = J p + e, e ~ N(0, sigma^2), p_hat = (J^T J)^-1 J^T y: the model with Sum r^2 = 2 (1 - ZNCC)
    (bottomf = 1) and n_valid = rows of J predicts Var(p_hat_t) = sigma^2 [(J^T J)^-1]_tt on
    average. Few rows (n = 40) make the n - 12 degrees of freedom matter (30 % at n)."""
    rng = np.random.default_rng(7)
    n, trials, sigma = 40, 6000, 0.3
    J = rng.normal(size=(n, 12)) * np.array([3.0] * 9 + [1.0, 1.5, 0.7])
    Hm = J.T @ J
    p_true = rng.normal(size=12)
    e = sigma * rng.normal(size=(trials, n))
    y = (J @ p_true)[None, :] + e
    p_hat = np.linalg.solve(Hm, J.T @ y.T).T
    r = y - p_hat @ J.T
    zncc = 1.0 - 0.5 * np.sum(r * r, axis=1)
    H = np.broadcast_to(Hm, (trials, 12, 12))
    pred = U.displacement_uncertainty(H, zncc, np.ones(trials, bool), np.full(trials, n), np.ones(trials), (8, 8, 8),
                                      noise_correction=False)  # fmt: skip
    emp_var = np.var(p_hat[:, 9:12], axis=0)
    pred_var = np.mean(pred**2, axis=0)
    assert np.allclose(pred_var, sigma**2 * np.diagonal(np.linalg.inv(Hm))[9:12], rtol=0.05)
    assert np.allclose(pred_var, emp_var, rtol=0.08)
