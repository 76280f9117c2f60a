"""The oracle's node-grid cleanup and the noise pattern against the reference's own outputs (CPU).

Golden vectors: ``tests/golden/make_golden_driver.py`` (``universal_median_test``,
``fill_nan_grid``, ``fill_nan_nearest``, ``noise_hessian_pattern`` of al_dvc 1.2.1).
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from oracle import cleanup_ref as C
from pyaldvc_amd import noise


@pytest.fixture(scope="module")
def gd(request):
    import os

    return np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_driver.npz"))


def test_median_test_matches_reference(gd):
    f, v, e = gd["mt_field"], gd["mt_valid"], gd["mt_edge_ok"]
    assert np.array_equal(C.universal_median_test(f, v, 2.0), gd["mt_flag"])
    assert np.array_equal(C.universal_median_test(f, v, 1.0, eps=0.05), gd["mt_flag_t1"])
    assert np.array_equal(C.universal_median_test(f, v, 2.0, edge_ok=e), gd["mt_flag_cut"])
    assert np.array_equal(C.universal_median_test(f[..., 0], None, 2.0), gd["mt_flag_novalid"])
    assert not C.universal_median_test(f, v, 0.0).any()


def test_rank_median_is_scipys_median_filter():
    """The neighbourhood median is ndimage's rank filter over the centre-less footprint."""
    rng = np.random.default_rng(3)
    a = rng.standard_normal((6, 7, 5))
    fp = np.ones((3, 3, 3), bool)
    fp[1, 1, 1] = False
    assert np.array_equal(C._rank_median(a, 3), ndimage.median_filter(a, footprint=fp, mode="nearest"))
    fp5 = np.ones((5, 5, 5), bool)
    fp5[2, 2, 2] = False
    assert np.array_equal(C._rank_median(a, 5), ndimage.median_filter(a, footprint=fp5, mode="nearest"))


def test_fill_matches_reference(gd):
    g, e = gd["fill_grid"], gd["mt_edge_ok"]
    np.testing.assert_allclose(C.fill_nan_grid(g), gd["fill_spring"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(C.fill_nan_grid(g, edge_ok=e), gd["fill_cut"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(C.fill_nan_nearest(g), gd["fill_nearest"])
    np.testing.assert_allclose(C.fill_nan_grid(gd["fill_iso_grid"]), gd["fill_iso"], rtol=0, atol=1e-12)


def test_spring_fill_is_harmonic_and_exact_for_linear_fields():
    """The spring model solves the discrete Laplace equation: a linear field is reproduced exactly
    in the interior of a hole, and every filled node is the mean of its grid neighbours."""
    z, y, x = np.meshgrid(np.arange(7.0), np.arange(8.0), np.arange(9.0), indexing="ij")
    lin = 0.3 * x - 0.7 * y + 1.1 * z + 2.0
    g = lin.copy()
    hole = np.zeros(g.shape, bool)
    hole[2:5, 3:6, 2:7] = True
    g[hole] = np.nan
    out = C.fill_nan_grid(g)
    np.testing.assert_allclose(out, lin, atol=1e-8)
    rng = np.random.default_rng(1)
    g = rng.standard_normal((6, 6, 6))
    g[hole[:6, :6, :6]] = np.nan
    out = C.fill_nan_grid(g)
    pad = np.pad(out, 1, mode="constant", constant_values=np.nan)
    nb = np.stack([pad[1 + dz : 7 + dz, 1 + dy : 7 + dy, 1 + dx : 7 + dx] for dz, dy, dx in
                   ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))])  # fmt: skip
    mean = np.nanmean(nb, axis=0)
    h = hole[:6, :6, :6]
    np.testing.assert_allclose(out[h], mean[h], atol=1e-8)


def test_noise_pattern_matches_reference(gd):
    import os

    g0 = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_local_step.npz"))
    assert np.array_equal(noise.noise_hessian_pattern((8, 8, 6), 1), gd["pattern_886_s1"])
    assert np.array_equal(noise.noise_hessian_pattern((12, 12, 9), 3), gd["pattern_12129_s3"])
    assert np.array_equal(noise.noise_hessian_pattern((6, 6, 6), 1), g0["pattern_666_s1"])
    assert np.array_equal(noise.noise_hessian_pattern((6, 5, 4), 2), g0["pattern_654_s2"])
    assert noise.STENCIL_NOISE_GAIN == float(g0["stencil_noise_gain"])


def test_noise_pattern_is_expected_hessian_of_white_gradient_noise():
    """Brute force: E[J J^T] for i.i.d. unit-variance gradient noise is sum_voxels of the
    per-voxel outer product with the noise covariance I_3."""
    half, stride = (3, 2, 4), 2
    pts = [(X, Y, Z) for Z in range(-4, 5, 2) for Y in range(-2, 3, 2) for X in range(-3, 4, 2)]
    E = np.zeros((12, 12))
    for X, Y, Z in pts:
        for c in range(3):  # gradient component c carries unit noise, the others none
            J = np.zeros(12)
            J[3 * c : 3 * c + 3] = (X, Y, Z)
            J[9 + c] = 1.0
            E += np.outer(J, J)
    np.testing.assert_allclose(noise.noise_hessian_pattern(half, stride), E)


def test_hex8_element_matrices_match_reference(gd):
    from pyaldvc_amd import fem

    for order in (1, 2, 3):
        box = fem.hex8_box_matrices((5.0, 6.0, 7.0), order)
        for key in ("K", "M", "G"):
            np.testing.assert_allclose(box[key], gd[f"hex8_{key}_o{order}"], rtol=0, atol=1e-14)


def test_hex8_element_matrices_closed_forms():
    """Row sums of K vanish (constants have no energy), M sums to the box volume, and G_j applied
    to a linear field's nodal values integrates its gradient: sum_a (G_j u)_a = slope_j * volume."""
    from pyaldvc_amd import fem

    h = np.array([2.0, 3.0, 5.0])
    box = fem.hex8_box_matrices(h, 2)
    np.testing.assert_allclose(box["K"].sum(axis=1), 0.0, atol=1e-13)
    assert abs(box["M"].sum() - h.prod()) < 1e-12
    corners = (fem.HEX8_SIGNS + 1.0) / 2.0 * h
    slope = np.array([0.7, -1.3, 0.4])
    u = corners @ slope
    for j in range(3):
        # G_j[a, b] = int dN_a/dx_j N_b: sum over a of dN_a vanishes, so use the transpose
        assert abs((box["G"][j].T @ u).sum() - slope[j] * h.prod()) < 1e-12
