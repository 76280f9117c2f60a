"""Volume preparation on the GPU against the reference (needs an MI355X).

``normalize_volume`` / ``prefilter_bspline`` / ``prepare_deformed`` outputs of al_dvc 1.2.1 on the
golden bead pair (``golden_local_step.npz``); the B-spline prefilter also against
``scipy.ndimage.spline_filter`` (the routine the reference calls) on odd shapes down to one voxel;
statistics against float64 NumPy on uint16 data and clamped VOIs.
"""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest
from scipy import ndimage

pytestmark = pytest.mark.gpu

V = pytest.importorskip("pyaldvc_amd.volumes")
K = pytest.importorskip("pyaldvc_amd.kernels")


def _ulps(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.max(np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64)))


def test_normalize_matches_reference(golden):
    for raw, ref in ((golden["f_raw"], golden["f"]), (golden["g_raw"], golden["g"])):
        got = V.normalize_volume(raw)
        assert got.dtype == np.float32
        # the reference rounds (v - mean) to float32 before dividing: at most a couple of ulps apart
        assert np.max(np.abs(got - ref)) <= 4 * np.finfo(np.float32).eps * np.max(np.abs(ref))
        m, s = np.mean(raw, dtype=np.float64), np.std(raw, dtype=np.float64)
        exact = ((raw.astype(np.float64) - m) / s).astype(np.float32)
        assert _ulps(got, exact) <= 1


def test_voi_statistics_uint16():
    rng = np.random.default_rng(4)
    vol = rng.integers(100, 60000, size=(21, 34, 45), dtype=np.uint16)
    voi = SimpleNamespace(x=(5, 40), y=(3, -1), z=(-2, 50))  # clamped to the volume like VOIRange
    mean, std = V.voi_mean_std(vol, voi)
    sub = vol[0:21, 3:34, 5:41].astype(np.float64)
    assert abs(mean - sub.mean()) <= 1e-9 * sub.mean()
    assert abs(std - sub.std()) <= 1e-9 * sub.std()
    got = V.normalize_volume(vol, voi)
    exact = ((vol.astype(np.float64) - sub.mean()) / sub.std()).astype(np.float32)
    assert _ulps(got, exact) <= 1
    flat = np.full((5, 6, 7), 3.0, np.float32)  # zero spread: the reference divides by 1
    assert np.array_equal(V.normalize_volume(flat), np.zeros_like(flat))


def test_bspline_prefilter_matches_reference_and_scipy(golden):
    got = V.prefilter_bspline(golden["g"])
    ref = golden["g_bspline"]
    assert np.max(np.abs(got - ref)) <= 2e-6 * np.max(np.abs(ref))
    rng = np.random.default_rng(8)
    for shape in ((1, 1, 1), (2, 3, 1), (5, 33, 2), (40, 9, 70), (65, 17, 34)):
        vol = rng.standard_normal(shape).astype(np.float32)
        exp = ndimage.spline_filter(vol, order=3, output=np.float32, mode="mirror")
        out = V.prefilter_bspline(vol)
        assert np.max(np.abs(out - exp)) <= 2e-6 * max(1.0, float(np.max(np.abs(exp)))), shape


def test_bspline_coefficients_interpolate_the_samples():
    """Sampling the coefficients with the cubic B-spline basis at the grid points returns the
    volume (the defining property), through the HIP sampler."""
    rng = np.random.default_rng(9)
    vol = ndimage.gaussian_filter(rng.standard_normal((24, 26, 28)), 1.0).astype(np.float32)
    coef = V.prefilter_bspline(vol)
    z, y, x = np.meshgrid(np.arange(2, 20), np.arange(2, 22), np.arange(2, 24), indexing="ij")
    s = K.sample_points(coef, z.ravel().astype(float), y.ravel().astype(float), x.ravel().astype(float), 2)
    assert np.max(np.abs(s - vol[z, y, x].ravel())) < 2e-5


def test_prepare_deformed_matches_reference(golden):
    dmask = np.ones(golden["g"].shape, dtype=bool)
    dmask[:11, :, :] = False
    got = V.prepare_deformed(golden["g"], "cubic", dmask)
    ref = golden["g_nan"]
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    fin = ~np.isnan(ref)
    assert np.array_equal(got[fin], ref[fin])
    gb = V.prepare_deformed(golden["g"], "bspline", dmask)
    assert np.array_equal(np.isnan(gb), ~dmask)
    assert np.max(np.abs(gb[dmask] - golden["g_bspline"][dmask])) <= 2e-6 * np.max(np.abs(golden["g_bspline"]))
    assert np.array_equal(V.prepare_deformed(golden["g"], "linear"), golden["g"])


def test_reference_bundle(golden):
    b = V.build_reference_bundle(golden["f"], golden["mask"])
    assert b.stored_gradients and b.has_mask
    for got, ref in zip((b.gx, b.gy, b.gz), (golden["gx"], golden["gy"], golden["gz"])):
        assert np.max(np.abs(got.cpu().numpy() - ref)) < 2e-6 * max(1.0, float(np.max(np.abs(ref))))
    assert np.array_equal(b.mask.cpu().numpy(), golden["mask"])
    b2 = V.build_reference_bundle(golden["f"], None, "on_the_fly")
    assert not b2.stored_gradients and not b2.has_mask
    with pytest.raises(ValueError):
        V.build_reference_bundle(golden["f"], np.ones((3, 3, 3), np.uint8))
