"""Initial-guess orchestration on the GPU (pyaldvc_amd/init_guess.py) against the oracle
(oracle/ncc_ref.py) and the reference's own test cases."""

from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage
from types import SimpleNamespace

from oracle import ncc_ref
from pyaldvc_amd import cleanup
from pyaldvc_amd import init_guess as IG
from pyaldvc_amd import synthetic as syn
from pyaldvc_amd.local import NodeMesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("factor", [2, 3, 4])
def test_block_downsample_matches_oracle(factor):
    v = np.random.default_rng(factor).normal(size=(37, 42, 51)).astype(np.float32)
    got = IG.block_downsample(v, factor).cpu().numpy()
    ref = ncc_ref.block_downsample(v, factor)
    assert got.shape == ref.shape
    # float64 block sums on the device, NumPy's float32 mean in the oracle: a few ulps apart
    assert np.max(np.abs(got - ref)) <= 4 * np.finfo(np.float32).eps * max(1.0, float(np.max(np.abs(ref))))


@pytest.mark.parametrize("shift", [(3, -2, 5), (-7, 4, 0)])
def test_phase_correlation_matches_oracle_on_a_cyclic_shift(shift):
    rng = np.random.default_rng(1)
    f = ndimage.gaussian_filter(rng.normal(size=(40, 48, 56)), 1.5).astype(np.float32)
    g = np.roll(f, (shift[2], shift[1], shift[0]), axis=(0, 1, 2))
    assert np.array_equal(IG.phase_correlation_shift(f, g), np.array(shift, dtype=float))


def _pair(shift, shape=(80, 72, 72), seed=12):
    f, g = syn.bead_pair_np(shape, 5000, 2.0, syn.translation(shift), seed=seed, noise=0.0)
    return syn.normalize_np(f), syn.normalize_np(g)


def _clean_host(d, grid_shape, threshold):
    d = np.array(d, dtype=np.float64).reshape(tuple(grid_shape) + (3,))
    valid = np.all(np.isfinite(d), axis=-1)
    if threshold > 0 and valid.sum() > 27:
        d[cleanup.universal_median_test(d, valid, threshold)] = np.nan
    return np.stack([cleanup.fill_nan_grid(d[..., c]) for c in range(3)], -1).reshape(-1, 3)


def test_phase_correlation_and_levels_match_oracle_on_a_bead_pair():
    f, g = _pair((11.3, -9.6, 6.4))
    assert np.array_equal(IG.phase_correlation_shift(f, g), ncc_ref.phase_correlation_shift(f, g))
    for shape in ((80, 72, 72), (80, 72, 40)):
        assert IG.auto_pyramid_levels(shape, (16, 16, 16), (4, 4, 4)) == ncc_ref.auto_pyramid_levels(shape, (16, 16, 16), (4, 4, 4))
    assert IG.auto_pyramid_levels(f.shape, (16, 16, 16), (4, 4, 4), f=f) == ncc_ref.auto_pyramid_levels(
        f.shape, (16, 16, 16), (4, 4, 4), f=f)  # fmt: skip


@pytest.mark.parametrize("use_shift", [False, True])
def test_pyramid_search_matches_oracle_and_recovers_the_large_shift(use_shift):
    """tests/test_integer_search.py:74-90: every node within 0.3 voxel of the imposed shift; the
    device path (HIP maps, device median test and spring fill) agrees with the oracle."""
    shift = (11.3, -9.6, 6.4)
    f, g = _pair(shift)
    ax = np.arange(24, 50, 8)
    gz, gy, gx = np.meshgrid(ax, ax, ax, indexing="ij")
    coords = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1).astype(np.int64)
    gs = ncc_ref.phase_correlation_shift(f, g) if use_shift else None
    levels = 0 if use_shift else ncc_ref.auto_pyramid_levels(f.shape, (16, 16, 16), (4, 4, 4))
    ref = ncc_ref.pyramid_search(f, g, coords, (4, 4, 4), (16, 16, 16), (4, 4, 4), levels, _clean_host, global_shift=gs)
    got = IG.pyramid_search(f, g, coords, (4, 4, 4), (16, 16, 16), (4, 4, 4), levels=levels, global_shift=gs)
    assert np.all(np.abs(got["disp"] - np.array(shift)) < 0.3)
    assert np.array_equal(got["ok"], ref["ok"])
    assert np.max(np.abs(got["disp"] - ref["disp"])) < 1e-3  # float32 maps on the device, float64 in the oracle
    assert [p["radius"] for p in got["per_level"]] == [p["radius"] for p in ref["per_level"]]


def test_split_nodes_take_their_initial_guess_from_their_own_side():
    """tests/test_subset_split.py:220-248: a node whose subset the cut reduced below 90 % is filled
    from its own side; without the split information its contaminated guess stays."""
    axis = np.arange(5.0)
    gz, gy, gx = np.meshgrid(axis, axis, axis, indexing="ij")
    coordinates = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
    grid = (5, 5, 5)
    n = coordinates.shape[0]
    ix = (np.arange(n) % grid[2]).reshape(grid)
    edge_ok = np.ones(grid + (3,), dtype=bool)
    edge_ok[:, :, 1, 0] = False  # the +x edge between ix = 1 and ix = 2 is cut everywhere
    mesh = NodeMesh(coordinates, grid, np.ones(n, dtype=bool), edge_ok.reshape(n, 3))
    disp = np.zeros((n, 3))
    disp[(ix <= 1).ravel(), 0] = 1.0
    disp[(ix >= 2).ravel(), 0] = -1.0
    disp[(ix == 1).ravel(), 0] = 0.0  # the contaminated peak: halfway between the two sides
    split_fraction = np.ones(n, dtype=np.float32)
    split_fraction[(ix == 1).ravel()] = 0.6
    para = SimpleNamespace(init_min_pce=0.0, init_outlier_threshold=0.0)
    ok = np.ones(n, dtype=bool)
    out, bad = IG.clean_initial_guess(disp.copy(), ok, None, mesh, para, split_fraction)
    cut = (ix == 1).ravel()
    assert bad[cut].all() and not bad[~cut].any()
    assert np.allclose(out[cut, 0], 1.0, atol=1e-6)
    assert np.allclose(out[(ix >= 2).ravel(), 0], -1.0)
    out2, bad2 = IG.clean_initial_guess(disp.copy(), ok, None, mesh, para, None)
    assert not bad2.any() and np.allclose(out2[cut, 0], 0.0)


def test_clean_initial_guess_median_test_matches_host():
    """Outliers planted in a smooth field are flagged and refilled as the host restatement does."""
    grid = (6, 7, 8)
    n = int(np.prod(grid))
    gz, gy, gx = np.meshgrid(*(np.arange(s, dtype=float) for s in grid), indexing="ij")
    coordinates = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
    disp = np.stack([0.1 * gx.ravel(), -0.05 * gy.ravel(), 0.02 * gz.ravel() + 1.0], 1)
    rng = np.random.default_rng(5)
    out_idx = rng.choice(n, 12, replace=False)
    disp[out_idx] += rng.normal(0, 4.0, (12, 3))
    ok = np.ones(n, dtype=bool)
    ok[rng.choice(n, 5, replace=False)] = False
    mesh = NodeMesh(coordinates, grid, np.ones(n, dtype=bool))
    para = SimpleNamespace(init_min_pce=0.0, init_outlier_threshold=2.0)
    out, bad = IG.clean_initial_guess(disp.copy(), ok, None, mesh, para)
    d = disp.copy()
    hb = ~ok
    d[hb] = np.nan
    flag = cleanup.universal_median_test(d.reshape(grid + (3,)), (~hb).reshape(grid), 2.0).ravel()
    hb |= flag
    d[hb] = np.nan
    ref = np.stack([cleanup.fill_nan_grid(d[:, c].reshape(grid)).ravel() for c in range(3)], 1)
    assert np.array_equal(bad, hb)
    assert np.max(np.abs(out - ref)) < 1e-8
